/*
 * oracle.c — TEST INFRASTRUCTURE ONLY. A plain-C restatement of the reference CPU renderer
 * (Rotsor/SpecRel), operation by operation, so the HIP backend can be checked bit for bit.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this file's
 * library; the product never does.
 *
 * Conventions (all as the Rust reference compiles them: IEEE f32, no contraction):
 *   - every float literal is an f32 literal; expressions keep Rust's left-to-right grouping;
 *   - Vec / scalar divides by multiplying with the reciprocal (minkowski.rs:124-129, 345-350);
 *   - f32 transcendentals are (float)f((double)x) with glibc's double routines;
 *   - `as i32` / `as usize` saturate, NaN -> 0; f32::max/min are maxNum/minNum;
 *     f32::clamp lets NaN through; rem_euclid uses exact fmod.
 *
 * Parity pins: every function here is checked in tests/ against the reference's own
 * known-answer tests (mesh4d.rs:1127-1160, star.rs:329-428, bloom.rs:131-151,
 * renderer.rs:1140-1168), brute force (trace and occlusion against a linear scan), closed
 * forms, and glibc / numpy float64 for the f32 helpers.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float x, y, z; } o3;
typedef struct { float t, x, y, z; } o4;

/* ---------------------------------------------------------------- scalar helpers */
static float ex(float x) { return (float)exp((double)x); }
static float sn(float x) { return (float)sin((double)x); }
static float cs(float x) { return (float)cos((double)x); }
static float at2(float y, float x) { return (float)atan2((double)y, (double)x); }
static float ac(float x) { return (float)acos((double)x); }
static float fmx(float a, float b) { return fmaxf(a, b); }
static float fmn(float a, float b) { return fminf(a, b); }
static float clampf_(float x, float lo, float hi) { /* core::f32::clamp */
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
static float remeu(float x, float m) { /* f32::rem_euclid */
    float r = fmodf(x, m);
    return r < 0.0f ? r + fabsf(m) : r;
}
static int32_t sat_i32(float x) {
    if (x != x) return 0;
    if (x >= 2147483647.0f) return 2147483647;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}
static uint64_t sat_usize(float x) {
    if (!(x > 0.0f)) return 0;
    if (x >= 18446744073709551616.0f) return UINT64_MAX;
    return (uint64_t)x;
}

float oracle_expf(float x) { return ex(x); }
float oracle_sinf(float x) { return sn(x); }
float oracle_cosf(float x) { return cs(x); }
float oracle_atan2f(float y, float x) { return at2(y, x); }
float oracle_acosf(float x) { return ac(x); }
float oracle_rem_euclid(float x, float m) { return remeu(x, m); }

#define TAU_F 6.28318530717958647692f
#define PI_F 3.14159265358979323846f
#define FRAC_1_PI_F 0.318309886183790671538f

/* ---------------------------------------------------------------- vectors (minkowski.rs:1-359) */
static o3 V3(float x, float y, float z) { o3 r = {x, y, z}; return r; }
static o4 V4(float t, float x, float y, float z) { o4 r = {t, x, y, z}; return r; }
static o3 add3(o3 a, o3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static o3 sub3(o3 a, o3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static o3 mul3(o3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
static float dot3(o3 a, o3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static float len3(o3 a) { return sqrtf(dot3(a, a)); }
static o3 norm3(o3 a) { /* Vec3::normalized */
    float l = len3(a);
    if (l > 1e-12f) return mul3(a, 1.0f / l);
    return V3(0.0f, 0.0f, 0.0f);
}
static o3 splat3(float v) { return V3(v, v, v); }
static o3 clamp3(o3 a, float lo, float hi) {
    return V3(clampf_(a.x, lo, hi), clampf_(a.y, lo, hi), clampf_(a.z, lo, hi));
}
static o4 add4(o4 a, o4 b) { return V4(a.t + b.t, a.x + b.x, a.y + b.y, a.z + b.z); }
static o4 sub4(o4 a, o4 b) { return V4(a.t - b.t, a.x - b.x, a.y - b.y, a.z - b.z); }
static o4 mul4(o4 a, float s) { return V4(a.t * s, a.x * s, a.y * s, a.z * s); }
static o4 neg4(o4 a) { return V4(-a.t, -a.x, -a.y, -a.z); }
static float mdot(o4 a, o4 b) { return -a.t * b.t + a.x * b.x + a.y * b.y + a.z * b.z; }
static float edot(o4 a, o4 b) { return a.t * b.t + a.x * b.x + a.y * b.y + a.z * b.z; }
static o3 spat(o4 a) { return V3(a.x, a.y, a.z); }
static o4 ld4(const float *p) { return V4(p[0], p[1], p[2], p[3]); }

/* rest_frame_cos (minkowski.rs:364-380) */
static float rest_frame_cos(o4 v, o4 n, o4 u) {
    o4 vp = add4(v, mul4(u, mdot(v, u)));
    float len_sq = mdot(vp, vp);
    if (len_sq > 1e-12f) return clampf_(mdot(vp, n) / sqrtf(len_sq), -1.0f, 1.0f);
    return 0.0f;
}

/* ---------------------------------------------------------------- geometry (mesh4d.rs) */
/* Aabb4D::intersect_ray (mesh4d.rs:113-146): returns 1 and *enter on a hit. */
static int aabb_ray(const float *mn, const float *mx, o4 o, o4 inv, float tmin, float tmax, float *enter) {
    float t0 = (mn[0] - o.t) * inv.t, t1 = (mx[0] - o.t) * inv.t;
    float x0 = (mn[1] - o.x) * inv.x, x1 = (mx[1] - o.x) * inv.x;
    float y0 = (mn[2] - o.y) * inv.y, y1 = (mx[2] - o.y) * inv.y;
    float z0 = (mn[3] - o.z) * inv.z, z1 = (mx[3] - o.z) * inv.z;
    float en = fmx(fmx(fmx(fmx(tmin, fmn(t0, t1)), fmn(x0, x1)), fmn(y0, y1)), fmn(z0, z1));
    float xt = fmn(fmn(fmn(fmn(tmax, fmx(t0, t1)), fmx(x0, x1)), fmx(y0, y1)), fmx(z0, z1));
    if (xt >= en) {
        *enter = en;
        return 1;
    }
    return 0;
}

/* SimplexGeom4D::intersect_fast (mesh4d.rs:175-211) */
static int intersect_fast(const specrel_geom *g, o4 o, o4 d, float tmin, float tmax, float *lam, o4 *ev,
                          float bary[4]) {
    o4 n = ld4(g->n), v0 = ld4(g->v0);
    float denom = edot(n, d);
    if (fabsf(denom) < 1e-9f) return 0;
    float lambda = edot(n, sub4(v0, o)) / denom;
    if (lambda <= tmin || lambda >= tmax) return 0;
    o4 event = add4(o, mul4(d, lambda));
    o4 q = sub4(event, v0);
    const float eps = -1e-5f;
    float u = edot(ld4(g->b1), q);
    if (u < eps || u > 1.0f - eps) return 0;
    float v = edot(ld4(g->b2), q);
    if (v < eps || u + v > 1.0f - eps) return 0;
    float w = edot(ld4(g->b3), q);
    if (w < eps || u + v + w > 1.0f - eps) return 0;
    *lam = lambda;
    *ev = event;
    bary[0] = 1.0f - u - v - w;
    bary[1] = u;
    bary[2] = v;
    bary[3] = w;
    return 1;
}

int oracle_intersect(const specrel_geom *g, const float o[4], const float d[4], float tmin, float tmax,
                     float out[9]) {
    float lam, b[4];
    o4 ev;
    if (!intersect_fast(g, ld4(o), ld4(d), tmin, tmax, &lam, &ev, b)) return 0;
    out[0] = lam;
    out[1] = ev.t; out[2] = ev.x; out[3] = ev.y; out[4] = ev.z;
    memcpy(out + 5, b, 16);
    return 1;
}

/* ---------------------------------------------------------------- BVH (bvh4d.rs:429-578) */
/* BvhNode4D::NOTHING (bvh4d.rs:30-31), "an internal node that no ray enters": no primitives and
   an empty box (Aabb4D::empty, mesh4d.rs:71-76). The slab test (mesh4d.rs:114-140) does not
   reject an empty box -- every axis gives (-inf, +inf) -- so the reference's traversal of a tree
   whose root is NOTHING (an arena holding nothing, scene_cache.rs:324) pushes node 0 as its own
   children and never ends, on the CPU (bvh4d.rs:445-516) and in the shader alike. The oracle
   follows the documented meaning instead: no ray enters it. */
static int is_nothing(const specrel_node *n) { return n->count == 0 && !(n->min[0] <= n->max[0]); }

static o4 safe_inv_dir(o4 d) { /* bvh4d.rs:430-441 */
    float c[4] = {d.t, d.x, d.y, d.z}, r[4];
    for (int i = 0; i < 4; ++i) {
        float v = c[i];
        if (fabsf(v) > 1e-12f) r[i] = 1.0f / v;
        else if (v >= 0.0f) r[i] = 1e12f;
        else r[i] = -1e12f;
    }
    return V4(r[0], r[1], r[2], r[3]);
}

typedef struct {
    int64_t slot;
    float lambda;
    o4 event;
    float bary[4];
} o_hit;

/* Distinct records a frame touches (oracle_render with counts): one bit per interior node whose
 * children's boxes were tested, per primitive record tested and per hit primitive whose attributes
 * were interpolated. Set for the worker threads of one render. */
typedef struct {
    _Atomic uint32_t *nodes, *prims, *hits;
} o_touch;
static _Thread_local const o_touch *tl_touch;
static void touch_bit(_Atomic uint32_t *bits, uint32_t i) { atomic_fetch_or(&bits[i >> 5], 1u << (i & 31u)); }

/* Bvh4D::trace (bvh4d.rs:445-520), up to the leaf slot of the winner. */
static int trace(const specrel_scene_arrays *s, o4 o, o4 d, float tmin, float tmax, o_hit *hit,
                 oracle_counts *c) {
    hit->slot = -1;
    if (s->num_nodes == 0) return 0;
    o4 inv = safe_inv_dir(d);
    const specrel_node *N = s->nodes;
    float root_enter;
    if (is_nothing(&N[0]) || !aabb_ray(N[0].min, N[0].max, o, inv, tmin, tmax, &root_enter)) return 0;
    uint32_t st_n[64];
    float st_t[64];
    size_t sp = 1;
    st_n[0] = 0;
    st_t[0] = root_enter;
    while (sp > 0) {
        sp -= 1;
        uint32_t ni = st_n[sp];
        float t_enter = st_t[sp];
        if (t_enter >= tmax) continue;
        const specrel_node *node = &N[ni];
        if (c) c->node_pops++;
        if (is_nothing(node)) continue; /* entered as a leaf without primitives */
        if (node->count > 0) {
            uint32_t first = node->right_or_first, end = first + node->count;
            for (uint32_t i = first; i < end; ++i) {
                float lam, b[4];
                o4 ev;
                if (c) c->prim_tests++;
                if (tl_touch) touch_bit(tl_touch->prims, i);
                if (intersect_fast(&s->geoms[i], o, d, tmin, tmax, &lam, &ev, b)) {
                    tmax = lam;
                    hit->slot = i;
                    hit->lambda = lam;
                    hit->event = ev;
                    memcpy(hit->bary, b, 16);
                }
            }
        } else {
            uint32_t li = node->left, ri = node->right_or_first;
            float tl, tr;
            if (c) c->box_tests += 2;
            if (tl_touch) touch_bit(tl_touch->nodes, ni);
            int hl = aabb_ray(N[li].min, N[li].max, o, inv, tmin, tmax, &tl);
            int hr = aabb_ray(N[ri].min, N[ri].max, o, inv, tmin, tmax, &tr);
            if (hl && hr) {
                if (sp + 2 <= 64) {
                    if (tl <= tr) {
                        st_n[sp] = ri; st_t[sp] = tr;
                        st_n[sp + 1] = li; st_t[sp + 1] = tl;
                    } else {
                        st_n[sp] = li; st_t[sp] = tl;
                        st_n[sp + 1] = ri; st_t[sp + 1] = tr;
                    }
                    sp += 2;
                }
            } else if (hl) {
                if (sp < 64) { st_n[sp] = li; st_t[sp] = tl; sp += 1; }
            } else if (hr) {
                if (sp < 64) { st_n[sp] = ri; st_t[sp] = tr; sp += 1; }
            }
        }
    }
    return hit->slot >= 0;
}

/* Bvh4D::is_occluded (bvh4d.rs:523-578) */
static int occluded(const specrel_scene_arrays *s, o4 o, o4 d, float tmin, float tmax, oracle_counts *c) {
    if (s->num_nodes == 0) return 0;
    o4 inv = safe_inv_dir(d);
    const specrel_node *N = s->nodes;
    float e;
    if (is_nothing(&N[0]) || !aabb_ray(N[0].min, N[0].max, o, inv, tmin, tmax, &e)) return 0;
    uint32_t st[64];
    size_t sp = 1;
    st[0] = 0;
    while (sp > 0) {
        sp -= 1;
        const uint32_t ni = st[sp];
        const specrel_node *node = &N[ni];
        if (c) c->shadow_node_pops++;
        if (is_nothing(node)) continue;
        if (node->count > 0) {
            uint32_t first = node->right_or_first, end = first + node->count;
            for (uint32_t i = first; i < end; ++i) {
                float lam, b[4];
                o4 ev;
                if (c) c->shadow_prim_tests++;
                if (tl_touch) touch_bit(tl_touch->prims, i);
                if (intersect_fast(&s->geoms[i], o, d, tmin, tmax, &lam, &ev, b)) return 1;
            }
        } else {
            uint32_t li = node->left, ri = node->right_or_first;
            float t;
            if (c) c->shadow_box_tests += 2;
            if (tl_touch) touch_bit(tl_touch->nodes, ni);
            int hl = aabb_ray(N[li].min, N[li].max, o, inv, tmin, tmax, &t);
            int hr = aabb_ray(N[ri].min, N[ri].max, o, inv, tmin, tmax, &t);
            if (hr && sp < 64) st[sp++] = ri;
            if (hl && sp < 64) st[sp++] = li;
        }
    }
    return 0;
}

int64_t oracle_trace(const specrel_scene_arrays *s, const float o[4], const float d[4], float tmin, float tmax,
                     float out[9], oracle_counts *c) {
    o_hit h;
    if (!trace(s, ld4(o), ld4(d), tmin, tmax, &h, c)) return -1;
    out[0] = h.lambda;
    out[1] = h.event.t; out[2] = h.event.x; out[3] = h.event.y; out[4] = h.event.z;
    memcpy(out + 5, h.bary, 16);
    return h.slot;
}

int64_t oracle_trace_brute(const specrel_scene_arrays *s, const float o[4], const float d[4], float tmin,
                           float tmax, float out[9]) {
    int64_t best = -1;
    float best_l = tmax;
    for (uint32_t i = 0; i < s->num_prims; ++i) {
        float r[9];
        if (oracle_intersect(&s->geoms[i], o, d, tmin, best_l, r)) {
            best = i;
            best_l = r[0];
            memcpy(out, r, sizeof r);
        }
    }
    return best;
}

int oracle_occluded(const specrel_scene_arrays *s, const float o[4], const float d[4], float tmin, float tmax) {
    return occluded(s, ld4(o), ld4(d), tmin, tmax, NULL);
}

int oracle_occluded_brute(const specrel_scene_arrays *s, const float o[4], const float d[4], float tmin,
                          float tmax) {
    for (uint32_t i = 0; i < s->num_prims; ++i) {
        float r[9];
        if (oracle_intersect(&s->geoms[i], o, d, tmin, tmax, r)) return 1;
    }
    return 0;
}

/* ---------------------------------------------------------------- hit record (mesh4d.rs:382-425) */
typedef struct {
    float lambda;
    o4 event;
    float tet_bary[4];
    o3 tri_bary;
    float uv_u, uv_v, proper_time, grad_u, grad_v, uvws_u, uvws_v;
    o4 normal, four_vel;
    float decided_at;
    uint32_t material;
} o_rec;

static float lerp4(const float *a, const float *b) { /* ((a0*w0 + a1*u) + a2*v) + a3*w */
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}

static void interpolate_hit(const specrel_attr *A, const o_hit *h, o_rec *r) {
    const float *b = h->bary;
    r->lambda = h->lambda;
    r->event = h->event;
    memcpy(r->tet_bary, b, 16);
    r->tri_bary = V3(lerp4(A->tb_x, b), lerp4(A->tb_y, b), lerp4(A->tb_z, b));
    r->uv_u = lerp4(A->uv_u, b);
    r->uv_v = lerp4(A->uv_v, b);
    r->proper_time = lerp4(A->tau, b);
    r->grad_u = lerp4(A->grad_u, b);
    r->grad_v = lerp4(A->grad_v, b);
    r->decided_at = lerp4(A->decided_at, b);
    r->uvws_u = A->misc[0];
    r->uvws_v = A->misc[1];
    r->normal = ld4(A->normal);
    r->four_vel = ld4(A->four_vel);
    memcpy(&r->material, &A->misc[2], 4);
}

/* ---------------------------------------------------------------- shading (renderer.rs) */
typedef struct {
    o3 base, checker;
    float emission, grid, wire, cavity;
    uint32_t pattern;
} o_mat;

static o_mat load_mat(const specrel_scene_arrays *s, uint32_t id) {
    o_mat m;
    const specrel_material *g = &s->materials[id];
    m.base = V3(g->base[0], g->base[1], g->base[2]);
    m.emission = g->base[3];
    m.checker = V3(g->checker[0], g->checker[1], g->checker[2]);
    m.grid = g->checker[3];
    m.wire = g->params[0];
    memcpy(&m.pattern, &g->params[1], 4);
    m.cavity = g->params[2];
    return m;
}

typedef struct {
    const specrel_scene_arrays *s;
    const specrel_frame *f;
    const uint8_t *lut;
    uint8_t glyphs[14][7];
    o4 origin, e0;
    float cone;
    oracle_counts *c;
} o_ctx;

/* hsv_to_rgb (renderer.rs:1081-1100) */
static o3 hsv_to_rgb(float h_deg, float s, float v) {
    float c = v * s;
    float h = remeu(h_deg / 60.0f, 6.0f);
    float x = c * (1.0f - fabsf(remeu(h, 2.0f) - 1.0f));
    float m = v - c;
    float r, g, b;
    if (h < 1.0f) { r = c; g = x; b = 0.0f; }
    else if (h < 2.0f) { r = x; g = c; b = 0.0f; }
    else if (h < 3.0f) { r = 0.0f; g = c; b = x; }
    else if (h < 4.0f) { r = 0.0f; g = x; b = c; }
    else if (h < 5.0f) { r = x; g = 0.0f; b = c; }
    else { r = c; g = 0.0f; b = x; }
    return V3(r + m, g + m, b + m);
}

/* shade_sky_gradient / shade_sky (renderer.rs:1056-1079) */
static o3 sky_gradient(o4 dir) {
    o3 s = norm3(spat(dir));
    float horizon = clampf_(s.y * 0.5f + 0.5f, 0.0f, 1.0f);
    o3 zen = V3(0.04f, 0.06f, 0.13f), hor = V3(0.12f, 0.16f, 0.25f);
    return add3(mul3(hor, 1.0f - horizon), mul3(zen, horizon));
}
static o3 shade_sky(o4 dir) {
    o3 s = norm3(spat(dir));
    o3 sky = sky_gradient(dir);
    float theta = ac(clampf_(s.y, -1.0f, 1.0f));
    float phi = at2(s.z, s.x);
    float gu = remeu(phi / PI_F * 12.0f, 1.0f);
    float gv = remeu(theta / PI_F * 6.0f, 1.0f);
    if (gu < 0.025f || gv < 0.025f) sky = add3(sky, V3(0.06f, 0.09f, 0.14f));
    return sky;
}
void oracle_shade_sky(const float dir[4], float out[3]) {
    o3 c = shade_sky(ld4(dir));
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}

static o3 distance_fog(o3 c, o4 dir, float lambda) { /* renderer.rs:691-695 */
    float fog = ex(-0.008f * lambda);
    o3 sky = sky_gradient(dir);
    return add3(mul3(c, fog), mul3(sky, 1.0f - fog));
}

/* analytic_decision_age (renderer.rs:319-325) */
static float decision_age(o4 event, o4 origin, o4 u) {
    o4 d = sub4(event, origin);
    float time_ago = mdot(d, u);
    float dist = sqrtf(fmx(time_ago * time_ago + mdot(d, d), 0.0f));
    return time_ago + dist;
}
float oracle_decision_age(const float e[4], const float o[4], const float u[4]) {
    return decision_age(ld4(e), ld4(o), ld4(u));
}

static o3 decision_age_color(float age) { /* renderer.rs:494-507 */
    if (age < 0.0f) return V3(0.14f, 0.14f, 0.18f);
    o3 col = hsv_to_rgb(fmn(age * 4.0f, 240.0f), 0.82f, 0.88f);
    if (age < 0.45f) {
        float f = 1.0f - age / 0.45f;
        col = add3(mul3(col, 1.0f - f * 0.7f), mul3(V3(1.0f, 0.96f, 0.65f), f * 0.9f));
    }
    if (fabsf(remeu(age * 0.5f, 1.0f) - 0.5f) > 0.465f) col = V3(0.95f, 0.98f, 1.0f);
    return col;
}

/* apply_undecided_haze (renderer.rs:514-541) */
static o3 haze(o3 base, o4 event, const float (*origins)[4], uint32_t n) {
    const float SHELL = 0.18f;
    float inside = -INFINITY;
    for (uint32_t i = 0; i < n; ++i) {
        o4 d = sub4(event, ld4(origins[i]));
        inside = fmx(inside, d.t - len3(spat(d)));
    }
    if (inside <= -SHELL) return base;
    o3 col = base;
    if (fabsf(inside) < SHELL) {
        float k = 1.0f - fabsf(inside) / SHELL;
        col = add3(col, mul3(V3(0.45f, 0.92f, 1.0f), 0.85f * k * k));
    }
    if (inside > 0.0f) {
        float ripple = 0.5f + 0.5f * sn(event.x * 3.5f + event.z * 3.5f - event.t * 4.0f) * cs(inside * 5.0f);
        o3 hz = add3(mul3(V3(0.32f, 0.18f, 0.62f), 1.0f - 0.35f * ripple),
                     mul3(V3(0.18f, 0.72f, 0.92f), 0.35f * ripple));
        float alpha = 0.55f + 0.25f * fmn(inside * 0.5f, 1.0f);
        col = add3(mul3(col, 1.0f - alpha), mul3(hz, alpha));
    }
    return col;
}
void oracle_haze(const float base[3], const float event[4], const float *origins, uint32_t n, float out[3]) {
    o3 c = haze(V3(base[0], base[1], base[2]), ld4(event), (const float (*)[4])origins, n);
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}

/* clock_tile_central_proper_time (renderer.rs:704-715) */
static float central_tau(float u_tile, float v_tile, float grid_scale, float local_tau, float gu, float gv) {
    float inv_grid = 1.0f / fmx(grid_scale, 1e-4f);
    float du = (u_tile - 0.5f) * inv_grid;
    float dv = (v_tile - 0.5f) * inv_grid;
    return local_tau - (du * gu + dv * gv);
}

static int on_hand(float dx, float dy, float ang, float length, float half_w, float back) {
    float hx = sn(ang), hy = cs(ang);
    float along = dx * hx + dy * hy;
    float perp = fabsf(dx * hy - dy * hx);
    return along >= back && along <= length && perp <= half_w;
}

/* sample_corner_mini_clock (renderer.rs:880-927) */
static o3 mini_clock(float mdx, float mdy, float mr, float mini_tau, const o_mat *m) {
    if (mr > 0.080f) return mul3(m->base, 1.15f);
    float angle = remeu(at2(mdx, mdy), TAU_F);
    float fast = remeu((mini_tau / 1.0f) * TAU_F, TAU_F);
    float main_a = remeu((mini_tau / 10.0f) * TAU_F, TAU_F);
    o3 col = V3(0.04f, 0.06f, 0.10f);
    if (angle <= fast && mr < 0.066f) col = add3(col, mul3(m->base, 0.25f));
    if (0.062f <= mr && mr <= 0.078f) {
        float sector = angle / (TAU_F / 10.0f);
        float dist_to_tick = fabsf(sector - roundf(sector));
        int32_t tick = sat_i32(roundf(sector)) % 10;
        int card = tick == 0 || tick == 5;
        float half_w = card ? 0.13f : 0.07f;
        if (dist_to_tick < half_w) return tick == 0 ? V3(1.0f, 0.88f, 0.15f) : V3(0.80f, 0.85f, 0.92f);
    }
    if (on_hand(mdx, mdy, fast, 0.068f, 0.0065f, -0.008f)) return V3(0.25f, 0.98f, 1.0f);
    if (on_hand(mdx, mdy, main_a, 0.052f, 0.0085f, -0.008f)) return V3(1.0f, 0.95f, 0.30f);
    if (mr < 0.012f) return V3(1.0f, 1.0f, 1.0f);
    return col;
}

/* sample_digital_tau (renderer.rs:946-981) */
static int digital_tau(const uint8_t glyphs[14][7], float u, float v, float tau) {
    float rel_u = (u - 0.28f) / (0.72f - 0.28f);
    float rel_v = (0.18f - v) / (0.18f - 0.06f);
    uint64_t col = sat_usize(floorf(rel_u * 35.0f));
    uint64_t row = sat_usize(floorf(rel_v * 7.0f));
    if (col >= 35 || row >= 7) return 0;
    uint64_t ci = col / 6, cc = col % 6;
    if (ci >= 6 || cc >= 5) return 0;
    int32_t t_tenths = sat_i32(roundf(tau * 10.0f));
    uint32_t abs_t = t_tenths < 0 ? (uint32_t)0 - (uint32_t)t_tenths : (uint32_t)t_tenths;
    uint32_t tenths = abs_t % 10, ones = (abs_t / 10) % 10, tens = (abs_t / 100) % 10;
    int g; /* glyph slot: 0-9 digits, 10 '+', 11 '-', 12 '.', 13 'S' */
    switch (ci) {
        case 0: g = t_tenths < 0 ? 11 : 10; break;
        case 1: g = (int)tens; break;
        case 2: g = (int)ones; break;
        case 3: g = 12; break;
        case 4: g = (int)tenths; break;
        default: g = 13; break;
    }
    return ((glyphs[g][row] >> (4 - cc)) & 1) != 0;
}

static o3 clock_face(const o_ctx *x, const o_rec *h, const o_mat *m) { /* renderer.rs:737-874 */
    float grid_scale = fmx(m->grid, 1e-4f);
    float inv_grid = 1.0f / grid_scale;
    float u = remeu(h->uv_u * grid_scale, 1.0f);
    float v = remeu(h->uv_v * grid_scale, 1.0f);
    float ctau = central_tau(u, v, grid_scale, h->proper_time, h->grad_u, h->grad_v);
    static const float corners[4][2] = {{0.135f, 0.865f}, {0.865f, 0.865f}, {0.135f, 0.135f}, {0.865f, 0.135f}};
    for (int k = 0; k < 4; ++k) {
        float mcx = corners[k][0], mcy = corners[k][1];
        float mdx = u - mcx, mdy = v - mcy;
        float mr = sqrtf(mdx * mdx + mdy * mdy);
        if (mr <= 0.095f) {
            float osc = ctau + ((mcx - 0.5f) * inv_grid) * h->grad_u + ((mcy - 0.5f) * inv_grid) * h->grad_v;
            float px = (mdx * inv_grid) * h->uvws_u;
            float py = (mdy * inv_grid) * h->uvws_v;
            float mr_phys = sqrtf(px * px + py * py);
            return mini_clock(mdx, mdy, mr, osc - mr_phys, m);
        }
    }
    if ((0.035f <= v && v <= 0.205f) && (0.25f <= u && u <= 0.75f)) {
        if (!(0.050f <= v && v <= 0.190f) || !(0.265f <= u && u <= 0.735f)) return mul3(m->base, 0.9f);
        if ((0.06f <= v && v <= 0.18f) && (0.28f <= u && u <= 0.72f) && digital_tau(x->glyphs, u, v, ctau))
            return V3(1.0f, 0.96f, 0.30f);
        return V3(0.03f, 0.04f, 0.07f);
    }
    float dx = u - 0.50f, dy = v - 0.57f;
    float r = sqrtf(dx * dx + dy * dy);
    if (r > 0.330f) {
        int32_t gu = sat_i32(floorf(u * 4.0f)), gv = sat_i32(floorf(v * 4.0f));
        int32_t sum = (int32_t)((uint32_t)gu + (uint32_t)gv);
        int32_t rem = ((sum % 2) + 2) % 2;
        return rem == 0 ? mul3(m->base, 0.78f) : mul3(m->checker, 0.78f);
    }
    if (r > 0.292f) return mul3(m->base, 1.15f);
    float dxp = (dx * inv_grid) * h->uvws_u, dyp = (dy * inv_grid) * h->uvws_v;
    float r_phys = sqrtf(dxp * dxp + dyp * dyp);
    float dial_tau = ctau - r_phys;
    float angle = remeu(at2(dx, dy), TAU_F);
    float main_a = remeu((dial_tau / 10.0f) * TAU_F, TAU_F);
    o3 dial = V3(0.04f, 0.06f, 0.10f);
    if (angle <= main_a && r < 0.236f) dial = add3(dial, mul3(m->base, 0.22f));
    if (0.220f <= r && r <= 0.286f) {
        float sector = angle / (TAU_F / 10.0f);
        float dist_to_tick = fabsf(sector - roundf(sector));
        int32_t tick = sat_i32(roundf(sector)) % 10;
        int card = tick == 0 || tick == 5;
        float half_w = card ? 0.105f : 0.055f;
        float inner_r = card ? 0.218f : 0.242f;
        if (dist_to_tick < half_w && r >= inner_r) {
            if (tick == 0) return V3(1.0f, 0.88f, 0.15f);
            if (card) return V3(1.0f, 1.0f, 1.0f);
            return V3(0.75f, 0.80f, 0.88f);
        }
    }
    float fast = remeu((dial_tau / 1.0f) * TAU_F, TAU_F);
    if (on_hand(dx, dy, fast, 0.254f, 0.012f, -0.025f)) return V3(0.25f, 0.98f, 1.0f);
    if (on_hand(dx, dy, main_a, 0.228f, 0.022f, -0.025f)) return V3(1.0f, 0.95f, 0.30f);
    if (on_hand(dx, dy, main_a, 0.236f, 0.030f, -0.025f)) return V3(0.02f, 0.02f, 0.04f);
    float slow = remeu((dial_tau / 100.0f) * TAU_F, TAU_F);
    if (on_hand(dx, dy, slow, 0.162f, 0.026f, -0.025f)) return V3(1.0f, 0.35f, 0.25f);
    if (r < 0.034f) return V3(1.0f, 1.0f, 1.0f);
    return dial;
}

static o3 light_clock(float uvu, float uvv, float tau, const o_mat *m) { /* renderer.rs:988-1052 */
    float u = remeu(uvu * m->grid, 1.0f);
    float v = remeu(uvv * m->grid, 1.0f);
    if (!(0.12f <= u && u <= 0.88f)) return mul3(m->base, 0.65f);
    if ((0.04f <= v && v <= 0.10f) || (0.90f <= v && v <= 0.96f)) return V3(0.75f, 0.90f, 0.98f);
    if (!(0.10f <= v && v <= 0.90f)) return V3(0.08f, 0.12f, 0.16f);
    float period = fmx(2.0f * m->cavity, 1e-4f);
    float cycle = remeu(tau / period, 1.0f);
    float bounce = cycle < 0.5f ? cycle * 2.0f : (1.0f - cycle) * 2.0f;
    float photon_v = 0.10f + 0.80f * bounce;
    float dv = fabsf(v - photon_v), du = fabsf(u - 0.5f);
    float dist = sqrtf(du * du * 0.7f + dv * dv);
    o3 col = V3(0.05f, 0.09f, 0.13f);
    float ruler = remeu((v - 0.10f) / 0.80f * 8.0f, 1.0f);
    if ((u < 0.22f || u > 0.78f) && ruler < 0.08f) col = add3(col, V3(0.20f, 0.35f, 0.35f));
    if (du < 0.015f) col = add3(col, V3(0.06f, 0.16f, 0.16f));
    if (dv < 0.032f) {
        float a = 1.0f - dv / 0.032f;
        col = add3(col, mul3(m->base, 0.9f * (a * a)));
    }
    if (dist < 0.11f) {
        float a = 1.0f - dist / 0.11f;
        col = add3(col, mul3(V3(1.0f, 1.0f, 0.85f), 1.4f * (a * a)));
    }
    return col;
}

static o3 surface_albedo(const o_ctx *x, const o_rec *h, const o_mat *m) { /* renderer.rs:544-580 */
    o3 albedo;
    if (m->pattern == SPECREL_PATTERN_CLOCK) {
        albedo = clock_face(x, h, m);
    } else if (m->pattern == SPECREL_PATTERN_LIGHT_CLOCK) {
        albedo = light_clock(h->uv_u, h->uv_v, h->proper_time, m);
    } else {
        int32_t gu = sat_i32(floorf(h->uv_u * m->grid)), gv = sat_i32(floorf(h->uv_v * m->grid));
        int32_t sum = (int32_t)((uint32_t)gu + (uint32_t)gv);
        int even = (((sum % 2) + 2) % 2) == 0;
        o3 col = even ? m->base : m->checker;
        float mtb = fmn(fmn(h->tri_bary.x, h->tri_bary.y), h->tri_bary.z);
        if (mtb < m->wire) {
            float et = clampf_(mtb / m->wire, 0.0f, 1.0f);
            col = add3(mul3(splat3(0.07f), 1.0f - et), mul3(col, et));
        }
        albedo = col;
    }
    if (m->pattern != SPECREL_PATTERN_CHECKER) {
        float ut = remeu(h->uv_u * m->grid, 1.0f), vt = remeu(h->uv_v * m->grid, 1.0f);
        float bd = fmn(fmn(ut, 1.0f - ut), fmn(vt, 1.0f - vt));
        if (bd < 0.025f) albedo = mul3(albedo, 0.35f);
    }
    return albedo;
}

/* compute_illumination + light_emission_event (renderer.rs:585-653); `parts` (may be NULL)
 * receives the intermediate values for the tests. */
static float illumination(const o_ctx *x, o4 dir, const o_rec *h, const o_mat *m, oracle_shade_parts *parts) {
    const specrel_frame *f = x->f;
    o3 light = V3(x->s->light_pos[1], x->s->light_pos[2], x->s->light_pos[3]);
    o3 s2l = sub3(light, spat(h->event));
    float dist = fmx(len3(s2l), 1e-4f);
    float t_emit;
    if (f->shadow_mode == SPECREL_SHADOW_CAMERA_SLICE) {
        o4 e0 = ld4(f->e[0]);
        o3 beta = fabsf(e0.t) > 1e-12f ? mul3(spat(e0), 1.0f / e0.t) : V3(0.0f, 0.0f, 0.0f);
        t_emit = h->event.t + dot3(beta, s2l);
    } else {
        t_emit = h->event.t - dist;
    }
    o4 emit = V4(t_emit, light.x, light.y, light.z);
    o4 to_light = mul4(sub4(emit, h->event), 1.0f / dist);
    float cl = fmx(rest_frame_cos(to_light, h->normal, h->four_vel), 0.0f);
    float cv = fmx(-rest_frame_cos(dir, h->normal, h->four_vel), 0.0f);
    if (parts) {
        parts->t_emit = t_emit;
        parts->dist = dist;
        parts->cos_light = cl;
        parts->cos_view = cv;
        parts->light_shadowed = 0;
    }
    if (m->pattern == SPECREL_PATTERN_CHECKER) {
        int in_shadow = 0;
        if (cl > 0.0f && f->shadow_mode != SPECREL_SHADOW_OFF) {
            if (x->c) x->c->shadow_rays++;
            in_shadow = occluded(x->s, h->event, to_light, 2e-3f, dist - 2e-3f, x->c);
        }
        if (parts) parts->light_shadowed = in_shadow;
        float sf = in_shadow ? 0.32f : 1.0f;
        float atten = 1.0f / (1.0f + 0.006f * dist * dist);
        return m->emission + 0.30f * cv + 0.80f * sf * cl * atten;
    }
    return 0.72f + 0.38f * cv + 0.22f * cl;
}

/* compute_flash_illumination (renderer.rs:658-688) */
static o3 flash_illumination(const o_ctx *x, const o_rec *h) {
    const specrel_frame *f = x->f;
    o3 total = V3(0.0f, 0.0f, 0.0f);
    for (uint32_t i = 0; i < f->num_flashes; ++i) {
        o4 at = ld4(f->flash_at[i]);
        o3 color = V3(f->flash_color_width[i][0], f->flash_color_width[i][1], f->flash_color_width[i][2]);
        float width = f->flash_color_width[i][3], intensity = f->flash_intensity[i];
        float dt = h->event.t - at.t;
        if (dt <= 0.0f) continue;
        float dist = fmx(len3(sub3(spat(at), spat(h->event))), 1e-4f);
        float shell_err = fabsf(dt - dist);
        float half_w = fmx(0.5f * width, 1e-3f);
        if (shell_err >= half_w) continue;
        float k = 1.0f - shell_err / half_w;
        float pulse = k * k;
        o4 to_flash = mul4(sub4(at, h->event), 1.0f / dist);
        float cf = fmx(rest_frame_cos(to_flash, h->normal, h->four_vel), 0.0f);
        int in_shadow = 0;
        if (cf > 0.0f && f->shadow_mode != SPECREL_SHADOW_OFF) {
            if (x->c) x->c->shadow_rays++;
            in_shadow = occluded(x->s, h->event, to_flash, 2e-3f, dist - 2e-3f, x->c);
        }
        float sf = in_shadow ? 0.25f : 1.0f;
        float atten = 1.0f / (1.0f + 0.015f * dist * dist);
        total = add3(total, mul3(color, intensity * pulse * atten * (0.25f + 0.75f * sf * cf)));
    }
    return total;
}

/* shade_hit (renderer.rs:414-489); `parts` (may be NULL) receives the intermediate values. */
static o3 shade_hit(const o_ctx *x, o4 dir, const o_rec *h, oracle_shade_parts *parts) {
    const specrel_frame *f = x->f;
    o_mat m = load_mat(x->s, h->material);
    switch (f->mode) {
        case SPECREL_MODE_SHADED: {
            o3 albedo = surface_albedo(x, h, &m);
            float lighting = illumination(x, dir, h, &m, parts);
            o3 lit = mul3(albedo, lighting);
            o3 fl = V3(0.0f, 0.0f, 0.0f);
            if (f->num_flashes > 0) {
                fl = flash_illumination(x, h);
                o3 a = add3(mul3(albedo, 0.65f), splat3(0.35f));
                lit = add3(lit, V3(a.x * fl.x, a.y * fl.y, a.z * fl.z));
            }
            if (parts) {
                parts->albedo[0] = albedo.x, parts->albedo[1] = albedo.y, parts->albedo[2] = albedo.z;
                parts->lighting = lighting;
                parts->flash[0] = fl.x, parts->flash[1] = fl.y, parts->flash[2] = fl.z;
                parts->lit[0] = lit.x, parts->lit[1] = lit.y, parts->lit[2] = lit.z;
            }
            return distance_fog(lit, dir, h->lambda);
        }
        case SPECREL_MODE_ISOCHRONES: {
            float phase = remeu(h->lambda * 0.12f, 1.0f);
            o3 base = hsv_to_rgb(phase * 300.0f, 0.8f, 0.85f);
            float tc = fabsf(remeu(h->event.t, 1.0f) - 0.5f);
            float uc = fabsf(remeu(h->proper_time, 1.0f) - 0.5f);
            if (tc > 0.45f) return V3(1.0f, 1.0f, 1.0f);
            if (uc > 0.46f) return V3(0.1f, 0.1f, 0.1f);
            return base;
        }
        case SPECREL_MODE_INSPECTOR: {
            const float *b = h->tet_bary;
            float min_b = fmn(fmn(fmn(b[0], b[1]), b[2]), b[3]);
            float edge = min_b < 0.035f ? 0.12f : 1.0f;
            float sb = clampf_((h->four_vel.t - 1.0f) * 0.8f, 0.0f, 1.0f);
            o3 tint = V3(0.25f + 0.65f * b[1] + 0.3f * sb, 0.25f + 0.65f * b[2],
                         0.35f + 0.55f * b[3] * (1.0f - 0.5f * sb));
            return mul3(tint, edge);
        }
        default: { /* DecisionAge */
            float age;
            if (h->decided_at > -1.0e9f * 0.5f) age = f->frontier_tau - h->decided_at;
            else age = decision_age(h->event, ld4(f->frontier_origin), ld4(f->e[0]));
            float cv = fmx(-rest_frame_cos(dir, h->normal, h->four_vel), 0.0f);
            return distance_fog(mul3(decision_age_color(age), 0.55f + 0.45f * cv), dir, h->lambda);
        }
    }
}

/* ---------------------------------------------------------------- stars (star.rs:114-286) */
static float planck(float lambda_um, float temperature) { /* star.rs:162-166 */
    float x = fmn(14388.0f / (lambda_um * fmx(temperature, 1.0f)), 80.0f);
    float l5 = (float)pow((double)lambda_um, 5.0); /* powi(5) of a constant, folded */
    return 1.0f / (l5 * (ex(x) - 1.0f));
}
static o3 blackbody(float T) { /* star.rs:171-180 */
    return V3(planck(0.610f, T) / planck(0.610f, 6504.0f), planck(0.550f, T) / planck(0.550f, 6504.0f),
              planck(0.465f, T) / planck(0.465f, 6504.0f));
}
void oracle_blackbody(float T, float out[3]) {
    o3 c = blackbody(T);
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}
float oracle_color_norm(float T, float I) { /* star.rs:117-124 with luminance (156-158) */
    o3 c = blackbody(T);
    float lum = 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z;
    return lum > 0.0f ? I / lum : 0.0f;
}

typedef struct {
    o4 event, u;
    float s_min, s_max;
} o_seg;
static o_seg load_seg(const float *w, uint32_t i) {
    const float *p = w + 4 * (size_t)i;
    o_seg s = {ld4(p), ld4(p + 4), p[8], p[9]};
    return s;
}

static float segment_cone_crossing(const o_seg *seg, o4 origin, o4 e0, float cone) { /* star.rs:185-210 */
    o4 d0 = sub4(seg->event, origin);
    o4 u = seg->u;
    if (cone == 0.0f) {
        float tau0 = -mdot(d0, e0);
        float gamma = -mdot(u, e0);
        return -tau0 / gamma;
    }
    float a = mdot(d0, d0);
    float b = mdot(d0, u);
    float sq = sqrtf(fmx(b * b + a, 0.0f));
    if (cone < 0.0f) return b > 0.0f ? -a / (b + sq) : b - sq;
    return b < 0.0f ? -a / (b - sq) : b + sq;
}
static float past_the_cone(o4 event, o4 origin, o4 e0, float cone) { /* star.rs:215-223 */
    o4 d = sub4(event, origin);
    float tau = -mdot(d, e0);
    if (cone == 0.0f) return tau;
    float rho = sqrtf(fmx(mdot(d, d) + tau * tau, 0.0f));
    return tau - cone * rho;
}
typedef struct {
    const float *words;
    uint32_t num_stars;
} o_stars;

/* worldline_cone_event (star.rs:228-235); partition_point as Rust's binary search
 * (slice::binary_search_by). */
static int cone_event2(const o_stars *S, uint32_t star, o4 origin, o4 e0, float cone, o4 *d_out, o4 *u_out) {
    const float *hdr = S->words + 4 * (size_t)star;
    uint32_t first, count;
    memcpy(&first, &hdr[2], 4);
    memcpy(&count, &hdr[3], 4);
    const float *segw = S->words + 4 * (size_t)S->num_stars;
    /* partition_point(|seg| past_the_cone(seg.event) <= 0) */
    size_t before;
    if (count == 0) {
        before = 0;
    } else {
        size_t size = count, base = 0;
        while (size > 1) {
            size_t half = size / 2, mid = base + half;
            o_seg sg = load_seg(segw, 3 * (first + (uint32_t)mid));
            int less = past_the_cone(sg.event, origin, e0, cone) <= 0.0f;
            base = less ? mid : base;
            size -= half;
        }
        o_seg sg = load_seg(segw, 3 * (first + (uint32_t)base));
        int less = past_the_cone(sg.event, origin, e0, cone) <= 0.0f;
        before = base + (less ? 1 : 0);
    }
    size_t idx = (before > 1 ? before : 1) - 1;
    if (idx >= count) return 0;
    o_seg sg = load_seg(segw, 3 * (first + (uint32_t)idx));
    float s = segment_cone_crossing(&sg, origin, e0, cone);
    if (!(s >= sg.s_min && s <= sg.s_max)) return 0;
    *d_out = add4(sub4(sg.event, origin), mul4(sg.u, s));
    *u_out = sg.u;
    return 1;
}

int oracle_cone_event(const float *w, uint32_t num_stars, uint32_t star, const float origin[4], const float e0[4],
                      float cone, float d_out[4], float u_out[4]) {
    o_stars S = {w, num_stars};
    o4 d, u;
    if (!cone_event2(&S, star, ld4(origin), ld4(e0), cone, &d, &u)) return 0;
    d_out[0] = d.t; d_out[1] = d.x; d_out[2] = d.y; d_out[3] = d.z;
    u_out[0] = u.t; u_out[1] = u.x; u_out[2] = u.y; u_out[3] = u.z;
    return 1;
}

/* star_light (star.rs:240-286) */
static o3 star_light(const o_stars *S, o4 origin, o4 dir, o4 e0, float cone, float width, float spread,
                     float lambda_max, oracle_counts *c) {
    o4 n = sub4(dir, mul4(e0, cone));
    o3 total = V3(0.0f, 0.0f, 0.0f);
    for (uint32_t i = 0; i < S->num_stars; ++i) {
        o4 d, u;
        if (c) c->star_pairs++;
        if (!cone_event2(S, i, origin, e0, cone, &d, &u)) continue;
        float tau = -mdot(d, e0);
        o4 r = add4(d, mul4(e0, -tau));
        float along = mdot(r, n);
        if (along <= 0.0f || along >= lambda_max) continue;
        o4 perp = sub4(r, mul4(n, along));
        float d2 = fmx(mdot(perp, perp), 0.0f);
        float sigma = 0.6f * (width + spread * along);
        float cut = 4.0f * sigma;
        if (d2 >= cut * cut) continue;
        float inv_2s2 = 1.0f / (2.0f * sigma * sigma);
        float kernel = ex(-d2 * inv_2s2) * inv_2s2 * FRAC_1_PI_F;
        float doppler = 1.0f;
        if (cone < 0.0f) {
            float lam = sqrtf(fmx(mdot(r, r), 0.0f));
            float du = mdot(d, u);
            doppler = du > 0.0f ? lam / du : 1.0f;
        }
        const float *hdr = S->words + 4 * (size_t)i;
        float T = hdr[0], cn = hdr[1];
        total = add3(total, mul3(blackbody(T * doppler), cn * kernel));
    }
    return total;
}

void oracle_star_light(const float *words, uint32_t num_stars, const float origin[4], const float dir[4],
                       const float e0[4], float cone, float width, float spread, float lambda_max, float out[3]) {
    o_stars S = {words, num_stars};
    o3 c = star_light(&S, ld4(origin), ld4(dir), ld4(e0), cone, width, spread, lambda_max, NULL);
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}

/* ---------------------------------------------------------------- bloom (bloom.rs:22-125) */
static void level_size(uint32_t w, uint32_t h, int k, uint32_t *ow, uint32_t *oh) {
    for (int i = 0; i < k; ++i) {
        w = (w + 1) / 2;
        h = (h + 1) / 2;
    }
    *ow = w;
    *oh = h;
}
static size_t level_offset(uint32_t w, uint32_t h, int k) {
    size_t off = 0;
    for (int i = 0; i < k; ++i) {
        uint32_t lw, lh;
        level_size(w, h, i, &lw, &lh);
        off += (size_t)lw * lh;
    }
    return off;
}
static int64_t clampi(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

static o3 upsample_at(const o3 *coarse, uint32_t cw, uint32_t ch, uint32_t x, uint32_t y) { /* bloom.rs:80-90 */
    uint32_t mx = x / 2, my = y / 2;
    int64_t ox = (x % 2 == 0) ? (int64_t)mx - 1 : (int64_t)mx + 1;
    int64_t oy = (y % 2 == 0) ? (int64_t)my - 1 : (int64_t)my + 1;
    uint32_t x1 = (uint32_t)clampi(ox, 0, (int64_t)cw - 1), y1 = (uint32_t)clampi(oy, 0, (int64_t)ch - 1);
    o3 r0 = add3(mul3(coarse[(size_t)my * cw + mx], 0.75f), mul3(coarse[(size_t)my * cw + x1], 0.25f));
    o3 r1 = add3(mul3(coarse[(size_t)y1 * cw + mx], 0.75f), mul3(coarse[(size_t)y1 * cw + x1], 0.25f));
    return add3(mul3(r0, 0.75f), mul3(r1, 0.25f));
}

void oracle_bloom(float *hdr3, uint32_t w, uint32_t h) {
    if (w == 0 || h == 0) return;
    o3 *img = (o3 *)hdr3;
    size_t total = level_offset(w, h, SPECREL_BLOOM_LEVELS + 1);
    o3 *pyr = (o3 *)calloc(total, sizeof(o3));
    memcpy(pyr, img, sizeof(o3) * (size_t)w * h);
    static const float DW[4] = {0.125f, 0.375f, 0.375f, 0.125f};
    for (int k = 1; k <= SPECREL_BLOOM_LEVELS; ++k) { /* downsample (bloom.rs:54-75) */
        uint32_t sw, sh, dw, dh;
        level_size(w, h, k - 1, &sw, &sh);
        level_size(w, h, k, &dw, &dh);
        const o3 *src = pyr + level_offset(w, h, k - 1);
        o3 *dst = pyr + level_offset(w, h, k);
        for (uint32_t y = 0; y < dh; ++y) {
            for (uint32_t x = 0; x < dw; ++x) {
                o3 acc = V3(0.0f, 0.0f, 0.0f);
                for (int j = 0; j < 4; ++j) {
                    size_t sy = (size_t)clampi(2 * (int64_t)y - 1 + j, 0, (int64_t)sh - 1);
                    o3 row = V3(0.0f, 0.0f, 0.0f);
                    for (int i = 0; i < 4; ++i) {
                        size_t sx = (size_t)clampi(2 * (int64_t)x - 1 + i, 0, (int64_t)sw - 1);
                        row = add3(row, mul3(src[sy * sw + sx], DW[i]));
                    }
                    acc = add3(acc, mul3(row, DW[j]));
                }
                dst[(size_t)y * dw + x] = acc;
            }
        }
    }
    for (int k = SPECREL_BLOOM_LEVELS - 1; k >= 1; --k) { /* bloom.rs:103-115 */
        uint32_t lw, lh, cw, ch;
        level_size(w, h, k, &lw, &lh);
        level_size(w, h, k + 1, &cw, &ch);
        o3 *fine = pyr + level_offset(w, h, k);
        const o3 *coarse = pyr + level_offset(w, h, k + 1);
        for (uint32_t y = 0; y < lh; ++y)
            for (uint32_t x = 0; x < lw; ++x)
                fine[(size_t)y * lw + x] = add3(fine[(size_t)y * lw + x], upsample_at(coarse, cw, ch, x, y));
    }
    uint32_t w1, h1;
    level_size(w, h, 1, &w1, &h1);
    const o3 *l1 = pyr + level_offset(w, h, 1);
    float halo = 0.04f / 6.0f;
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            o3 *px = &img[(size_t)y * w + x];
            *px = add3(mul3(*px, 1.0f - 0.04f), mul3(upsample_at(l1, w1, h1, x, y), halo));
        }
    free(pyr);
}

/* ---------------------------------------------------------------- tone map (renderer.rs:1102-1129) */
void oracle_gamma_lut(uint8_t out[SPECREL_LUT_SIZE]) {
    float g = 1.0f / 2.2f;
    for (int i = 0; i < SPECREL_LUT_SIZE; ++i) {
        float x = (float)i / (float)(SPECREL_LUT_SIZE - 1);
        float p = (float)pow((double)x, (double)g);
        float v = clampf_(p * 255.0f + 0.5f, 0.0f, 255.0f);
        out[i] = (uint8_t)sat_usize(v);
    }
}
static uint32_t pack_px(o3 c, const uint8_t *lut) {
    o3 mp = clamp3(V3(c.x / (1.0f + c.x * 0.35f), c.y / (1.0f + c.y * 0.35f), c.z / (1.0f + c.z * 0.35f)), 0.0f,
                   1.0f);
    float scale = (float)(SPECREL_LUT_SIZE - 1);
    uint64_t ir = sat_usize(mp.x * scale), ig = sat_usize(mp.y * scale), ib = sat_usize(mp.z * scale);
    return ((uint32_t)lut[ir] << 16) | ((uint32_t)lut[ig] << 8) | (uint32_t)lut[ib];
}
void oracle_pack(const float *hdr3, uint32_t n, uint32_t *out) {
    uint8_t lut[SPECREL_LUT_SIZE];
    oracle_gamma_lut(lut);
    for (uint32_t i = 0; i < n; ++i) out[i] = pack_px(V3(hdr3[3 * i], hdr3[3 * i + 1], hdr3[3 * i + 2]), lut);
}

/* ---------------------------------------------------------------- glyphs (hud.rs:473-532) */
static void load_glyphs(uint8_t g[14][7]) {
    static const uint8_t tab[14][7] = {
        {14, 17, 19, 21, 25, 17, 14}, {4, 12, 4, 4, 4, 4, 14},   {14, 17, 1, 2, 4, 8, 31},
        {30, 1, 1, 14, 1, 1, 30},     {2, 6, 10, 18, 31, 2, 2},  {31, 16, 30, 1, 1, 17, 14},
        {14, 16, 16, 30, 17, 17, 14}, {31, 1, 2, 4, 8, 8, 8},    {14, 17, 17, 14, 17, 17, 14},
        {14, 17, 17, 15, 1, 1, 14},   {0, 4, 4, 31, 4, 4, 0},    {0, 0, 0, 31, 0, 0, 0},
        {0, 0, 0, 0, 0, 12, 12},      {15, 16, 16, 14, 1, 1, 30}};
    memcpy(g, tab, sizeof tab);
}
int oracle_digital_tau(float u, float v, float tau) {
    uint8_t g[14][7];
    load_glyphs(g);
    return digital_tau(g, u, v, tau);
}

/* ---------------------------------------------------------------- render_frame (renderer.rs:327-412) */
typedef struct {
    o_ctx base;
    o_stars stars;
    uint32_t w, h;
    float half_w, half_h, pixel_pitch;
    int draw_stars;
    o4 ray_origin;
    float *hdr;
    uint32_t *hit_index;
    float *lambda, *tau;
    const uint32_t *rows;
    uint32_t nrows;
    atomic_uint next_row;
    pthread_mutex_t lock;
    oracle_counts *counts;
    const o_touch *touch;
} o_job;

/* The ray of pixel (px, py) (renderer.rs:334-384): normalised device coordinates scaled by the
 * frame's half extents, the local direction normalised (twice on the light cones, as
 * past_light_ray_direction / future_light_ray_direction normalise again, minkowski.rs:564-575),
 * then in the tetrad's frame. *sx, *sy receive the scaled coordinates. */
static o4 pixel_ray(const specrel_frame *f, uint32_t w, uint32_t h, float half_w, float half_h, uint32_t px,
                    uint32_t py, float *sx_out, float *sy_out) {
    float ndc_y = 1.0f - 2.0f * (((float)py + 0.5f) / (float)h);
    float sy = ndc_y * half_h;
    float ndc_x = 2.0f * (((float)px + 0.5f) / (float)w) - 1.0f;
    float sx = ndc_x * half_w;
    o3 local_dir = norm3(V3(sx, sy, 1.0f));
    o4 e0 = ld4(f->e[0]), e1 = ld4(f->e[1]), e2 = ld4(f->e[2]), e3 = ld4(f->e[3]);
    o4 dir;
    if (f->slice_mode == SPECREL_VIEW_PAST_CONE) { /* past_light_ray_direction */
        o3 n = norm3(local_dir);
        dir = add4(add4(add4(neg4(e0), mul4(e1, n.x)), mul4(e2, n.y)), mul4(e3, n.z));
    } else if (f->slice_mode == SPECREL_VIEW_FRONTIER) { /* future_light_ray_direction */
        o3 n = norm3(local_dir);
        dir = add4(add4(add4(e0, mul4(e1, n.x)), mul4(e2, n.y)), mul4(e3, n.z));
    } else { /* local_spatial_to_world */
        dir = add4(add4(mul4(e1, local_dir.x), mul4(e2, local_dir.y)), mul4(e3, local_dir.z));
    }
    *sx_out = sx;
    *sy_out = sy;
    return dir;
}

/* One image row: HDR into hdr_row, hit / lambda / tau (when asked) at row `out_row` of those
 * arrays (the frame row, or the row's place in a sample). */
static void shade_row(o_job *J, uint32_t py, uint32_t out_row, float *hdr_row, oracle_counts *c) {
    const specrel_frame *f = J->base.f;
    o_ctx x = J->base;
    x.c = c;
    o_rec rec;
    o4 e0 = ld4(f->e[0]);
    for (uint32_t px = 0; px < J->w; ++px) {
        float sx, sy;
        o4 dir = pixel_ray(f, J->w, J->h, J->half_w, J->half_h, px, py, &sx, &sy);
        if (c) c->rays++;
        o_hit h;
        int found = trace(J->base.s, J->ray_origin, dir, 1e-3f, 180.0f, &h, c);
        o3 color;
        if (found) {
            if (c) c->hits++;
            if (tl_touch) touch_bit(tl_touch->hits, (uint32_t)h.slot);
            interpolate_hit(&J->base.s->attrs[h.slot], &h, &rec);
            color = shade_hit(&x, dir, &rec, NULL);
            if (f->num_haze > 0) color = haze(color, rec.event, (const float (*)[4])f->haze_origin, f->num_haze);
        } else {
            color = shade_sky(dir);
        }
        if (J->draw_stars) {
            float inv_len = 1.0f / sqrtf(sx * sx + sy * sy + 1.0f);
            float spread = J->pixel_pitch * inv_len * sqrtf(inv_len);
            float lmax = found ? h.lambda : INFINITY;
            color = add3(color, star_light(&J->stars, J->ray_origin, dir, e0, J->base.cone, 0.0f, spread, lmax, c));
        }
        hdr_row[3 * px] = color.x;
        hdr_row[3 * px + 1] = color.y;
        hdr_row[3 * px + 2] = color.z;
        size_t pi = (size_t)out_row * J->w + px;
        if (J->hit_index) J->hit_index[pi] = found ? J->base.s->order[h.slot] : SPECREL_NO_HIT;
        if (J->lambda) J->lambda[pi] = found ? h.lambda : INFINITY;
        if (J->tau) J->tau[pi] = found ? rec.proper_time : 0.0f;
    }
}

static void *worker(void *arg) {
    o_job *J = (o_job *)arg;
    tl_touch = J->touch;
    oracle_counts local;
    memset(&local, 0, sizeof local);
    for (;;) {
        uint32_t k = atomic_fetch_add(&J->next_row, 1);
        if (k >= J->nrows) break;
        uint32_t py = J->rows ? J->rows[k] : k;
        const uint32_t out_row = J->rows ? k : py;
        shade_row(J, py, out_row, J->hdr + (size_t)out_row * J->w * 3, J->counts ? &local : NULL);
    }
    if (J->counts) {
        pthread_mutex_lock(&J->lock);
        uint64_t *a = (uint64_t *)J->counts, *b = (uint64_t *)&local;
        for (size_t i = 0; i < sizeof(oracle_counts) / 8; ++i) a[i] += b[i];
        pthread_mutex_unlock(&J->lock);
    }
    tl_touch = NULL;
    return NULL;
}

static void setup_job(o_job *J, const specrel_scene_arrays *s, const specrel_frame *f, const uint8_t *lut) {
    memset(J, 0, sizeof *J);
    J->base.s = s;
    J->base.f = f;
    J->base.lut = lut;
    load_glyphs(J->base.glyphs);
    J->w = f->size[0];
    J->h = f->size[1];
    J->half_w = f->half_wh[0];
    J->half_h = f->half_wh[1];
    J->pixel_pitch = 2.0f * J->half_h / (float)J->h;
    o4 cam = ld4(f->cam_pos), e0 = ld4(f->e[0]);
    J->base.e0 = e0;
    switch (f->slice_mode) {
        case SPECREL_VIEW_FRONTIER: J->ray_origin = sub4(cam, mul4(e0, f->frontier_view_offset)); break;
        case SPECREL_VIEW_PROPHECY: J->ray_origin = add4(cam, mul4(e0, f->prophecy_delta)); break;
        default: J->ray_origin = cam; break;
    }
    J->base.origin = J->ray_origin;
    J->base.cone = f->slice_mode == SPECREL_VIEW_PAST_CONE ? -1.0f : (f->slice_mode == SPECREL_VIEW_FRONTIER ? 1.0f : 0.0f);
    J->stars.words = s->star_words;
    J->stars.num_stars = s->num_stars;
    J->draw_stars = f->mode == SPECREL_MODE_SHADED && s->num_stars > 0;
    pthread_mutex_init(&J->lock, NULL);
}

static void run_workers(o_job *J, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads == 1) {
        worker(J);
        return;
    }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, worker, J);
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    free(th);
}

int oracle_render(const specrel_scene_arrays *s, const specrel_frame *f, int nthreads, uint32_t *rgb,
                  uint32_t *hit_index, float *lambda, float *tau, float *hdr_out, oracle_counts *counts) {
    uint32_t w = f->size[0], h = f->size[1];
    uint8_t lut[SPECREL_LUT_SIZE];
    oracle_gamma_lut(lut);
    o_job J;
    setup_job(&J, s, f, lut);
    J.hdr = (float *)malloc(sizeof(float) * 3 * (size_t)w * h);
    J.hit_index = hit_index;
    J.lambda = lambda;
    J.tau = tau;
    J.nrows = h;
    J.counts = counts;
    o_touch touch = {NULL, NULL, NULL};
    const size_t wn = ((size_t)s->num_nodes + 31) / 32, wp = ((size_t)s->num_prims + 31) / 32;
    if (counts) {
        touch.nodes = (_Atomic uint32_t *)calloc(wn + 1, 4);
        touch.prims = (_Atomic uint32_t *)calloc(wp + 1, 4);
        touch.hits = (_Atomic uint32_t *)calloc(wp + 1, 4);
        J.touch = &touch;
    }
    atomic_init(&J.next_row, 0);
    run_workers(&J, nthreads);
    if (counts) {
        counts->distinct_nodes = counts->distinct_prims = counts->distinct_hits = 0;
        for (size_t i = 0; i < wn; ++i) counts->distinct_nodes += (uint64_t)__builtin_popcount(touch.nodes[i]);
        for (size_t i = 0; i < wp; ++i) counts->distinct_prims += (uint64_t)__builtin_popcount(touch.prims[i]);
        for (size_t i = 0; i < wp; ++i) counts->distinct_hits += (uint64_t)__builtin_popcount(touch.hits[i]);
        free((void *)touch.nodes), free((void *)touch.prims), free((void *)touch.hits);
    }
    if (hdr_out) memcpy(hdr_out, J.hdr, sizeof(float) * 3 * (size_t)w * h);
    oracle_bloom(J.hdr, w, h);
    if (rgb)
        for (size_t i = 0; i < (size_t)w * h; ++i)
            rgb[i] = pack_px(V3(J.hdr[3 * i], J.hdr[3 * i + 1], J.hdr[3 * i + 2]), lut);
    free(J.hdr);
    pthread_mutex_destroy(&J.lock);
    return 0;
}

int oracle_render_rows(const specrel_scene_arrays *s, const specrel_frame *f, int nthreads, uint32_t row0,
                       uint32_t step, uint32_t nrows, float *hdr_rows, uint32_t *hit_rows, float *lambda_rows,
                       float *tau_rows, oracle_counts *counts) {
    uint8_t lut[SPECREL_LUT_SIZE];
    oracle_gamma_lut(lut);
    o_job J;
    setup_job(&J, s, f, lut);
    uint32_t *rows = (uint32_t *)malloc(sizeof(uint32_t) * (nrows ? nrows : 1));
    for (uint32_t k = 0; k < nrows; ++k) rows[k] = row0 + k * step;
    J.rows = rows;
    J.nrows = nrows;
    J.hdr = hdr_rows;
    J.hit_index = hit_rows;
    J.lambda = lambda_rows;
    J.tau = tau_rows;
    J.counts = counts;
    atomic_init(&J.next_row, 0);
    run_workers(&J, nthreads);
    free(rows);
    pthread_mutex_destroy(&J.lock);
    return 0;
}

/* ---------------------------------------------------------------- forward state for the FP64 adjoint reference */
static void pixel_state(o_job *J, uint32_t px, uint32_t py, oracle_pixel_state *out) {
    const specrel_scene_arrays *s = J->base.s;
    const specrel_frame *f = J->base.f;
    memset(out, 0, sizeof *out);
    out->slot = -1;
    /* the ray of shade_row (past light cone) */
    float ndc_y = 1.0f - 2.0f * (((float)py + 0.5f) / (float)J->h);
    float sy = ndc_y * J->half_h;
    float ndc_x = 2.0f * (((float)px + 0.5f) / (float)J->w) - 1.0f;
    float sx = ndc_x * J->half_w;
    o3 local_dir = norm3(V3(sx, sy, 1.0f));
    o3 n = norm3(local_dir);
    out->local_dir[0] = n.x;
    out->local_dir[1] = n.y;
    out->local_dir[2] = n.z;
    o4 e0 = ld4(f->e[0]), e1 = ld4(f->e[1]), e2 = ld4(f->e[2]), e3 = ld4(f->e[3]);
    o4 dir = add4(add4(add4(neg4(e0), mul4(e1, n.x)), mul4(e2, n.y)), mul4(e3, n.z));
    o_hit h;
    if (!trace(s, J->ray_origin, dir, 1e-3f, 180.0f, &h, NULL)) return;
    o_rec rec;
    interpolate_hit(&s->attrs[h.slot], &h, &rec);
    o_mat m = load_mat(s, rec.material);
    out->slot = h.slot;
    out->material = rec.material;
    if (h.bary[0] > 0.0f && h.bary[1] > 0.0f && h.bary[2] > 0.0f && h.bary[3] > 0.0f) out->flags |= ORACLE_PX_INSIDE;
    out->pattern = (int32_t)m.pattern;
    o3 nw;
    if (m.pattern == SPECREL_PATTERN_CLOCK || m.pattern == SPECREL_PATTERN_LIGHT_CLOCK) {
        nw = surface_albedo(&J->base, &rec, &m);
        out->dark = 1.0f;
    } else {
        int32_t gu = sat_i32(floorf(rec.uv_u * m.grid)), gv = sat_i32(floorf(rec.uv_v * m.grid));
        int32_t sum = (int32_t)((uint32_t)gu + (uint32_t)gv);
        o3 col0 = ((((sum % 2) + 2) % 2) == 0) ? m.base : m.checker;
        float dark = 1.0f;
        if (m.pattern != SPECREL_PATTERN_CHECKER) {
            float ut = remeu(rec.uv_u * m.grid, 1.0f), vt = remeu(rec.uv_v * m.grid, 1.0f);
            float bd = fmn(fmn(ut, 1.0f - ut), fmn(vt, 1.0f - vt));
            if (bd < 0.025f) dark = 0.35f;
        }
        out->col0[0] = col0.x;
        out->col0[1] = col0.y;
        out->col0[2] = col0.z;
        out->dark = dark;
        nw = mul3(col0, dark);
    }
    out->nw_albedo[0] = nw.x;
    out->nw_albedo[1] = nw.y;
    out->nw_albedo[2] = nw.z;
    /* the f32 branch taken at every kink of the differentiable part */
    {
        o3 tb = rec.tri_bary;
        float mtb = fmn(fmn(tb.x, tb.y), tb.z);
        out->mtb_k = (tb.x <= tb.y) ? (tb.x <= tb.z ? 0 : 2) : (tb.y <= tb.z ? 1 : 2);
        if (m.pattern != SPECREL_PATTERN_CLOCK && m.pattern != SPECREL_PATTERN_LIGHT_CLOCK && mtb < m.wire) {
            float et = mtb / m.wire;
            if (et > 0.0f && et < 1.0f) out->flags |= ORACLE_PX_WIRE;
        }
        o3 light = V3(s->light_pos[1], s->light_pos[2], s->light_pos[3]);
        o3 s2l = sub3(light, spat(rec.event));
        float len = len3(s2l), dist = fmx(len, 1e-4f);
        if (len > 1e-4f) out->flags |= ORACLE_PX_DIST_FREE;
        float t_emit = rec.event.t - dist;
        o4 to_light = mul4(sub4(V4(t_emit, light.x, light.y, light.z), rec.event), 1.0f / dist);
        o4 vs[2] = {to_light, dir};
        for (int k = 0; k < 2; ++k) {
            o4 v = vs[k];
            float sd = mdot(v, rec.four_vel);
            o4 vp = add4(v, mul4(rec.four_vel, sd));
            float q = mdot(vp, vp);
            if (q > 1e-12f) {
                float r = mdot(vp, rec.normal) / sqrtf(q);
                if (r > -1.0f && r < 1.0f) out->flags |= k == 0 ? ORACLE_PX_RL_FREE : ORACLE_PX_RV_FREE;
            }
        }
        if (fmx(rest_frame_cos(to_light, rec.normal, rec.four_vel), 0.0f) > 0.0f) out->flags |= ORACLE_PX_CL;
        if (fmx(-rest_frame_cos(dir, rec.normal, rec.four_vel), 0.0f) > 0.0f) out->flags |= ORACLE_PX_CV;
        o3 sp = spat(dir);
        float sl = len3(sp);
        if (sl > 1e-12f) {
            o3 su = mul3(sp, 1.0f / sl);
            float hz = su.y * 0.5f + 0.5f;
            if (hz > 0.0f && hz < 1.0f) out->flags |= ORACLE_PX_HZ_FREE;
        }
    }
    /* the shadow decision of compute_illumination (null shadows) */
    if (m.pattern == SPECREL_PATTERN_CHECKER && f->shadow_mode == SPECREL_SHADOW_NULL) {
        o3 light = V3(s->light_pos[1], s->light_pos[2], s->light_pos[3]);
        o3 s2l = sub3(light, spat(rec.event));
        float dist = fmx(len3(s2l), 1e-4f);
        float t_emit = rec.event.t - dist;
        o4 to_light = mul4(sub4(V4(t_emit, light.x, light.y, light.z), rec.event), 1.0f / dist);
        float cl = fmx(rest_frame_cos(to_light, rec.normal, rec.four_vel), 0.0f);
        if (cl > 0.0f) out->in_shadow = occluded(s, rec.event, to_light, 2e-3f, dist - 2e-3f, NULL);
    }
    o3 color = shade_hit(&J->base, dir, &rec, NULL);
    out->color[0] = color.x;
    out->color[1] = color.y;
    out->color[2] = color.z;
}

typedef struct {
    o_job *J;
    oracle_pixel_state *out;
    atomic_uint next;
} o_state_job;

static void *state_worker(void *arg) {
    o_state_job *S = (o_state_job *)arg;
    for (;;) {
        uint32_t py = atomic_fetch_add(&S->next, 1);
        if (py >= S->J->h) break;
        for (uint32_t px = 0; px < S->J->w; ++px) pixel_state(S->J, px, py, &S->out[(size_t)py * S->J->w + px]);
    }
    return NULL;
}

int oracle_pixel_states(const specrel_scene_arrays *s, const specrel_frame *f, int nthreads, oracle_pixel_state *out) {
    if (f->slice_mode != SPECREL_VIEW_PAST_CONE || f->mode != SPECREL_MODE_SHADED) return -1;
    uint8_t lut[SPECREL_LUT_SIZE];
    oracle_gamma_lut(lut);
    o_job J;
    setup_job(&J, s, f, lut);
    o_state_job S;
    S.J = &J;
    S.out = out;
    atomic_init(&S.next, 0);
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, state_worker, &S);
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    free(th);
    pthread_mutex_destroy(&J.lock);
    return 0;
}

/* ---------------------------------------------------------------- test entries (pinning tests) */
/* shade_hit on a hit record given field by field (renderer.rs:414-489), with the scene's light,
 * materials and BVH (for the shadow rays) and the frame's mode, shadow mode, camera and flashes.
 * out: the colour before haze and stars; parts (may be NULL): the intermediate values. */
int oracle_shade_record(const specrel_scene_arrays *s, const specrel_frame *f, const float dir[4],
                        const oracle_record *r, float out[3], oracle_shade_parts *parts) {
    if (r->material >= s->num_materials) return -1;
    uint8_t lut[SPECREL_LUT_SIZE];
    oracle_gamma_lut(lut);
    o_job J;
    setup_job(&J, s, f, lut);
    o_rec rec;
    rec.lambda = r->lambda;
    rec.event = ld4(r->event);
    memcpy(rec.tet_bary, r->tet_bary, 16);
    rec.tri_bary = V3(r->tri_bary[0], r->tri_bary[1], r->tri_bary[2]);
    rec.uv_u = r->uv[0];
    rec.uv_v = r->uv[1];
    rec.proper_time = r->proper_time;
    rec.grad_u = r->tau_uv_grad[0];
    rec.grad_v = r->tau_uv_grad[1];
    rec.uvws_u = r->uv_world_scale[0];
    rec.uvws_v = r->uv_world_scale[1];
    rec.normal = ld4(r->normal);
    rec.four_vel = ld4(r->four_velocity);
    rec.decided_at = r->decided_at;
    rec.material = r->material;
    if (parts) memset(parts, 0, sizeof *parts);
    o3 c = shade_hit(&J.base, ld4(dir), &rec, parts);
    out[0] = c.x, out[1] = c.y, out[2] = c.z;
    pthread_mutex_destroy(&J.lock);
    return 0;
}

/* The ray origin and direction of pixel (px, py) of frame f (renderer.rs:334-384). */
void oracle_pixel_ray(const specrel_frame *f, uint32_t px, uint32_t py, float origin[4], float dir[4]) {
    o_job J;
    uint8_t lut[SPECREL_LUT_SIZE] = {0};
    specrel_scene_arrays empty;
    memset(&empty, 0, sizeof empty);
    setup_job(&J, &empty, f, lut);
    float sx, sy;
    o4 d = pixel_ray(f, J.w, J.h, J.half_w, J.half_h, px, py, &sx, &sy);
    origin[0] = J.ray_origin.t, origin[1] = J.ray_origin.x, origin[2] = J.ray_origin.y, origin[3] = J.ray_origin.z;
    dir[0] = d.t, dir[1] = d.x, dir[2] = d.y, dir[3] = d.z;
    pthread_mutex_destroy(&J.lock);
}

void oracle_sky_gradient(const float dir[4], float out[3]) {
    o3 c = sky_gradient(ld4(dir));
    out[0] = c.x, out[1] = c.y, out[2] = c.z;
}

void oracle_distance_fog(const float color[3], const float dir[4], float lambda, float out[3]) {
    o3 c = distance_fog(V3(color[0], color[1], color[2]), ld4(dir), lambda);
    out[0] = c.x, out[1] = c.y, out[2] = c.z;
}

/* The f32 helpers over arrays: fn 0 exp, 1 sin, 2 cos, 3 atan2 (y = a, x = b), 4 acos. */
void oracle_math_vec(int fn, const float *a, const float *b, float *out, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) {
        switch (fn) {
            case 0: out[i] = ex(a[i]); break;
            case 1: out[i] = sn(a[i]); break;
            case 2: out[i] = cs(a[i]); break;
            case 3: out[i] = at2(a[i], b[i]); break;
            default: out[i] = ac(a[i]); break;
        }
    }
}
