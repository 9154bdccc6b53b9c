// FP64 forward-mode reference for the image adjoint. TEST INFRASTRUCTURE ONLY: the GPU
// adjoint (specrel_adjoint) is checked against it; nothing in the product links or calls it.
//
// For every valid pixel the shaded colour is evaluated again in double precision with dual
// numbers carrying 13 tangents: the observer event (4), the observer's rapidity vector (3) and
// rotation vector (3) -- through the tetrad they define, not through tetrad legs -- and the boost
// rapidity (3) of the object the pixel sees. The discrete forward decisions (hit slot, shadow,
// albedo cell and grid line) and the branch taken at every kink of the shading (wireframe band,
// max(.,0) of the cosines, clamps) come from the f32 oracle (oracle_pixel_states), which is
// bit-exact with the renderer: the derivative is the one of the branch the forward evaluated. The model is the renderer's (renderer.rs:414-489, 544-580, 585-653,
// 691-695; mesh4d.rs:175-211, 380-425) with every float operation in double.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle.h"

namespace {

constexpr int NT = 13;  // event 0-3, rapidity 4-6, rotation 7-9, object boost 10-12

struct D {
    double v;
    double d[NT];
};
D cst(double v) {
    D r;
    r.v = v;
    std::memset(r.d, 0, sizeof r.d);
    return r;
}
D var(double v, int k) {
    D r = cst(v);
    r.d[k] = 1.0;
    return r;
}
D operator+(const D &a, const D &b) {
    D r;
    r.v = a.v + b.v;
    for (int k = 0; k < NT; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
D operator-(const D &a, const D &b) {
    D r;
    r.v = a.v - b.v;
    for (int k = 0; k < NT; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
D operator-(const D &a) {
    D r;
    r.v = -a.v;
    for (int k = 0; k < NT; ++k) r.d[k] = -a.d[k];
    return r;
}
D operator*(const D &a, const D &b) {
    D r;
    r.v = a.v * b.v;
    for (int k = 0; k < NT; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
D operator*(const D &a, double s) {
    D r;
    r.v = a.v * s;
    for (int k = 0; k < NT; ++k) r.d[k] = a.d[k] * s;
    return r;
}
D operator*(double s, const D &a) { return a * s; }
D operator+(const D &a, double s) {
    D r = a;
    r.v += s;
    return r;
}
D operator+(double s, const D &a) { return a + s; }
D operator-(const D &a, double s) { return a + (-s); }
D operator-(double s, const D &a) { return (-a) + s; }
D operator/(const D &a, const D &b) {
    D r;
    r.v = a.v / b.v;
    for (int k = 0; k < NT; ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
    return r;
}
D operator/(const D &a, double s) { return a * (1.0 / s); }
D operator/(double s, const D &a) { return cst(s) / a; }
D func(const D &a, double f, double df) {
    D r;
    r.v = f;
    for (int k = 0; k < NT; ++k) r.d[k] = a.d[k] * df;
    return r;
}
D sqrt(const D &a) {
    double s = std::sqrt(a.v);
    return func(a, s, 0.5 / s);
}
D exp(const D &a) {
    double e = std::exp(a.v);
    return func(a, e, e);
}
D sinh(const D &a) { return func(a, std::sinh(a.v), std::cosh(a.v)); }
D cosh(const D &a) { return func(a, std::cosh(a.v), std::sinh(a.v)); }
D sin(const D &a) { return func(a, std::sin(a.v), std::cos(a.v)); }
D cos(const D &a) { return func(a, std::cos(a.v), -std::sin(a.v)); }

struct V4 {
    D c[4];  // t, x, y, z
};
V4 v4c(const float *p) {
    V4 r;
    for (int i = 0; i < 4; ++i) r.c[i] = cst(p[i]);
    return r;
}
V4 add(const V4 &a, const V4 &b) {
    V4 r;
    for (int i = 0; i < 4; ++i) r.c[i] = a.c[i] + b.c[i];
    return r;
}
V4 sub(const V4 &a, const V4 &b) {
    V4 r;
    for (int i = 0; i < 4; ++i) r.c[i] = a.c[i] - b.c[i];
    return r;
}
V4 scale(const V4 &a, const D &s) {
    V4 r;
    for (int i = 0; i < 4; ++i) r.c[i] = a.c[i] * s;
    return r;
}
D edot(const V4 &a, const V4 &b) { return a.c[0] * b.c[0] + a.c[1] * b.c[1] + a.c[2] * b.c[2] + a.c[3] * b.c[3]; }
D mdot(const V4 &a, const V4 &b) { return -(a.c[0] * b.c[0]) + a.c[1] * b.c[1] + a.c[2] * b.c[2] + a.c[3] * b.c[3]; }

// 4x4 matrices of duals, row-major, acting on (t, x, y, z).
struct M4 {
    D m[4][4];
};
M4 ident() {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = cst(i == j ? 1.0 : 0.0);
    return r;
}
V4 apply(const M4 &M, const V4 &x) {
    V4 r;
    for (int i = 0; i < 4; ++i) r.c[i] = M.m[i][0] * x.c[0] + M.m[i][1] * x.c[1] + M.m[i][2] * x.c[2] + M.m[i][3] * x.c[3];
    return r;
}
M4 mul(const M4 &A, const M4 &B) {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = A.m[i][0] * B.m[0][j] + A.m[i][1] * B.m[1][j] + A.m[i][2] * B.m[2][j] + A.m[i][3] * B.m[3][j];
    return r;
}
// K(phi): the boost generator with K[0][i] = K[i][0] = phi_i.
M4 boost_gen(const D phi[3]) {
    M4 K;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) K.m[i][j] = cst(0.0);
    for (int i = 0; i < 3; ++i) K.m[0][i + 1] = K.m[i + 1][0] = phi[i];
    return K;
}
// exp(K(phi)) = I + sinh(r)/r K + (cosh r - 1)/r^2 K^2.
M4 boost(const D phi[3]) {
    M4 K = boost_gen(phi), K2 = mul(K, K), r = ident();
    D r2 = phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2];
    D c1, c2;
    if (r2.v < 1e-12) {
        c1 = 1.0 + r2 * (1.0 / 6.0) + r2 * r2 * (1.0 / 120.0);
        c2 = 0.5 + r2 * (1.0 / 24.0) + r2 * r2 * (1.0 / 720.0);
    } else {
        D rr = sqrt(r2);
        c1 = sinh(rr) / rr;
        c2 = (cosh(rr) - 1.0) / r2;
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = r.m[i][j] + c1 * K.m[i][j] + c2 * K2.m[i][j];
    return r;
}
// Rodrigues rotation of the spatial axes by rotation vector w, as a 4x4 matrix.
M4 rotation(const D w[3]) {
    M4 W;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) W.m[i][j] = cst(0.0);
    W.m[1][2] = -w[2];
    W.m[1][3] = w[1];
    W.m[2][1] = w[2];
    W.m[2][3] = -w[0];
    W.m[3][1] = -w[1];
    W.m[3][2] = w[0];
    M4 W2 = mul(W, W), r = ident();
    D a2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    D s1, s2;
    if (a2.v < 1e-12) {
        s1 = 1.0 - a2 * (1.0 / 6.0) + a2 * a2 * (1.0 / 120.0);
        s2 = 0.5 - a2 * (1.0 / 24.0) + a2 * a2 * (1.0 / 720.0);
    } else {
        D a = sqrt(a2);
        s1 = sin(a) / a;
        s2 = (1.0 - cos(a)) / a2;
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = r.m[i][j] + s1 * W.m[i][j] + s2 * W2.m[i][j];
    return r;
}

// rest_frame_cos on the branch the f32 forward took: free (differentiable) or clamped/degenerate
// (locally constant).
D rest_frame_cos(const V4 &v, const V4 &n, const V4 &u, bool free_branch) {
    D s = mdot(v, u);
    V4 vp = add(v, scale(u, s));
    D q = mdot(vp, vp);
    if (!(q.v > 0.0)) return cst(0.0);
    D r = mdot(vp, n) / sqrt(q);
    if (free_branch) return r;
    return cst(r.v <= -1.0 ? -1.0 : (r.v >= 1.0 ? 1.0 : (q.v > 1e-12 ? r.v : 0.0)));
}

uint32_t fnv(uint32_t h, uint32_t w) { return (h ^ w) * 16777619u; }
uint32_t fbits(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    return u;
}

struct Ctx {
    const specrel_scene_arrays *s;
    const specrel_frame *f;
    const double *theta;  // event[4], rapidity[3], rotation[3]
    const uint32_t *ofirst, *ocount;
    const float *ocenter;
    uint32_t n_obj;
    bool frame_values;
};

uint32_t find_object(const Ctx &c, uint32_t prim) {
    uint32_t lo = 0, hi = c.n_obj;
    while (hi > lo) {
        uint32_t mid = (lo + hi) >> 1;
        if (c.ofirst[mid] <= prim) lo = mid + 1;
        else hi = mid;
    }
    if (lo == 0) return ~0u;
    return prim - c.ofirst[lo - 1] < c.ocount[lo - 1] ? lo - 1 : ~0u;
}

// The colour of a hit pixel as a function of the 13 tangents.
void pixel_color(const Ctx &c, const oracle_pixel_state &st, uint32_t obj, D out[3]) {
    const specrel_scene_arrays *s = c.s;
    D E[4], phi[3], om[3], del[3];
    for (int k = 0; k < 4; ++k) E[k] = var(c.theta[k], k);
    for (int k = 0; k < 3; ++k) phi[k] = var(c.theta[4 + k], 4 + k);
    for (int k = 0; k < 3; ++k) om[k] = var(c.theta[7 + k], 7 + k);
    for (int k = 0; k < 3; ++k) del[k] = var(0.0, 10 + k);
    M4 L = mul(boost(phi), rotation(om));
    // With frame_values the values are the frame's f32 event and legs (what the renderer used)
    // and the tangents those of the double-precision parameterisation; otherwise both come from
    // the parameters in double.
    V4 e[4];
    for (int a = 0; a < 4; ++a)
        for (int i = 0; i < 4; ++i) {
            e[a].c[i] = L.m[i][a];
            if (c.frame_values) e[a].c[i].v = (double)c.f->e[a][i];
        }
    V4 o;
    for (int i = 0; i < 4; ++i) {
        o.c[i] = E[i];
        if (c.frame_values) o.c[i].v = (double)c.f->cam_pos[i];
    }
    V4 d;
    for (int i = 0; i < 4; ++i)
        d.c[i] = -e[0].c[i] + e[1].c[i] * (double)st.local_dir[0] + e[2].c[i] * (double)st.local_dir[1] +
                 e[3].c[i] * (double)st.local_dir[2];
    const specrel_geom &g = s->geoms[st.slot];
    const specrel_attr &A = s->attrs[st.slot];
    V4 v0 = v4c(g.v0), N = v4c(g.n), B1 = v4c(g.b1), B2 = v4c(g.b2), B3 = v4c(g.b3);
    V4 nrm = v4c(A.normal), u4 = v4c(A.four_vel);
    if (obj != ~0u) {
        // The object's worldtube boosted about its centre: first order in the boost is exact here.
        M4 K = boost_gen(del), K2 = mul(K, K), Lb = ident(), Li = ident();
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                Lb.m[i][j] = Lb.m[i][j] + K.m[i][j] + K2.m[i][j] * 0.5;
                Li.m[i][j] = Li.m[i][j] - K.m[i][j] + K2.m[i][j] * 0.5;
            }
        V4 cen = v4c(c.ocenter + 4 * obj);
        v0 = add(cen, apply(Lb, sub(v0, cen)));
        N = apply(Li, N);
        B1 = apply(Li, B1);
        B2 = apply(Li, B2);
        B3 = apply(Li, B3);
        nrm = apply(Lb, nrm);
        u4 = apply(Lb, u4);
    }
    // intersect_fast (mesh4d.rs:175-211)
    D den = edot(N, d);
    D lam = edot(N, sub(v0, o)) / den;
    V4 X = add(o, scale(d, lam));
    V4 q = sub(X, v0);
    D uu = edot(B1, q), vv = edot(B2, q), ww = edot(B3, q);
    D b[4] = {1.0 - uu - vv - ww, uu, vv, ww};
    // albedo (renderer.rs:544-580): smooth only through the wireframe blend
    const specrel_material &mat = s->materials[st.material];
    D alb[3];
    if (st.pattern == SPECREL_PATTERN_CLOCK || st.pattern == SPECREL_PATTERN_LIGHT_CLOCK) {
        for (int k = 0; k < 3; ++k) alb[k] = cst(st.nw_albedo[k]);
    } else {
        const float *tbs[3] = {A.tb_x, A.tb_y, A.tb_z};
        D tb[3];
        for (int k = 0; k < 3; ++k) tb[k] = b[0] * (double)tbs[k][0] + b[1] * (double)tbs[k][1] + b[2] * (double)tbs[k][2] + b[3] * (double)tbs[k][3];
        D mtb = tb[st.mtb_k];
        double wire = mat.params[0];
        for (int k = 0; k < 3; ++k) {
            D col = cst(st.col0[k]);
            if (st.flags & ORACLE_PX_WIRE) {
                D et = mtb / wire;
                col = 0.07 * (1.0 - et) + col * et;
            }
            alb[k] = col * (double)st.dark;
        }
    }
    // compute_illumination (renderer.rs:585-653), null shadows
    D light[3] = {cst(s->light_pos[1]), cst(s->light_pos[2]), cst(s->light_pos[3])};
    D s2l[3] = {light[0] - X.c[1], light[1] - X.c[2], light[2] - X.c[3]};
    D len = sqrt(s2l[0] * s2l[0] + s2l[1] * s2l[1] + s2l[2] * s2l[2]);
    D dist = (st.flags & ORACLE_PX_DIST_FREE) ? len : cst(1e-4);
    D t_emit = X.c[0] - dist;
    V4 tl;
    tl.c[0] = (t_emit - X.c[0]) / dist;
    for (int k = 0; k < 3; ++k) tl.c[k + 1] = s2l[k] / dist;
    D rl = rest_frame_cos(tl, nrm, u4, st.flags & ORACLE_PX_RL_FREE);
    D rv = rest_frame_cos(d, nrm, u4, st.flags & ORACLE_PX_RV_FREE);
    D cl = (st.flags & ORACLE_PX_CL) ? rl : cst(0.0);
    D cv = (st.flags & ORACLE_PX_CV) ? -rv : cst(0.0);
    D lighting;
    if (st.pattern == SPECREL_PATTERN_CHECKER) {
        double sf = st.in_shadow ? 0.32 : 1.0;
        D atten = 1.0 / (1.0 + 0.006 * dist * dist);
        lighting = (double)mat.base[3] + 0.30 * cv + 0.80 * sf * cl * atten;
    } else {
        lighting = 0.72 + 0.38 * cv + 0.22 * cl;
    }
    // distance_fog (renderer.rs:691-695) with sky_gradient (renderer.rs:1056-1062)
    D fog = exp(-0.008 * lam);
    D sl = sqrt(d.c[1] * d.c[1] + d.c[2] * d.c[2] + d.c[3] * d.c[3]);
    D hz = d.c[2] / sl * 0.5 + 0.5;
    if (!(st.flags & ORACLE_PX_HZ_FREE)) hz = cst(hz.v <= 0.0 ? 0.0 : (hz.v >= 1.0 ? 1.0 : hz.v));
    const double top[3] = {0.12, 0.16, 0.25}, bot[3] = {0.04, 0.06, 0.13};
    for (int k = 0; k < 3; ++k) {
        D sky = top[k] * (1.0 - hz) + bot[k] * hz;
        out[k] = alb[k] * lighting * fog + sky * (1.0 - fog);
    }
}

}  // namespace

extern "C" int oracle_adjoint_ref(const specrel_scene_arrays *s, const specrel_frame *f, const double theta[10],
                                  const float *target, const uint8_t *mask_in, const uint32_t *obj_first,
                                  const uint32_t *obj_count, const float *obj_center, uint32_t n_obj,
                                  int frame_values, int strict, int nthreads, double *loss_out, double *grad_out,
                                  uint64_t counts_out[6], uint8_t *mask_out, double *max_rel_color_err) {
    const uint32_t w = f->size[0], h = f->size[1];
    const size_t np = (size_t)w * h;
    std::vector<oracle_pixel_state> st(np);
    if (oracle_pixel_states(s, f, nthreads, st.data()) != 0) return -1;
    std::vector<float> hdr(np * 3);
    oracle_render(s, f, nthreads, nullptr, nullptr, nullptr, nullptr, hdr.data(), nullptr);
    Ctx c{s, f, theta, obj_first, obj_count, obj_center, n_obj, frame_values != 0};
    const uint32_t MISS = 0xFFFFFFFFu, STAR = 0xFFFFFFFEu;
    std::vector<uint32_t> sig(np, MISS), obj(np, ~0u);
    for (size_t i = 0; i < np; ++i) {
        if (st[i].slot < 0) continue;
        bool same = fbits(st[i].color[0]) == fbits(hdr[3 * i]) && fbits(st[i].color[1]) == fbits(hdr[3 * i + 1]) &&
                    fbits(st[i].color[2]) == fbits(hdr[3 * i + 2]);
        obj[i] = find_object(c, s->order[st[i].slot]);
        uint32_t hh = fnv(fnv(2166136261u, obj[i]), st[i].material);
        hh = fnv(fnv(fnv(hh, fbits(st[i].nw_albedo[0])), fbits(st[i].nw_albedo[1])), fbits(st[i].nw_albedo[2]));
        hh = fnv(hh, st[i].in_shadow ? 1u : 0u);
        const specrel_attr &A = s->attrs[st[i].slot];
        for (int k = 0; k < 4; ++k) hh = fnv(hh, fbits(A.normal[k]));
        for (int k = 0; k < 4; ++k) hh = fnv(hh, fbits(A.four_vel[k]));
        if (hh >= STAR) hh &= 0x7FFFFFFFu;
        sig[i] = same ? hh : STAR;
    }
    std::vector<uint8_t> valid(np, 0);
    uint64_t cnt[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t py = 0; py < h; ++py)
        for (uint32_t px = 0; px < w; ++px) {
            size_t i = (size_t)py * w + px;
            int cls;
            if (mask_in) cls = mask_in[i] ? 0 : (sig[i] == MISS ? 1 : (sig[i] == STAR ? 2 : 5));
            else if (sig[i] == MISS) cls = 1;
            else if (sig[i] == STAR) cls = 2;
            else if (px == 0 || py == 0 || px + 1 == w || py + 1 == h) cls = 3;
            else {
                cls = 0;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        size_t ni = (size_t)(py + dy) * w + (px + dx);
                        if (sig[ni] == sig[i]) continue;
                        if (sig[ni] == MISS || sig[ni] == STAR || obj[ni] != obj[i]) cls = 4;
                        else if (strict && cls == 0) cls = 5;
                    }
                if (cls == 0 && !(st[i].flags & ORACLE_PX_INSIDE)) cls = 5;
            }
            cnt[cls]++;
            valid[i] = cls == 0;
            if (mask_out) mask_out[i] = cls == 0;
        }
    const size_t ng = 10 + 3 * (size_t)n_obj;
    int nt = nthreads < 1 ? 1 : nthreads;
    std::vector<std::vector<double>> part(nt, std::vector<double>(ng + 1, 0.0));
    std::vector<double> maxerr(nt, 0.0);
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            std::vector<double> &g = part[t];
            for (size_t i = t; i < np; i += nt) {
                if (!valid[i]) continue;
                if (st[i].slot < 0 || sig[i] == STAR) {  // a caller's mask may hold misses: loss only
                    for (int k = 0; k < 3; ++k) {
                        double r = (double)hdr[3 * i + k] - (double)target[3 * i + k];
                        g[ng] += r * r;
                    }
                    continue;
                }
                D col[3];
                pixel_color(c, st[i], obj[i], col);
                for (int k = 0; k < 3; ++k) {
                    double r = col[k].v - (double)target[3 * i + k];
                    g[ng] += r * r;
                    for (int p = 0; p < 10; ++p) g[p] += 2.0 * r * col[k].d[p];
                    if (obj[i] != ~0u)
                        for (int a = 0; a < 3; ++a) g[10 + 3 * obj[i] + a] += 2.0 * r * col[k].d[10 + a];
                    double ref = std::fabs((double)st[i].color[k]);
                    double e = std::fabs(col[k].v - (double)st[i].color[k]) / (ref > 1e-3 ? ref : 1e-3);
                    if (e > maxerr[t]) maxerr[t] = e;
                }
            }
        });
    for (auto &x : th) x.join();
    std::vector<double> g(ng + 1, 0.0);
    double me = 0.0;
    for (int t = 0; t < nt; ++t) {
        for (size_t k = 0; k <= ng; ++k) g[k] += part[t][k];
        if (maxerr[t] > me) me = maxerr[t];
    }
    *loss_out = g[ng];
    for (size_t k = 0; k < ng; ++k) grad_out[k] = g[k];
    for (int k = 0; k < 6; ++k) counts_out[k] = cnt[k];
    if (max_rel_color_err) *max_rel_color_err = me;
    return 0;
}
