// Adjoint of the shaded image: the gradient of L = sum over valid pixels of |I - target|^2, where
// I is the linear (pre-bloom) HDR radiance, with respect to the observer event, the observer's
// tetrad legs and, per object, the rapidity of a Lorentz boost of its whole worldtube about a
// reference event. The forward render's per-pixel hit slot is the checkpoint: each pixel's
// intersection, interpolation and shading (with its shadow ray) are recomputed from it with the
// renderer's own float operations and then differentiated by hand.
//
// Differentiable model (renderer.rs:414-489 in RelativisticShaded mode, null or no shadows, no
// flashes or haze): the ray (renderer.rs:369-384), intersect_fast (mesh4d.rs:175-211), the
// interpolated triangle barycentrics (mesh4d.rs:380-425) through the wireframe blend of
// sample_surface_albedo (renderer.rs:544-580), compute_illumination (renderer.rs:585-653:
// rest-frame cosines, attenuation) and distance_fog (renderer.rs:691-695). Everything else
// the albedo depends on (checker cells, clock faces, grid lines) and the shadow test are locally
// constant; pixels at their discontinuities, silhouettes and misses are masked by the reduction.
#include "kernels.hpp"

namespace dev {

constexpr uint32_t SIG_MISS = 0xFFFFFFFFu, SIG_STAR = 0xFFFFFFFEu, NO_OBJ = 0xFFFFFFFFu;

// The backward pass runs in double on the f32 inputs of the forward (frame, records, attributes):
// the f32 forward fixes every discrete choice and the branch at every kink, the continuous
// intermediates are recomputed in double so that badly conditioned pixels (grazing or distant
// hits) do not lose the gradient to f32 cancellation.
struct D4 {
    double t, x, y, z;
};
__device__ __forceinline__ D4 d4(V4 a) { return D4{a.t, a.x, a.y, a.z}; }
__device__ __forceinline__ D4 d4(float4 a) { return D4{a.x, a.y, a.z, a.w}; }
__device__ __forceinline__ D4 operator+(D4 a, D4 b) { return D4{a.t + b.t, a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D4 operator-(D4 a, D4 b) { return D4{a.t - b.t, a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D4 operator*(D4 a, double s) { return D4{a.t * s, a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double edot(D4 a, D4 b) { return a.t * b.t + a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double mdot(D4 a, D4 b) { return -a.t * b.t + a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D4 eta(D4 x) { return D4{-x.t, x.x, x.y, x.z}; }
__device__ __forceinline__ D4 boost_gen(int a, D4 x) {
    D4 r = D4{a == 1 ? x.x : (a == 2 ? x.y : x.z), 0.0, 0.0, 0.0};
    if (a == 1) r.x = x.t;
    else if (a == 2) r.y = x.t;
    else r.z = x.t;
    return r;
}

// rest_frame_cos (minkowski.rs:364-380) on its free branch, and its backward for gradient g.
__device__ double rest_frame_cos_d(D4 v, D4 n, D4 u) {
    double s = mdot(v, u);
    D4 vp = v + u * s;
    return mdot(vp, n) / sqrt(mdot(vp, vp));
}
__device__ void rest_frame_cos_grad(D4 v, D4 n, D4 u, double g, D4 &dv, D4 &dn, D4 &du) {
    if (g == 0.0) return;
    double s = mdot(v, u);
    D4 vp = v + u * s;
    double q = mdot(vp, vp), p = mdot(vp, n), sq = sqrt(q);
    double dp = g / sq, dq = -0.5 * g * p / (q * sq);
    D4 dvp = eta(n) * dp + eta(vp) * (2.0 * dq);
    dn = dn + eta(vp) * dp;
    double ds = edot(dvp, u);
    dv = dv + dvp + eta(u) * ds;
    du = du + dvp * s + eta(v) * ds;
}

// The f32 forward's branch of rest_frame_cos: neither degenerate nor clamped.
__device__ __forceinline__ bool rfc_free(V4 v, V4 n, V4 u) {
    float s = mdot(v, u);
    V4 vp = v + u * s;
    float q = mdot(vp, vp);
    if (!(q > 1e-12f)) return false;
    float r = mdot(vp, n) / sqrtf(q);
    return r > -1.0f && r < 1.0f;
}

__device__ __forceinline__ uint32_t fnv(uint32_t h, uint32_t w) { return (h ^ w) * 16777619u; }

__device__ __forceinline__ uint32_t find_object(const uint32_t *first, const uint32_t *count, uint32_t n,
                                                uint32_t prim) {
    uint32_t lo = 0, hi = n;
    while (hi > lo) {
        uint32_t mid = (lo + hi) >> 1;
        if (first[mid] <= prim) lo = mid + 1;
        else hi = mid;
    }
    if (lo == 0) return NO_OBJ;
    uint32_t k = lo - 1;
    return prim - first[k] < count[k] ? k : NO_OBJ;
}

// Pass 1: per pixel, the loss term (every pixel) and, for hits, the gradient with respect to the
// ray origin and direction and to the hit object's boost, plus the pixel's shading signature
// (object, albedo without the wireframe blend, shadow flag, facet normal and four-velocity)
// for the mask.
template <int SK>
__global__ __launch_bounds__(256) void k_adj_pixel(AdjArgs a) {
    const uint32_t t = blockIdx.x, vi = find_view(a.V.level_start[0], a.V.n, t);
    const specrel_frame &f = a.V.frames[vi];
    const ViewDerived v = derive(f, 0);
    const uint32_t lt = t - a.V.level_start[0][vi], tiles_x = (v.w + TILE - 1) / TILE;
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const uint32_t px = (lt % tiles_x) * TILE + (wave & 1) * 8 + (lane & 7);
    const uint32_t py = (lt / tiles_x) * TILE + (wave >> 1) * 8 + (lane >> 3);
    __shared__ uint2 s_stack[StackSel<SK>::lds_entries];
    typename StackSel<SK>::type stk;
    typename StackSel<SK>::Private stk_priv;
    StackSel<SK>::init(stk, s_stack, tid, stk_priv);
    if (px >= v.w || py >= v.h) return;
    const size_t lp = (size_t)py * v.w + px, li = a.V.pix_start[vi] + lp;  // in the view, in the batch
    double out[ADJ_PIX];
    for (int k = 0; k < ADJ_PIX; ++k) out[k] = 0.0;
    uint32_t sig = SIG_MISS, obj = NO_OBJ;
    bool inside = false;
    const float4 hv = a.hdr[a.V.pyr_start[vi] + lp];
    const V3 hdr = V3{hv.x, hv.y, hv.z};
    const V3 tgt = V3{a.target[3 * li], a.target[3 * li + 1], a.target[3 * li + 2]};
    const V3 res = hdr - tgt;
    out[11] = (double)res.x * res.x + (double)res.y * res.y + (double)res.z * res.z;
    const uint32_t slot = a.slot[li];
    if (slot != SPECREL_NO_HIT) {
        const V4 o = v.origin, dir = ray_dir(v, f, px, py);
        const float4 *g = a.S.geoms + 5 * (size_t)slot;
        Hit hit;
        hit.slot = (int32_t)slot;
        intersect(g, o, dir, 1e-3f, 180.0f, hit.lambda, hit.event, hit.bary);
        inside = hit.bary[0] > 0.0f && hit.bary[1] > 0.0f && hit.bary[2] > 0.0f && hit.bary[3] > 0.0f;
        Rec rec;
        const float4 *A = a.S.attrs + 12 * (size_t)slot;
        interpolate(A, hit, rec);
        const Mat m = load_mat(a.S, rec.material);
        const V4 X = rec.event, nrm = rec.normal, u4 = rec.four_vel;
        const float lam = rec.lambda;

        // Forward, as shade_hit + compute_illumination + distance_fog evaluate it.
        const V3 alb = albedo(a.T, rec, m);
        V3 light = V3{a.S.light.y, a.S.light.z, a.S.light.w};
        V3 s2l = light - spatial(X);
        float len = length(s2l), dist = fmax_(len, 1e-4f);
        float t_emit = X.t - dist;
        V4 delta = from_ts(t_emit, light) - X;
        float inv = 1.0f / dist;
        V4 to_light = delta * inv;
        float rl = rest_frame_cos(to_light, nrm, u4), cl = fmax_(rl, 0.0f);
        float rv = rest_frame_cos(dir, nrm, u4), cv = fmax_(-rv, 0.0f);
        const bool checker = m.pattern == SPECREL_PATTERN_CHECKER;
        bool in_shadow = false;
        float lighting, atten = 0.0f, sf = 1.0f;
        if (checker) {
            if (cl > 0.0f && f.shadow_mode != SPECREL_SHADOW_OFF) {
                Counters C = {};
                TraceShadow<false, typename StackSel<SK>::type> sh{a.S, C, stk};
                in_shadow = sh.query(0u, X, to_light, 2e-3f, dist - 2e-3f);
            }
            sf = in_shadow ? 0.32f : 1.0f;
            atten = 1.0f / (1.0f + 0.006f * dist * dist);
            lighting = m.emission + 0.30f * cv + 0.80f * sf * cl * atten;
        } else {
            lighting = 0.72f + 0.38f * cv + 0.22f * cl;
        }
        const V3 lit = alb * lighting;
        const float fog = expf_(-0.008f * lam);
        const V3 sky = sky_gradient(dir);
        const V3 color = lit * fog + sky * (1.0f - fog);

        // The albedo without its wireframe blend, and d albedo / d (smallest triangle barycentric).
        // The albedo without its wireframe blend (cell colour times grid-line factor), and whether
        // the forward is inside the blend band (where the albedo moves with the triangle
        // barycentrics).
        V3 nw = alb, col0 = V3{0.0f, 0.0f, 0.0f};
        float dark = 1.0f;
        bool wire_branch = false;
        int mtb_k = 0;
        if (m.pattern != SPECREL_PATTERN_CLOCK && m.pattern != SPECREL_PATTERN_LIGHT_CLOCK) {
            int32_t gu = f2i_sat(floorf(rec.uv_u * m.grid)), gv = f2i_sat(floorf(rec.uv_v * m.grid));
            int32_t sum = (int32_t)((uint32_t)gu + (uint32_t)gv);
            col0 = (((sum % 2) + 2) % 2) == 0 ? m.base : m.checker;
            if (!checker) {
                float ut = rem_euclid_(rec.uv_u * m.grid, 1.0f), vt = rem_euclid_(rec.uv_v * m.grid, 1.0f);
                float bd = fmin_(fmin_(ut, 1.0f - ut), fmin_(vt, 1.0f - vt));
                if (bd < 0.025f) dark = 0.35f;
            }
            nw = col0 * dark;
            V3 tb = rec.tri_bary;
            float mtb = fmin_(fmin_(tb.x, tb.y), tb.z);
            mtb_k = (tb.x <= tb.y) ? (tb.x <= tb.z ? 0 : 2) : (tb.y <= tb.z ? 1 : 2);
            if (mtb < m.wire) {
                float et = mtb / m.wire;
                wire_branch = et > 0.0f && et < 1.0f;
            }
        }
        obj = find_object(a.obj_first, a.obj_count, a.n_obj, a.S.order[slot]);
        uint32_t h = fnv(fnv(2166136261u, obj), rec.material);
        h = fnv(fnv(fnv(h, __float_as_uint(nw.x)), __float_as_uint(nw.y)), __float_as_uint(nw.z));
        h = fnv(h, in_shadow ? 1u : 0u);
        // the facet: normal and four-velocity are constant per tetrahedron (flat shading)
        h = fnv(fnv(fnv(fnv(h, __float_as_uint(nrm.t)), __float_as_uint(nrm.x)), __float_as_uint(nrm.y)),
                __float_as_uint(nrm.z));
        h = fnv(fnv(fnv(fnv(h, __float_as_uint(u4.t)), __float_as_uint(u4.x)), __float_as_uint(u4.y)),
                __float_as_uint(u4.z));
        if (h >= SIG_STAR) h &= 0x7FFFFFFFu;
        const bool same = __float_as_uint(color.x) == __float_as_uint(hdr.x) &&
                          __float_as_uint(color.y) == __float_as_uint(hdr.y) &&
                          __float_as_uint(color.z) == __float_as_uint(hdr.z);
        sig = same ? h : SIG_STAR;  // a differing pixel has starlight added (star.rs:237-286)
        if (same) {
            // Continuous intermediates again, in double, from the same f32 inputs.
            const D4 e0 = d4(lv(f.e[0])), e1 = d4(lv(f.e[1])), e2 = d4(lv(f.e[2])), e3 = d4(lv(f.e[3]));
            const float sxf = pix_sx(v, px), syf = pix_sy(v, py);
            const V3 nl = normalized(normalized(V3{sxf, syf, 1.0f}));
            const D4 o64 = d4(o);
            const D4 d = (e0 * -1.0) + e1 * (double)nl.x + e2 * (double)nl.y + e3 * (double)nl.z;
            const D4 v0 = d4(g[0]), N = d4(g[1]), B1 = d4(g[2]), B2 = d4(g[3]), B3 = d4(g[4]);
            const D4 nr = d4(nrm), uu = d4(u4);
            const double den = edot(N, d), lam64 = edot(N, v0 - o64) / den;
            const D4 X64 = o64 + d * lam64, q = X64 - v0;
            const double bu = edot(B1, q), bv = edot(B2, q), bw = edot(B3, q);
            const double bb[4] = {1.0 - bu - bv - bw, bu, bv, bw};
            double tb64[3];
            for (int k = 0; k < 3; ++k) {
                const float4 c = A[2 + k];
                tb64[k] = bb[0] * c.x + bb[1] * c.y + bb[2] * c.z + bb[3] * c.w;
            }
            double alb64[3] = {alb.x, alb.y, alb.z}, dadm[3] = {0.0, 0.0, 0.0};
            if (wire_branch) {
                const double et = tb64[mtb_k] / (double)m.wire;
                const double c0[3] = {col0.x, col0.y, col0.z};
                for (int k = 0; k < 3; ++k) {
                    alb64[k] = (0.07 * (1.0 - et) + c0[k] * et) * (double)dark;
                    dadm[k] = (c0[k] - 0.07) * (double)dark / (double)m.wire;
                }
            } else if (m.pattern != SPECREL_PATTERN_CLOCK && m.pattern != SPECREL_PATTERN_LIGHT_CLOCK) {
                alb64[0] = (double)col0.x * dark;
                alb64[1] = (double)col0.y * dark;
                alb64[2] = (double)col0.z * dark;
            }
            const double Lx = a.S.light.y, Ly = a.S.light.z, Lz = a.S.light.w;
            const double sx = Lx - X64.x, sy = Ly - X64.y, sz = Lz - X64.z;
            const double len64 = sqrt(sx * sx + sy * sy + sz * sz);
            const bool dist_free = len > 1e-4f;
            const double dist64 = dist_free ? len64 : 1e-4;
            const double inv64 = 1.0 / dist64;
            const D4 delta64 = D4{(X64.t - dist64) - X64.t, sx, sy, sz};
            const D4 tl64 = delta64 * inv64;
            const bool rl_free = rfc_free(to_light, nrm, u4), rv_free = rfc_free(dir, nrm, u4);
            const bool cl_on = rl > 0.0f, cv_on = -rv > 0.0f;
            const double rl64 = rl_free ? rest_frame_cos_d(tl64, nr, uu) : (double)rl;
            const double rv64 = rv_free ? rest_frame_cos_d(d, nr, uu) : (double)rv;
            const double cl64 = cl_on ? rl64 : 0.0, cv64 = cv_on ? -rv64 : 0.0;
            double lighting64, atten64 = 0.0;
            if (checker) {
                atten64 = 1.0 / (1.0 + 0.006 * dist64 * dist64);
                lighting64 = (double)m.emission + 0.30 * cv64 + 0.80 * (double)sf * cl64 * atten64;
            } else {
                lighting64 = 0.72 + 0.38 * cv64 + 0.22 * cl64;
            }
            const double fog64 = exp(-0.008 * lam64);
            const double sl64 = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
            double hz64 = d.y / sl64 * 0.5 + 0.5;
            bool hz_free = false;
            {
                V3 sp = spatial(dir);
                float sl = length(sp);
                if (sl > 1e-12f) {
                    float hz = sp.y * (1.0f / sl) * 0.5f + 0.5f;
                    hz_free = hz > 0.0f && hz < 1.0f;
                    if (!hz_free) hz64 = hz <= 0.0f ? 0.0 : 1.0;
                }
            }
            const double top[3] = {0.12, 0.16, 0.25}, bot[3] = {0.04, 0.06, 0.13};
            const double tg[3] = {tgt.x, tgt.y, tgt.z};
            double gc[3], lit64[3], sky64[3];
            for (int k = 0; k < 3; ++k) {
                lit64[k] = alb64[k] * lighting64;
                sky64[k] = top[k] * (1.0 - hz64) + bot[k] * hz64;
                gc[k] = 2.0 * (lit64[k] * fog64 + sky64[k] * (1.0 - fog64) - tg[k]);
            }
            // color = lit * fog + sky * (1 - fog)
            double dfog = 0.0, dhz = 0.0, dl = 0.0, dmtb = 0.0;
            for (int k = 0; k < 3; ++k) {
                const double dlit = gc[k] * fog64, dsky = gc[k] * (1.0 - fog64);
                dfog += gc[k] * (lit64[k] - sky64[k]);
                dhz += dsky * (bot[k] - top[k]);
                dl += dlit * alb64[k];
                dmtb += dlit * lighting64 * dadm[k];
            }
            double dlam = dfog * fog64 * -0.008;
            D4 dd = D4{0.0, 0.0, 0.0, 0.0}, dX = D4{0.0, 0.0, 0.0, 0.0};
            if (hz_free) {  // hz = d.y / |d.xyz| * 0.5 + 0.5
                const double sxn = d.x / sl64, syn = d.y / sl64, szn = d.z / sl64, dsy = 0.5 * dhz / sl64;
                dd.x += -sxn * syn * dsy;
                dd.y += (1.0 - syn * syn) * dsy;
                dd.z += -szn * syn * dsy;
            }
            double dcv, dcl, datten = 0.0;
            if (checker) {
                dcv = 0.30 * dl;
                dcl = 0.80 * (double)sf * atten64 * dl;
                datten = 0.80 * (double)sf * cl64 * dl;
            } else {
                dcv = 0.38 * dl;
                dcl = 0.22 * dl;
            }
            double ddist = datten * (-atten64 * atten64 * 0.012 * dist64);
            D4 dtl = D4{0.0, 0.0, 0.0, 0.0}, dn = D4{0.0, 0.0, 0.0, 0.0}, du = D4{0.0, 0.0, 0.0, 0.0};
            if (rl_free && cl_on) rest_frame_cos_grad(tl64, nr, uu, dcl, dtl, dn, du);
            if (rv_free && cv_on) rest_frame_cos_grad(d, nr, uu, -dcv, dd, dn, du);
            // to_light = delta / dist, delta = (-dist, s2l) as a function of the event
            const double dinv = edot(dtl, delta64);
            const D4 ddelta = dtl * inv64;
            if (dist_free) {
                ddist += -dinv * inv64 * inv64 - ddelta.t;
                const double k = ddist / len64;
                dX.x -= ddelta.x + sx * k;
                dX.y -= ddelta.y + sy * k;
                dX.z -= ddelta.z + sz * k;
            } else {
                dX.x -= ddelta.x;
                dX.y -= ddelta.y;
                dX.z -= ddelta.z;
            }
            // barycentrics: b = (1 - u - v - w, u, v, w), tri_bary_k = sum_i A[2 + k][i] b_i
            double db[4];
            {
                const float4 c = A[2 + mtb_k];
                db[0] = dmtb * c.x;
                db[1] = dmtb * c.y;
                db[2] = dmtb * c.z;
                db[3] = dmtb * c.w;
            }
            const double du_ = db[1] - db[0], dv_ = db[2] - db[0], dw_ = db[3] - db[0];
            const D4 dq = B1 * du_ + B2 * dv_ + B3 * dw_;
            const D4 dB1 = q * du_, dB2 = q * dv_, dB3 = q * dw_;
            dX = dX + dq;
            D4 dv0 = dq * -1.0;
            // X = o + d * lambda
            D4 dorig = dX;
            dd = dd + dX * lam64;
            dlam += edot(dX, d);
            // lambda = edot(N, v0 - o) / edot(N, d)
            const double dnum = dlam / den, dden = -dlam * lam64 / den;
            const D4 dN = (v0 - o64) * dnum + d * dden;
            dv0 = dv0 + N * dnum;
            dorig = dorig - N * dnum;
            dd = dd + N * dden;
            out[0] = dorig.t, out[1] = dorig.x, out[2] = dorig.y, out[3] = dorig.z;
            out[4] = dd.t, out[5] = dd.x, out[6] = dd.y, out[7] = dd.z;
            if (obj != NO_OBJ) {
                // The object's worldtube boosted about c: v0 -> c + L (v0 - c), N, B_i -> L^-T (.),
                // normal and four-velocity -> L (.), with dL/d rapidity_a = K_a at zero.
                const D4 rel = v0 - d4(a.obj_center[obj]);
                for (int ax = 1; ax <= 3; ++ax)
                    out[7 + ax] = edot(dv0, boost_gen(ax, rel)) - edot(dN, boost_gen(ax, N)) -
                                  edot(dB1, boost_gen(ax, B1)) - edot(dB2, boost_gen(ax, B2)) -
                                  edot(dB3, boost_gen(ax, B3)) + edot(dn, boost_gen(ax, nr)) +
                                  edot(du, boost_gen(ax, uu));
            }
        }
    }
    double *p = a.pix + ADJ_PIX * li;
    for (int k = 0; k < ADJ_PIX; ++k) p[k] = out[k];
    a.sig[li] = sig;
    a.obj[li] = obj;
    a.inside[li] = inside ? 1 : 0;
}

template __global__ void k_adj_pixel<16>(AdjArgs);

// Pass 2: one block per 16x16 tile of every view. A pixel is valid when it hit inside its
// tetrahedron, carries no starlight and its 8 neighbours lie in the image on the same object and
// carry no starlight (silhouettes are masked); with `strict` its neighbours must also have its
// shading signature (no facet, texture or shadow edge nearby: for finite differences); or, with a
// caller's mask, when the mask says so. Valid pixels' terms are summed in double in a fixed order
// per block.
__global__ __launch_bounds__(256) void k_adj_reduce(AdjArgs a) {
    const uint32_t t = blockIdx.x, vi = find_view(a.V.level_start[0], a.V.n, t);
    const specrel_frame &f = a.V.frames[vi];
    const ViewDerived v = derive(f, 0);
    const uint32_t lt = t - a.V.level_start[0][vi], tiles_x = (v.w + TILE - 1) / TILE, tid = threadIdx.x;
    const uint32_t px = (lt % tiles_x) * TILE + (tid & 15), py = (lt / tiles_x) * TILE + (tid >> 4);
    const size_t base = a.V.pix_start[vi];
    __shared__ double s_red[256];
    __shared__ uint32_t s_obj[256];
    __shared__ double s_g[256 * 3];
    __shared__ uint32_t s_cnt[ADJ_CLASSES];
    if (tid < ADJ_CLASSES) s_cnt[tid] = 0u;
    double val[ADJ_CAM];
    for (int k = 0; k < ADJ_CAM; ++k) val[k] = 0.0;
    uint32_t my_obj = NO_OBJ;
    double go[3] = {0.0, 0.0, 0.0};
    __syncthreads();
    if (px < v.w && py < v.h) {
        const size_t li = base + (size_t)py * v.w + px;
        const uint32_t sig = a.sig[li];
        int cls;  // 0 valid, 1 miss, 2 starlight, 3 image border, 4 silhouette, 5 edge (of the tet; strict: shading)
        if (a.mask_in) {
            cls = a.mask_in[li] ? 0 : (sig == SIG_MISS ? 1 : (sig == SIG_STAR ? 2 : 5));
        } else if (sig == SIG_MISS) {
            cls = 1;
        } else if (sig == SIG_STAR) {
            cls = 2;
        } else if (px == 0 || py == 0 || px + 1 == v.w || py + 1 == v.h) {
            cls = 3;
        } else {
            cls = 0;
            const uint32_t mine = a.obj[li];
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const size_t ni = base + (size_t)(py + dy) * v.w + (px + dx);
                    const uint32_t s2 = a.sig[ni];
                    if (s2 == sig) continue;
                    if (s2 == SIG_MISS || s2 == SIG_STAR || a.obj[ni] != mine) cls = 4;
                    else if (a.strict && cls == 0) cls = 5;
                }
            if (cls == 0 && !a.inside[li]) cls = 5;
        }
        atomicAdd(&s_cnt[cls], 1u);
        if (a.mask_out) a.mask_out[li] = cls == 0 ? 1 : 0;
        const double *p = a.pix + ADJ_PIX * li;
        if (cls == 0 || (a.mask_in && a.mask_in[li])) {
            // d/d event = d/d origin; d = -e0 + n.x e1 + n.y e2 + n.z e3 (past light cone)
            float sx = pix_sx(v, px), sy = pix_sy(v, py);
            V3 n = normalized(normalized(V3{sx, sy, 1.0f}));
            val[0] = p[11];
            for (int k = 0; k < 4; ++k) {
                val[1 + k] = p[k];
                val[5 + k] = -p[4 + k];
                val[9 + k] = p[4 + k] * n.x;
                val[13 + k] = p[4 + k] * n.y;
                val[17 + k] = p[4 + k] * n.z;
            }
            my_obj = a.obj[li];
            go[0] = p[8];
            go[1] = p[9];
            go[2] = p[10];
        }
    }
    double *out = a.partials + (size_t)t * a.stride;
    for (int k = 0; k < ADJ_CAM; ++k) {
        s_red[tid] = val[k];
        __syncthreads();
        for (uint32_t st = 128; st > 0; st >>= 1) {
            if (tid < st) s_red[tid] += s_red[tid + st];
            __syncthreads();
        }
        if (tid == 0) out[k] = s_red[0];
        __syncthreads();
    }
    s_obj[tid] = my_obj;
    s_g[3 * tid] = go[0];
    s_g[3 * tid + 1] = go[1];
    s_g[3 * tid + 2] = go[2];
    __syncthreads();
    for (uint32_t k = tid; k < a.n_obj; k += 256) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (uint32_t i = 0; i < 256; ++i)
            if (s_obj[i] == k) {
                acc[0] += s_g[3 * i];
                acc[1] += s_g[3 * i + 1];
                acc[2] += s_g[3 * i + 2];
            }
        out[ADJ_CAM + 3 * k] = acc[0];
        out[ADJ_CAM + 3 * k + 1] = acc[1];
        out[ADJ_CAM + 3 * k + 2] = acc[2];
    }
    if (tid < ADJ_CLASSES) a.block_counts[(size_t)t * ADJ_CLASSES + tid] = s_cnt[tid];
}

// Pass 3: each view's tiles added up: a block takes 16 components of one view, 16 threads per
// component stride over the tiles in order, and their 16 partial sums are added in order
// (the result does not depend on scheduling). Components past `stride` are the pixel classes.
__global__ __launch_bounds__(256) void k_adj_final(AdjArgs a) {
    const uint32_t v = blockIdx.x, tid = threadIdx.x, c = blockIdx.y * 16 + (tid & 15), lane = tid >> 4;
    const uint32_t b0 = a.V.level_start[0][v], b1 = a.V.level_start[0][v + 1];
    __shared__ double s_part[256];
    double acc = 0.0;
    if (c < a.stride) {
        for (uint32_t b = b0 + lane; b < b1; b += 16) acc += a.partials[(size_t)b * a.stride + c];
    } else if (c < a.stride + ADJ_CLASSES) {
        uint64_t n = 0;
        for (uint32_t b = b0 + lane; b < b1; b += 16) n += a.block_counts[(size_t)b * ADJ_CLASSES + (c - a.stride)];
        acc = (double)n;  // exact: pixel counts stay far below 2^53
    }
    s_part[tid] = acc;
    __syncthreads();
    if (lane == 0) {
        double sum = 0.0;
        for (uint32_t j = 0; j < 16; ++j) sum += s_part[(tid & 15) + 16 * j];
        if (c < a.stride) a.result[(size_t)v * a.stride + c] = sum;
        else if (c < a.stride + ADJ_CLASSES)
            a.result_counts[(size_t)v * ADJ_CLASSES + (c - a.stride)] = (unsigned long long)sum;
    }
}

}  // namespace dev
