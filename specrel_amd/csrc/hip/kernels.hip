// Kernels of the gfx950 renderer: per-view star pass, the fused trace/shade kernel, and the
// bloom pyramid + tone map (bloom.rs:93-125, renderer.rs:1116-1129).
#include "kernels.hpp"

namespace dev {

// Camera-derived per-view values, computed exactly as render_frame does (renderer.rs:334-358).
struct ViewDerived {
    V4 origin, e0, e1, e2, e3;
    float cone, half_w, half_h, pitch;
    uint32_t w, h;
    bool draw_stars;
};

__device__ __forceinline__ ViewDerived derive(const specrel_frame &f, uint32_t num_stars) {
    ViewDerived v;
    v.e0 = lv(f.e[0]);
    v.e1 = lv(f.e[1]);
    v.e2 = lv(f.e[2]);
    v.e3 = lv(f.e[3]);
    V4 cam = lv(f.cam_pos);
    if (f.slice_mode == SPECREL_VIEW_FRONTIER) v.origin = cam - v.e0 * f.frontier_view_offset;
    else if (f.slice_mode == SPECREL_VIEW_PROPHECY) v.origin = cam + v.e0 * f.prophecy_delta;
    else v.origin = cam;
    v.cone = f.slice_mode == SPECREL_VIEW_PAST_CONE ? -1.0f : (f.slice_mode == SPECREL_VIEW_FRONTIER ? 1.0f : 0.0f);
    v.w = f.size[0];
    v.h = f.size[1];
    v.half_w = f.half_wh[0];
    v.half_h = f.half_wh[1];
    v.pitch = 2.0f * v.half_h / (float)v.h;
    v.draw_stars = f.mode == SPECREL_MODE_SHADED && num_stars > 0;
    return v;
}

__device__ __forceinline__ float pix_sx(const ViewDerived &v, uint32_t px) {
    float ndc_x = 2.0f * (((float)px + 0.5f) / (float)v.w) - 1.0f;
    return ndc_x * v.half_w;
}
__device__ __forceinline__ float pix_sy(const ViewDerived &v, uint32_t py) {
    float ndc_y = 1.0f - 2.0f * (((float)py + 0.5f) / (float)v.h);
    return ndc_y * v.half_h;
}

// Star pass: for every star, its cone event and the view-only parts of star_light
// (star.rs:251-283), plus its direction in the camera rest frame for tile culling.
__global__ __launch_bounds__(256) void k_star_prep(const float4 *words, uint32_t num_stars,
                                                   const specrel_frame *frame, StarPrep *out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_stars) return;
    const specrel_frame &f = *frame;
    ViewDerived v = derive(f, num_stars);
    StarPrep p;
    p.flags = 0;
    p.r = make_float4(0, 0, 0, 0);
    p.bb_cn = make_float4(0, 0, 0, 0);
    p.dir = make_double4(0, 0, 0, 0);
    p.pad[0] = p.pad[1] = p.pad[2] = 0;
    V4 d, u;
    if (cone_event(words, num_stars, i, v.origin, v.e0, v.cone, d, u)) {
        float tau = -mdot(d, v.e0);
        V4 r = d + v.e0 * (-tau);
        float doppler = 1.0f;
        if (v.cone < 0.0f) {
            float lam = sqrtf(fmax_(mdot(r, r), 0.0f));
            float du = mdot(d, u);
            doppler = du > 0.0f ? lam / du : 1.0f;
        }
        float4 hdr = words[i];
        V3 bb = blackbody(hdr.x * doppler);
        p.r = make_float4(r.t, r.x, r.y, r.z);
        p.bb_cn = make_float4(bb.x, bb.y, bb.z, hdr.y);
        p.flags = 1u;
        // Direction of r in the camera frame (r is orthogonal to e0 up to rounding).
        double rt = r.t, rx = r.x, ry = r.y, rz = r.z;
        double lx = -rt * (double)v.e1.t + rx * (double)v.e1.x + ry * (double)v.e1.y + rz * (double)v.e1.z;
        double ly = -rt * (double)v.e2.t + rx * (double)v.e2.x + ry * (double)v.e2.y + rz * (double)v.e2.z;
        double lz = -rt * (double)v.e3.t + rx * (double)v.e3.x + ry * (double)v.e3.y + rz * (double)v.e3.z;
        double ln = sqrt(lx * lx + ly * ly + lz * lz);
        bool finite = isfinite(r.t) && isfinite(r.x) && isfinite(r.y) && isfinite(r.z) && isfinite(bb.x) &&
                      isfinite(bb.y) && isfinite(bb.z) && isfinite(hdr.y) && ln > 0.0;
        if (!finite) p.flags |= 2u;  // evaluated for every pixel, never culled
        else p.dir = make_double4(lx / ln, ly / ln, lz / ln, 0.0);
    }
    out[i] = p;
}

constexpr int TILE = 16;
constexpr int STAR_CAP = 2048;

template <bool STATS>
__global__ __launch_bounds__(256) void k_render(RenderArgs a) {
    const specrel_frame &f = *a.frame;
    const ViewDerived v = derive(f, a.num_stars);
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const uint32_t x0 = blockIdx.x * TILE, y0 = blockIdx.y * TILE;
    const uint32_t px = x0 + (wave & 1) * 8 + (lane & 7);
    const uint32_t py = y0 + (wave >> 1) * 8 + (lane >> 3);

    __shared__ uint32_t s_list[STAR_CAP];
    __shared__ uint32_t s_wave[4];
    __shared__ uint32_t s_total;
    Counters C = {};

    const bool cull = v.draw_stars && a.cull;
    if (cull) {
        // Bounding cone of the tile's pixel directions (pixel centres) in the camera frame.
        uint32_t xa = x0, xb = min(x0 + TILE, v.w) - 1, ya = y0, yb = min(y0 + TILE, v.h) - 1;
        double cd[4][3];
        uint32_t cxs[4] = {xa, xb, xa, xb}, cys[4] = {ya, ya, yb, yb};
        double cx = 0, cy = 0, cz = 0;
        for (int k = 0; k < 4; ++k) {
            double sx = pix_sx(v, cxs[k]), sy = pix_sy(v, cys[k]);
            double n = sqrt(sx * sx + sy * sy + 1.0);
            cd[k][0] = sx / n;
            cd[k][1] = sy / n;
            cd[k][2] = 1.0 / n;
            cx += cd[k][0];
            cy += cd[k][1];
            cz += cd[k][2];
        }
        double cn = sqrt(cx * cx + cy * cy + cz * cz);
        cx /= cn;
        cy /= cn;
        cz /= cn;
        double cos_a = 1.0;
        for (int k = 0; k < 4; ++k) cos_a = fmin(cos_a, cd[k][0] * cx + cd[k][1] * cy + cd[k][2] * cz);
        // A star reaches a pixel only if tan(angle to the pixel ray) < 2.4 spread <= 2.4 pitch
        // (sigma = 0.6 spread lambda, cut = 4 sigma); 5% and a quarter pixel of slack on top.
        double beta = (1.05 * 2.4 + 0.25) * (double)v.pitch;
        double ang = acos(fmin(1.0, cos_a)) + beta;
        double cos_lim = ang >= 3.141592653589793 ? -2.0 : cos(ang);
        if (tid == 0) s_total = 0;
        __syncthreads();
        for (uint32_t base = 0; base < a.num_stars; base += 256) {
            uint32_t idx = base + tid;
            bool keep = false;
            if (idx < a.num_stars) {
                const StarPrep &p = a.stars[idx];
                if (p.flags & 2u) keep = true;
                else if (p.flags & 1u) keep = p.dir.x * cx + p.dir.y * cy + p.dir.z * cz >= cos_lim - 1e-12;
            }
            uint64_t mask = __ballot(keep);
            uint32_t before = __popcll(mask & ((1ull << lane) - 1ull));
            if (lane == 0) s_wave[wave] = __popcll(mask);
            __syncthreads();
            uint32_t off = s_total;
            for (uint32_t k = 0; k < wave; ++k) off += s_wave[k];
            if (keep && off + before < STAR_CAP) s_list[off + before] = idx;
            __syncthreads();
            if (tid == 0) s_total += s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
            __syncthreads();
        }
    }

    if (px < v.w && py < v.h) {
        float sx = pix_sx(v, px), sy = pix_sy(v, py);
        V3 local_dir = normalized(V3{sx, sy, 1.0f});
        V4 dir;
        if (f.slice_mode == SPECREL_VIEW_PAST_CONE) {
            V3 n = normalized(local_dir);
            dir = -v.e0 + v.e1 * n.x + v.e2 * n.y + v.e3 * n.z;
        } else if (f.slice_mode == SPECREL_VIEW_FRONTIER) {
            V3 n = normalized(local_dir);
            dir = v.e0 + v.e1 * n.x + v.e2 * n.y + v.e3 * n.z;
        } else {
            dir = v.e1 * local_dir.x + v.e2 * local_dir.y + v.e3 * local_dir.z;
        }
        Hit hit;
        bool found = !a.S.empty && trace<STATS>(a.S, v.origin, dir, 1e-3f, 180.0f, hit, C);
        V3 color;
        Rec rec;
        if (found) {
            interpolate(a.S.attrs + 12 * (size_t)hit.slot, hit, rec);
            color = shade_hit<STATS>(a.S, a.T, f, dir, rec, C);
            if (f.num_haze > 0) color = haze(color, rec.event, f);
        } else {
            color = shade_sky(dir);
        }
        if (v.draw_stars) {
            float inv_len = 1.0f / sqrtf(sx * sx + sy * sy + 1.0f);
            float spread = v.pitch * inv_len * sqrtf(inv_len);
            float lmax = found ? hit.lambda : __builtin_inff();
            V4 n = dir - v.e0 * v.cone;
            V3 total = V3{0.0f, 0.0f, 0.0f};
            if (cull && s_total <= (uint32_t)STAR_CAP) {
                for (uint32_t k = 0; k < s_total; ++k) star_pixel<STATS>(a.stars[s_list[k]], n, spread, lmax, total, C);
            } else {
                for (uint32_t k = 0; k < a.num_stars; ++k) star_pixel<STATS>(a.stars[k], n, spread, lmax, total, C);
            }
            color += total;
        }
        size_t pi = (size_t)py * v.w + px;
        a.hdr[pi] = make_float4(color.x, color.y, color.z, 0.0f);
        if (a.want_hits) {
            a.hit[pi] = found ? a.S.order[hit.slot] : SPECREL_NO_HIT;
            a.lambda[pi] = found ? hit.lambda : __builtin_inff();
            a.tau[pi] = found ? rec.tau : 0.0f;
        }
    }

    if (STATS) {
        __shared__ unsigned long long s_c[10];
        if (tid < 10) s_c[tid] = 0;
        __syncthreads();
        bool in = px < v.w && py < v.h;
        atomicAdd(&s_c[0], (unsigned long long)(in ? 1 : 0));
        atomicAdd(&s_c[1], (unsigned long long)C.node_pops);
        atomicAdd(&s_c[2], (unsigned long long)C.box_tests);
        atomicAdd(&s_c[3], (unsigned long long)C.prim_tests);
        atomicAdd(&s_c[5], (unsigned long long)C.shadow_rays);
        atomicAdd(&s_c[6], (unsigned long long)C.shadow_node_pops);
        atomicAdd(&s_c[7], (unsigned long long)C.shadow_box_tests);
        atomicAdd(&s_c[8], (unsigned long long)C.shadow_prim_tests);
        atomicAdd(&s_c[9], (unsigned long long)C.star_tests);
        if (in && a.want_hits && a.hit[(size_t)py * v.w + px] != SPECREL_NO_HIT) atomicAdd(&s_c[4], 1ull);
        __syncthreads();
        if (tid < 10) atomicAdd(&a.stats[tid], s_c[tid]);
    }
}

template __global__ void k_render<false>(RenderArgs);
template __global__ void k_render<true>(RenderArgs);

// ---------------------------------------------------------------------------------------
// Bloom (bloom.rs:22-125): the pyramid lives after the full-resolution HDR image.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ V3 ld3(const float4 *p, size_t i) {
    float4 a = p[i];
    return V3{a.x, a.y, a.z};
}

__global__ __launch_bounds__(256) void k_bloom_down(const float4 *src, uint32_t sw, uint32_t sh, float4 *dst,
                                                    uint32_t dw, uint32_t dh) {
    uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= dw || y >= dh) return;
    const float W[4] = {0.125f, 0.375f, 0.375f, 0.125f};
    V3 acc = V3{0.0f, 0.0f, 0.0f};
    for (int j = 0; j < 4; ++j) {
        int64_t syy = 2 * (int64_t)y - 1 + j;
        size_t sy = (size_t)(syy < 0 ? 0 : (syy > (int64_t)sh - 1 ? (int64_t)sh - 1 : syy));
        V3 row = V3{0.0f, 0.0f, 0.0f};
        for (int i = 0; i < 4; ++i) {
            int64_t sxx = 2 * (int64_t)x - 1 + i;
            size_t sx = (size_t)(sxx < 0 ? 0 : (sxx > (int64_t)sw - 1 ? (int64_t)sw - 1 : sxx));
            row += ld3(src, sy * sw + sx) * W[i];
        }
        acc += row * W[j];
    }
    dst[(size_t)y * dw + x] = make_float4(acc.x, acc.y, acc.z, 0.0f);
}

// bloom::upsample_at (bloom.rs:80-90)
__device__ __forceinline__ V3 upsample_at(const float4 *coarse, uint32_t cw, uint32_t ch, uint32_t x, uint32_t y) {
    uint32_t mx = x / 2, my = y / 2;
    int64_t ox = (x % 2 == 0) ? (int64_t)mx - 1 : (int64_t)mx + 1;
    int64_t oy = (y % 2 == 0) ? (int64_t)my - 1 : (int64_t)my + 1;
    uint32_t x1 = (uint32_t)(ox < 0 ? 0 : (ox > (int64_t)cw - 1 ? (int64_t)cw - 1 : ox));
    uint32_t y1 = (uint32_t)(oy < 0 ? 0 : (oy > (int64_t)ch - 1 ? (int64_t)ch - 1 : oy));
    V3 r0 = ld3(coarse, (size_t)my * cw + mx) * 0.75f + ld3(coarse, (size_t)my * cw + x1) * 0.25f;
    V3 r1 = ld3(coarse, (size_t)y1 * cw + mx) * 0.75f + ld3(coarse, (size_t)y1 * cw + x1) * 0.25f;
    return r0 * 0.75f + r1 * 0.25f;
}

__global__ __launch_bounds__(256) void k_bloom_up(float4 *fine, uint32_t w, uint32_t h, const float4 *coarse,
                                                  uint32_t cw, uint32_t ch) {
    uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= w || y >= h) return;
    size_t i = (size_t)y * w + x;
    V3 c = ld3(fine, i) + upsample_at(coarse, cw, ch, x, y);
    fine[i] = make_float4(c.x, c.y, c.z, 0.0f);
}

// Final mix (bloom.rs:116-124) and pack_rgb_with_lut (renderer.rs:1116-1129).
__global__ __launch_bounds__(256) void k_composite(const float4 *hdr, uint32_t w, uint32_t h, const float4 *l1,
                                                   uint32_t w1, uint32_t h1, const uint8_t *lut, uint32_t *out) {
    uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= w || y >= h) return;
    size_t i = (size_t)y * w + x;
    float halo = 0.04f / 6.0f;
    V3 c = ld3(hdr, i) * (1.0f - 0.04f) + upsample_at(l1, w1, h1, x, y) * halo;
    V3 m = clamp3(V3{c.x / (1.0f + c.x * 0.35f), c.y / (1.0f + c.y * 0.35f), c.z / (1.0f + c.z * 0.35f)}, 0.0f, 1.0f);
    const float scale = 4095.0f;
    uint32_t r = lut[f2u_sat(m.x * scale)], g = lut[f2u_sat(m.y * scale)], b = lut[f2u_sat(m.z * scale)];
    out[i] = (r << 16) | (g << 8) | b;
}

}  // namespace dev
