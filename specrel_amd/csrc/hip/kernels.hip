// Kernels of the gfx950 renderer: per-view star pass and star binning, the primary-ray trace,
// the shading pass, and the bloom pyramid + tone map (bloom.rs:93-125, renderer.rs:1116-1129).
#include "kernels.hpp"

namespace dev {

// Star pass: for every star, its cone event and the view-only parts of star_light
// (star.rs:251-283), plus its unit direction in the camera rest frame for the tile culling.
// Star i of view `view`, whose frame is f.
__device__ __forceinline__ void star_prep_one(const float4 *words, uint32_t num_stars, const specrel_frame &f,
                                              uint32_t view, uint32_t i, StarPrep *out, StarDir *dirs) {
    if (i >= num_stars || f.mode != SPECREL_MODE_SHADED) return;
    ViewDerived v = derive(f, num_stars);
    out += (size_t)view * num_stars;
    dirs += (size_t)view * num_stars;
    StarPrep p;
    p.flags = 0;
    p.r = make_float4(0, 0, 0, 0);
    p.bb_cn = make_float4(0, 0, 0, 0);
    p.pad[0] = p.pad[1] = p.pad[2] = 0;
    StarDir sd = {0.0f, 0.0f, 0.0f, 0u};
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
        else sd = StarDir{(float)(lx / ln), (float)(ly / ln), (float)(lz / ln), 0u};
        sd.flags = p.flags;
    }
    out[i] = p;
    dirs[i] = sd;
}
// Grid: (star blocks, views).
__global__ __launch_bounds__(256) void k_star_prep(const float4 *words, uint32_t num_stars, Views V, StarPrep *out,
                                                   StarDir *dirs) {
    star_prep_one(words, num_stars, V.frames[blockIdx.y], blockIdx.y, blockIdx.x * blockDim.x + threadIdx.x, out,
                  dirs);
}
// The same for the one view of a render launched directly whose star pass comes first: its block 0
// also does k_reset_frame's work (the counters cleared, the frame, passed by value, written to the
// frame table for the later kernels), and every block reads the frame from the arguments. One
// launch instead of two.
__global__ __launch_bounds__(256) void k_star_prep_first(const float4 *words, uint32_t num_stars, StarPrep *out,
                                                         StarDir *dirs, uint32_t *counters, uint32_t n, FrameWords fw,
                                                         uint32_t *dst) {
    if (blockIdx.x == 0) {
        const uint32_t i = threadIdx.x;
        if (i < n) counters[i] = 0u;
        if (i < FrameWords::N) dst[i] = fw.w[i];
    }
    star_prep_one(words, num_stars, *reinterpret_cast<const specrel_frame *>(fw.w), 0u,
                  blockIdx.x * blockDim.x + threadIdx.x, out, dirs);
}

// Bounding cone of the pixel rays (pixel centres) of the image rectangle [xa,xb] x [ya,yb], widened
// by the largest angle at which a star can still reach a pixel: tan(angle) < 2.4 spread <= 2.4 pitch
// (sigma = 0.6 spread lambda, cut at 4 sigma; star.rs:261-267), plus 5% and half a pixel of slack.
// Returned as a float axis and a float cosine bound lowered by 1e-6, so a float dot product
// (error < 3e-7) can only keep more stars, never fewer.
__device__ void tile_cone(const ViewDerived &v, uint32_t xa, uint32_t xb, uint32_t ya, uint32_t yb, float &cx,
                          float &cy, float &cz, float &cos_lim) {
    double cd[4][3];
    uint32_t xs[4] = {xa, xb, xa, xb}, ys[4] = {ya, ya, yb, yb};
    double sx_ = 0, sy_ = 0, sz_ = 0;
    for (int k = 0; k < 4; ++k) {
        double sx = pix_sx(v, xs[k]), sy = pix_sy(v, ys[k]);
        double n = sqrt(sx * sx + sy * sy + 1.0);
        cd[k][0] = sx / n;
        cd[k][1] = sy / n;
        cd[k][2] = 1.0 / n;
        sx_ += cd[k][0];
        sy_ += cd[k][1];
        sz_ += cd[k][2];
    }
    double cn = sqrt(sx_ * sx_ + sy_ * sy_ + sz_ * sz_);
    sx_ /= cn;
    sy_ /= cn;
    sz_ /= cn;
    double cos_a = 1.0;
    for (int k = 0; k < 4; ++k) cos_a = fmin(cos_a, cd[k][0] * sx_ + cd[k][1] * sy_ + cd[k][2] * sz_);
    double beta = (1.05 * 2.4 + 0.5) * (double)v.pitch;
    double ang = acos(fmin(1.0, cos_a)) + beta;
    cx = (float)sx_;
    cy = (float)sy_;
    cz = (float)sz_;
    cos_lim = ang >= 3.0 ? -2.0f : (float)(cos(ang) - 1e-6);
}

__device__ __forceinline__ bool star_in_cone(const StarDir &s, float cx, float cy, float cz, float cos_lim) {
    if (s.flags & 2u) return true;
    if (!(s.flags & 1u)) return false;
    return s.x * cx + s.y * cy + s.z * cz >= cos_lim;
}

// Ordered compaction of candidates (in increasing index) into list by a block of W waves: every
// thread of the block calls it with the same `n`; keep(i) decides candidate i. Returns the total
// kept (which may exceed cap, in which case the list holds only the first cap entries).
template <int W, class Keep>
__device__ uint32_t ordered_compact(uint32_t n, Keep keep, uint32_t *list, uint32_t cap, uint32_t *s_wave,
                                    uint32_t *s_total) {
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (tid == 0) *s_total = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n; base += 64 * W) {
        uint32_t idx = base + tid;
        uint32_t item = 0;
        bool k = idx < n && keep(idx, item);
        uint64_t mask = __ballot(k);
        uint32_t before = __popcll(mask & ((1ull << lane) - 1ull));
        if (lane == 0) s_wave[wave] = __popcll(mask);
        __syncthreads();
        uint32_t off = *s_total;
        for (uint32_t w = 0; w < wave; ++w) off += s_wave[w];
        if (k && off + before < cap) list[off + before] = item;
        __syncthreads();
        if (tid == 0)
            for (int w = 0; w < W; ++w) *s_total += s_wave[w];
        __syncthreads();
    }
    return *s_total;
}

// The star list of a 16x16 tile (second culling level), in index order: culled from its super
// tile's list (or from every star when that list overflowed) by the tile's cone; when the tile's
// own list overflows, the super-tile list is used as is. Every thread of the block (W waves) calls
// it. The culling keeps every star the reference's per-pixel test can accept.
struct StarList {
    const uint32_t *src;  // super-tile list (nullptr: every star)
    uint32_t n_src, n_tile;
    bool use_tile;  // the tile's own list (in LDS) is complete
};
template <int W>
__device__ StarList tile_stars(const ShadeArgs &a, uint32_t vi, const ViewDerived &v, uint32_t t, uint32_t x0,
                               uint32_t y0, uint32_t *s_list, uint32_t *s_wave, uint32_t *s_total) {
    StarList L{nullptr, 0u, 0u, false};
    if (!(v.draw_stars && a.cull)) return L;
    const StarDir *dirs = a.dirs + (size_t)vi * a.num_stars;
    const uint32_t nsx = (v.w + SUPER - 1) / SUPER;
    const uint32_t st = a.V.super_start[vi] + (y0 / SUPER) * nsx + x0 / SUPER;
    const uint32_t sc = a.super_counts[st];
    const bool all = sc > SUPER_CAP;
    L.src = all ? nullptr : a.super_lists + (size_t)st * SUPER_CAP;
    L.n_src = all ? a.num_stars : sc;
    const float4 cone = a.tile_cones[t];  // tile_cone of this tile, from k_star_bin
    const float cx = cone.x, cy = cone.y, cz = cone.z, cl = cone.w;
    const uint32_t *src = L.src;
    L.n_tile = ordered_compact<W>(
        L.n_src,
        [&](uint32_t k, uint32_t &item) {
            item = src ? src[k] : k;
            return star_in_cone(dirs[item], cx, cy, cz, cl);
        },
        s_list, TILE_CAP, s_wave, s_total);
    L.use_tile = L.n_tile <= TILE_CAP;
    return L;
}

// One pixel of the shading pass (not the queueing pass): re-intersect the winning slot (the same
// operations as the traversal: the same lambda, event and barycentrics), interpolate, shade with
// the shadow answers of `sh`, haze, the sky for a miss, the stars of the tile's list; writes the
// linear HDR pixel and the requested hit index, lambda and tau.
template <bool STATS, class SH>
__device__ void shade_pixel(const ShadeArgs &a, uint32_t vi, const specrel_frame &f, const ViewDerived &v,
                            uint32_t px, uint32_t py, uint32_t slot, SH &sh, const StarList &L,
                            const uint32_t *s_list, Counters &C) {
    const StarPrep *stars = a.stars + (size_t)vi * a.num_stars;
    const size_t li = (size_t)py * v.w + px, pi = a.V.pix_start[vi] + li;
    V4 dir = ray_dir(v, f, px, py);
    const bool found = slot != SPECREL_NO_HIT;
    V3 color = V3{0.0f, 0.0f, 0.0f};
    Rec rec;
    float lambda = 0.0f;
    if (found) {
        Hit hit;
        hit.slot = (int32_t)slot;
        intersect(a.S.geoms + 5 * (size_t)slot, v.origin, dir, 1e-3f, 180.0f, hit.lambda, hit.event, hit.bary);
        lambda = hit.lambda;
        interpolate(a.S.attrs + 12 * (size_t)slot, hit, rec);
        if (STATS) touch(a.S.touch_hits, slot);
        color = shade_hit(a.S, a.T, f, dir, rec, sh);
        if (f.num_haze > 0) color = haze(color, rec.event, f);
    } else {
        color = shade_sky(dir);
    }
    if (v.draw_stars) {
        float sx = pix_sx(v, px), sy = pix_sy(v, py);
        float inv_len = 1.0f / sqrtf(sx * sx + sy * sy + 1.0f);
        float spread = v.pitch * inv_len * sqrtf(inv_len);
        float lmax = found ? lambda : __builtin_inff();
        V4 n = dir - v.e0 * v.cone;
        V3 total = V3{0.0f, 0.0f, 0.0f};
        if (a.cull && L.use_tile) {
            for (uint32_t k = 0; k < L.n_tile; ++k) star_pixel<STATS>(stars[s_list[k]], n, spread, lmax, total, C);
        } else if (a.cull && L.src) {
            for (uint32_t k = 0; k < L.n_src; ++k) star_pixel<STATS>(stars[L.src[k]], n, spread, lmax, total, C);
        } else {
            for (uint32_t k = 0; k < a.num_stars; ++k) star_pixel<STATS>(stars[k], n, spread, lmax, total, C);
        }
        color += total;
    }
    a.hdr[a.V.pyr_start[vi] + li] = make_float4(color.x, color.y, color.z, 0.0f);
    if (a.want_hits) {
        a.hit[pi] = found ? a.S.order[slot] : SPECREL_NO_HIT;
        a.lambda[pi] = found ? lambda : __builtin_inff();
        a.tau[pi] = found ? rec.tau : 0.0f;
    }
    if (STATS && found) C.hits++;
}

// The stars of a cone in index order (ordered_compact for the star test, K candidates per thread
// and round): in a round, thread t of the 256-thread block tests stars base + 256 j + t for
// j < K, their direction records loaded together (coalesced, the index clamped to the last star
// so every load is in bounds); one barrier per round shares the waves' counts per j, and the kept
// stars are placed j-major, wave by wave, lane by lane: increasing index, the same list as
// ordered_compact's in n / (256 K) rounds instead of n / 256.
template <int K>
__device__ uint32_t compact_stars(const StarDir *dirs, uint32_t n, float cx, float cy, float cz, float cl,
                                  uint32_t *list, uint32_t cap, uint32_t (*s_cnt)[4]) {
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t total = 0;
    for (uint32_t base = 0; base < n; base += 256 * K) {
        StarDir d[K];
#pragma unroll
        for (int j = 0; j < K; ++j) d[j] = dirs[min(base + 256 * j + tid, n - 1)];
        uint64_t mask[K];
        bool keep[K];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            keep[j] = base + 256 * j + tid < n && star_in_cone(d[j], cx, cy, cz, cl);
            mask[j] = __ballot(keep[j]);
            if (lane == 0) s_cnt[j][wave] = (uint32_t)__popcll(mask[j]);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < K; ++j) {
            uint32_t pos = total + (uint32_t)__popcll(mask[j] & below);
            for (uint32_t w = 0; w < wave; ++w) pos += s_cnt[j][w];
            if (keep[j] && pos < cap) list[pos] = base + 256 * j + tid;
            total += s_cnt[j][0] + s_cnt[j][1] + s_cnt[j][2] + s_cnt[j][3];
        }
        __syncthreads();
    }
    return total;
}

__device__ void tile_order_block(const uint32_t *cost, uint32_t n_tiles, uint32_t *order);

// First culling level: one block per 64x64-pixel super tile keeps, in index order, the stars that
// can reach any of its pixels. Its second wave meanwhile computes the cones of its 16x16 tiles for
// the shading pass's second level. With order_tiles, a leading block writes the next render's
// tile order (the sort of k_tile_order) beside the binning.
// Grid: every super tile of every view (+ 1 with order_tiles).
__global__ __launch_bounds__(256) void k_star_bin(Views V, const StarDir *dirs, uint32_t num_stars, uint32_t *lists,
                                                  uint32_t *counts, float4 *cones, const uint32_t *order_cost,
                                                  uint32_t order_tiles, uint32_t *order) {
    if (order_tiles && blockIdx.x == 0) {
        tile_order_block(order_cost, order_tiles, order);
        return;
    }
    const uint32_t gst = blockIdx.x - (order_tiles ? 1u : 0u), vi = find_view(V.super_start, V.n, gst);
    const specrel_frame &f = V.frames[vi];
    if (f.mode != SPECREL_MODE_SHADED) return;
    const ViewDerived v = derive(f, num_stars);
    dirs += (size_t)vi * num_stars;
    const uint32_t nsx = (v.w + SUPER - 1) / SUPER;
    const uint32_t st = gst - V.super_start[vi], x0 = (st % nsx) * SUPER, y0 = (st / nsx) * SUPER;
    __shared__ float s_cone[4];
    __shared__ uint32_t s_cnt[8][4];
    constexpr uint32_t PER_SIDE = SUPER / TILE;
    if (threadIdx.x == 0) {
        tile_cone(v, x0, min(x0 + SUPER, v.w) - 1, y0, min(y0 + SUPER, v.h) - 1, s_cone[0], s_cone[1], s_cone[2],
                  s_cone[3]);
    } else if (threadIdx.x >= 64 && threadIdx.x < 64 + PER_SIDE * PER_SIDE) {
        const uint32_t k = threadIdx.x - 64, tx0 = x0 + (k % PER_SIDE) * TILE, ty0 = y0 + (k / PER_SIDE) * TILE;
        if (tx0 < v.w && ty0 < v.h) {
            float c[4];
            tile_cone(v, tx0, min(tx0 + TILE, v.w) - 1, ty0, min(ty0 + TILE, v.h) - 1, c[0], c[1], c[2], c[3]);
            const uint32_t tiles_x = (v.w + TILE - 1) / TILE;
            cones[V.level_start[0][vi] + (ty0 / TILE) * tiles_x + tx0 / TILE] = make_float4(c[0], c[1], c[2], c[3]);
        }
    }
    __syncthreads();
    const float cx = s_cone[0], cy = s_cone[1], cz = s_cone[2], cl = s_cone[3];
    uint32_t *list = lists + (size_t)gst * SUPER_CAP;
    uint32_t n = num_stars ? compact_stars<8>(dirs, num_stars, cx, cy, cz, cl, list, SUPER_CAP, s_cnt) : 0u;
    if (threadIdx.x == 0) counts[gst] = n;
}

// Primary rays: closest hit per pixel (the leaf slot, or SPECREL_NO_HIT). With SH >= 1, a shaded
// view's hit also traces its shadow queries at once (the ones compute_illumination and
// compute_flash_illumination make, renderer.rs:585-688, from the same record the shading pass
// rebuilds) and stores their answers for the shading pass (k_shade PASS 2). With SH == 2 the
// block then shades its own pixels (the shading pass fused in: the tile's star list culled into
// the LDS the traversal stacks used, each pixel shaded with its occlusion bits in registers), so
// the shading of the frame's light tiles runs while its heaviest tiles are still tracing; the
// star pass must have run before. TV: the traversal variant (TV_PLAIN, TV_MERGED, TV_WIDE4,
// TV_LDSTOP), each visiting the same sequence.
template <bool STATS, int SK, int TV, int SH, int W>
__global__ __launch_bounds__(64 * W) void k_trace(TraceArgs a) {
    static_assert(W == 1 || W == 2 || W == 4, "a block is 1, 2 or 4 quarters of one tile");
    static_assert(!(STATS && (TV == TV_WIDE4 || TV == TV_LDSTOP)), "instrumented renders use TV_PLAIN");
    static_assert(!(STATS && SH == 2), "instrumented renders shade in their own pass");
    static_assert(SK <= 0 || SH != 2 || StackSel<SK, 64 * W>::lds_entries * 8 >= TILE_CAP * 4,
                  "the fused shading's star list fits in the stacks' LDS");
    constexpr bool M = TV == TV_MERGED;
    const uint64_t clock0 = a.clock ? wall_clock64() : 0ull;
    if constexpr (TV == TV_LDSTOP) top_fill(a.S);
    using SS = StackSel<SK, 64 * W>;
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    // The wave's scheduling slot: blocks start in slot order, and the order table puts the tiles
    // that were slowest in the last frame first (each pixel's work does not depend on it).
    const uint32_t slot_q = blockIdx.x * W + (tid >> 6), q = slot_q & 3;
    const uint32_t t = a.order ? a.order[slot_q >> 2] : slot_q >> 2, vi = find_view(a.V.level_start[0], a.V.n, t);
    // The slowest tiles of the last render issue first on their SIMD while light waves share it.
    if ((slot_q >> 2) < a.prio_slots) __builtin_amdgcn_s_setprio(3);
    const specrel_frame &f = a.V.frames[vi];
    const ViewDerived v = derive(f, 0);
    const uint32_t lt = t - a.V.level_start[0][vi], tiles_x = (v.w + TILE - 1) / TILE;
    const uint32_t tx = (lt % tiles_x) * TILE, ty = (lt / tiles_x) * TILE;
    uint32_t px, py;
    if (a.wave_shape == 1) {  // 16x4 bands
        px = tx + (lane & 15);
        py = ty + q * 4 + (lane >> 4);
    } else if (a.wave_shape == 2) {  // 4x16 bands
        px = tx + q * 4 + (lane & 3);
        py = ty + (lane >> 2);
    } else {  // 8x8 quarters
        px = tx + (q & 1) * 8 + (lane & 7);
        py = ty + (q >> 1) * 8 + (lane >> 3);
    }
    __shared__ uint2 s_stack[SS::lds_entries];
    typename SS::type stk;
    typename SS::Private stk_priv;
    SS::init(stk, s_stack, tid, stk_priv);
    Counters C = {};
    uint32_t out_slot = SPECREL_NO_HIT, out_bits = 0u;  // for the fused shading
    if (px < v.w && py < v.h) {
        V4 dir = ray_dir(v, f, px, py);
        Hit hit;
        bool found = false;
        if (!a.S.empty) {
            bool redo = true;
            if (!STATS && !M && a.postpone)
                found = trace_pp<TV>(a.S, v.origin, dir, 1e-3f, 180.0f, hit, C, stk, (int)a.postpone, redo);
            if (redo)
                found = trace<STATS, typename SS::type, M, TV>(a.S, v.origin, dir, 1e-3f, 180.0f, hit, C, stk);
        }
        const size_t pi = (size_t)a.V.pix_start[vi] + (size_t)py * v.w + px;
        out_slot = found ? (uint32_t)hit.slot : SPECREL_NO_HIT;
        a.slot[pi] = out_slot;
        if constexpr (SH >= 1) {
            uint32_t bits = 0u;
            if (found && f.mode == SPECREL_MODE_SHADED && f.shadow_mode != SPECREL_SHADOW_OFF) {
                Rec rec;
                interpolate(a.S.attrs + 12 * (size_t)hit.slot, hit, rec);
                if (STATS) touch(a.S.touch_hits, (uint32_t)hit.slot);
                const Mat m = load_mat(a.S, rec.material);
                RecordShadow<STATS, typename SS::type, TV> sh{a.S, C, stk, 0u};
                illumination(a.S, f, dir, rec, m, sh);
                if (f.num_flashes > 0) flashes(a.S, f, rec, sh);
                bits = sh.mask;
            }
            a.mask[pi] = bits;
            out_bits = bits;
        }
    }
    if constexpr (SH == 2) {
        // Every wave of the block is past its traversals: the stacks' LDS holds the tile's star list.
        uint32_t *s_list = (uint32_t *)s_stack;
        __shared__ uint32_t s_wave[W], s_total;
        __syncthreads();
        const ViewDerived vs = derive(f, a.shade.num_stars);
        const StarList L = tile_stars<W>(a.shade, vi, vs, t, tx, ty, s_list, s_wave, &s_total);
        if (px < v.w && py < v.h) {
            LookupShadow sh{out_bits};
            shade_pixel<false>(a.shade, vi, f, vs, px, py, out_slot, sh, L, s_list, C);
        }
    }
    if (a.cost) {  // the quarter's cost: its lanes' longest traversal (primary and shadow steps)
        uint32_t c = C.steps;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) c = max(c, (uint32_t)__shfl_xor((int)c, off));
        if (lane == 0) a.cost[4 * (size_t)t + q] = c;
    }
    if (a.clock && lane == 0) {
        specrel_wave_clock w;
        w.start = clock0;
        w.end = wall_clock64();
        w.tile = t;
        w.quarter = q;
        w.hw_id = __builtin_amdgcn_s_getreg((31 << 11) | 4);  // HW_ID, all 32 bits
        w.xcc = __builtin_amdgcn_s_getreg((31 << 11) | 20);  // XCC_ID
        a.clock[slot_q] = w;
    }
    if (STATS) {
        __shared__ unsigned long long s_c[4];
        if (tid < 4) s_c[tid] = 0;
        __syncthreads();
        atomicAdd(&s_c[0], (unsigned long long)((px < v.w && py < v.h) ? 1 : 0));
        atomicAdd(&s_c[1], (unsigned long long)C.node_pops);
        atomicAdd(&s_c[2], (unsigned long long)C.box_tests);
        atomicAdd(&s_c[3], (unsigned long long)C.prim_tests);
        __syncthreads();
        if (tid < 4) atomicAdd(&a.stats[tid], s_c[tid]);
        if constexpr (SH >= 1) {
            atomicAdd(&a.stats[5], (unsigned long long)C.shadow_rays);
            atomicAdd(&a.stats[6], (unsigned long long)C.shadow_node_pops);
            atomicAdd(&a.stats[7], (unsigned long long)C.shadow_box_tests);
            atomicAdd(&a.stats[8], (unsigned long long)C.shadow_prim_tests);
        }
    }
}

// Persistent primary-ray kernel: every lane takes the next ray of a global queue (the 16x16 tiles
// of every view in order, 64 rays of one 8x8 quarter per grab) as soon as its own ray ends, so
// the waves stay full until the queue drains. Each ray's traversal is trace()'s, step for step.
template <bool STATS, int SK>
__global__ __launch_bounds__(256) void k_trace_p(TraceArgs a) {
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    __shared__ uint2 s_stack[StackSel<SK>::lds_entries];
    typename StackSel<SK>::type stk;
    typename StackSel<SK>::Private stk_priv;
    StackSel<SK>::init(stk, s_stack, tid, stk_priv);
    Counters C = {};
    uint64_t rays = 0;
    const uint32_t total = a.num_tiles * 256u;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t next = 0, end = 0, gv = 0;
    bool drained = false, active = false;
    RayState r;
    while (true) {
        uint64_t idle = __ballot(!active);
        if (idle != 0 && !drained) {
            if (next == end) {
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(a.queue, 64u);
                base = __shfl(base, 0);
                if (base >= total) drained = true;
                else {
                    next = base;
                    end = min(base + 64u, total);
                    gv = find_view(a.V.level_start[0], a.V.n, base >> 8);
                }
            }
            if (!drained) {
                uint32_t avail = end - next, rank = __popcll(idle & lt);
                if (!active && rank < avail) {
                    const specrel_frame &f = a.V.frames[gv];
                    const ViewDerived v = derive(f, 0);
                    uint32_t ray = next + rank, q = (ray >> 6) & 3u, l = ray & 63u;
                    uint32_t tl = (ray >> 8) - a.V.level_start[0][gv], tiles_x = (v.w + TILE - 1) / TILE;
                    uint32_t px = (tl % tiles_x) * TILE + (q & 1) * 8 + (l & 7);
                    uint32_t py = (tl / tiles_x) * TILE + (q >> 1) * 8 + (l >> 3);
                    if (px < v.w && py < v.h) {
                        r.pix = a.V.pix_start[gv] + py * v.w + px;
                        r.o = v.origin;
                        r.d = ray_dir(v, f, px, py);
                        r.tmax = 180.0f;
                        if (STATS) rays++;
                        if (trace_begin(a.S, r, 1e-3f, stk)) active = true;
                        else a.slot[r.pix] = SPECREL_NO_HIT;
                    }
                }
                next += min((uint32_t)__popcll(idle), avail);
            }
        }
        if (__ballot(active) == 0) {
            if (drained) break;
            continue;
        }
        if (active && !trace_step<STATS>(a.S, r, 1e-3f, C, stk)) {
            a.slot[r.pix] = r.best >= 0 ? (uint32_t)r.best : SPECREL_NO_HIT;
            active = false;
        }
    }
    if (STATS) {
        atomicAdd(&a.stats[0], (unsigned long long)rays);
        atomicAdd(&a.stats[1], (unsigned long long)C.node_pops);
        atomicAdd(&a.stats[2], (unsigned long long)C.box_tests);
        atomicAdd(&a.stats[3], (unsigned long long)C.prim_tests);
    }
}

// Persistent any-hit kernel over the queued shadow queries; an occluded query sets bit `id`
// of its pixel's mask.
template <bool STATS, int SK>
__global__ __launch_bounds__(256) void k_occlude_p(OccludeArgs a) {
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    __shared__ uint2 s_stack[StackSel<SK>::lds_entries];
    typename StackSel<SK>::type stk;
    typename StackSel<SK>::Private stk_priv;
    StackSel<SK>::init(stk, s_stack, tid, stk_priv);
    Counters C = {};
    uint64_t queries = 0;
    const uint32_t total = min(*a.count, a.cap);
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t next = 0, end = 0, pixel = 0, id = 0;
    float tmin = 0.0f;
    bool drained = false, active = false;
    RayState r;
    while (true) {
        uint64_t idle = __ballot(!active);
        if (idle != 0 && !drained) {
            if (next == end) {
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(a.next, 64u);
                base = __shfl(base, 0);
                if (base >= total) drained = true;
                else {
                    next = base;
                    end = min(base + 64u, total);
                }
            }
            if (!drained) {
                uint32_t avail = end - next, rank = __popcll(idle & lt);
                if (!active && rank < avail) {
                    const ShadowQuery q = a.queue[next + rank];
                    r.o = f4v(q.o);
                    r.d = f4v(q.d);
                    r.tmax = q.tmax;
                    tmin = q.tmin;
                    pixel = q.pixel;
                    id = q.id;
                    if (STATS) queries++;
                    active = trace_begin(a.S, r, tmin, stk);
                }
                next += min((uint32_t)__popcll(idle), avail);
            }
        }
        if (__ballot(active) == 0) {
            if (drained) break;
            continue;
        }
        if (active) {
            int st = occlude_step<STATS>(a.S, r, tmin, C, stk);
            if (st == 1) atomicOr(&a.mask[pixel], 1u << id);
            if (st != 0) active = false;
        }
    }
    if (STATS) {
        atomicAdd(&a.stats[5], (unsigned long long)queries);
        atomicAdd(&a.stats[6], (unsigned long long)C.shadow_node_pops);
        atomicAdd(&a.stats[7], (unsigned long long)C.shadow_box_tests);
        atomicAdd(&a.stats[8], (unsigned long long)C.shadow_prim_tests);
    }
}

// Order of the next frame's primary-ray tiles: by decreasing cost of their slowest quarter (64
// buckets of 8 steps, a counting sort; the order inside a bucket is whatever the atomics give,
// which changes nothing but the schedule). order becomes a permutation of the tiles.
__device__ inline uint32_t tile_bucket(const uint32_t *cost, uint32_t t) {
    const uint4 c = reinterpret_cast<const uint4 *>(cost)[t];
    return min(63u, max(max(c.x, c.y), max(c.z, c.w)) >> 3);
}

// Lane b of a full wave: the sum of v over lanes b+1..63 (the buckets of higher cost).
__device__ inline uint32_t sum_above(uint32_t v, uint32_t lane) {
    uint32_t s = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_down(s, off, 64);
        s += lane + off < 64 ? o : 0u;
    }
    return s - v;
}

// One block (any size) sorts every tile: up to 4,096 tiles.
__device__ void tile_order_block(const uint32_t *cost, uint32_t n_tiles, uint32_t *order) {
    __shared__ uint32_t hist[64];
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    if (tid < 64) hist[tid] = 0;
    __syncthreads();
    for (uint32_t t = tid; t < n_tiles; t += nt) atomicAdd(&hist[tile_bucket(cost, t)], 1u);
    __syncthreads();
    if (tid < 64) {  // wave 0
        const uint32_t h = hist[tid];
        hist[tid] = sum_above(h, tid);
    }
    __syncthreads();
    for (uint32_t t = tid; t < n_tiles; t += nt) order[atomicAdd(&hist[tile_bucket(cost, t)], 1u)] = t;
}

__global__ __launch_bounds__(1024) void k_tile_order(const uint32_t *cost, uint32_t n_tiles, uint32_t *order) {
    tile_order_block(cost, n_tiles, order);
}

// Larger grids: TILE_ORDER_CHUNK tiles per 256-thread block in two launches. k_tile_hist adds
// each block's bucket counts into total[64]; k_tile_scatter turns total into bucket starts
// (every block the same), reserves its block's range in each bucket with cursor[64] and writes
// its tiles there. total and cursor start at zero (the frame's k_reset clears them).
__device__ inline void chunk_hist(const uint32_t *cost, uint32_t n_tiles, uint32_t *hist) {
    const uint32_t tid = threadIdx.x, t0 = blockIdx.x * TILE_ORDER_CHUNK;
    if (tid < 64) hist[tid] = 0;
    __syncthreads();
    for (uint32_t t = t0 + tid; t < min(n_tiles, t0 + TILE_ORDER_CHUNK); t += 256)
        atomicAdd(&hist[tile_bucket(cost, t)], 1u);
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_tile_hist(const uint32_t *cost, uint32_t n_tiles, uint32_t *total) {
    __shared__ uint32_t hist[64];
    chunk_hist(cost, n_tiles, hist);
    if (threadIdx.x < 64 && hist[threadIdx.x]) atomicAdd(&total[threadIdx.x], hist[threadIdx.x]);
}

__global__ __launch_bounds__(256) void k_tile_scatter(const uint32_t *cost, uint32_t n_tiles, const uint32_t *total,
                                                      uint32_t *cursor, uint32_t *order) {
    __shared__ uint32_t hist[64];
    chunk_hist(cost, n_tiles, hist);
    const uint32_t tid = threadIdx.x, t0 = blockIdx.x * TILE_ORDER_CHUNK;
    if (tid < 64) {  // wave 0
        const uint32_t h = hist[tid];
        const uint32_t start = sum_above(total[tid], tid);
        hist[tid] = start + (h ? atomicAdd(&cursor[tid], h) : 0u);
    }
    __syncthreads();
    for (uint32_t t = t0 + tid; t < min(n_tiles, t0 + TILE_ORDER_CHUNK); t += 256)
        order[atomicAdd(&hist[tile_bucket(cost, t)], 1u)] = t;
}

// Row order and zero costs, after a change of the view sizes.
__global__ void k_tile_init(uint32_t *order, uint32_t *cost, uint32_t n_tiles) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tiles) return;
    order[t] = t;
    reinterpret_cast<uint4 *>(cost)[t] = make_uint4(0, 0, 0, 0);
}

// specrel_patch_scene: node i's new bounds go to the child-box pair its parent holds
// (pslot = parent << 1 | right side).
__global__ void k_patch_nodes(const uint32_t *pslot, const float4 *bounds, uint32_t n, float4 *nbox) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    size_t at = 4 * (size_t)(pslot[i] >> 1) + 2 * (pslot[i] & 1u);
    nbox[at] = bounds[2 * (size_t)i];
    nbox[at + 1] = bounds[2 * (size_t)i + 1];
}

// specrel_update_scene: whole rows (child boxes and entries) of the listed interior nodes.
__global__ void k_patch_rows(const uint32_t *idx, const float4 *rows, const uint2 *info, uint32_t n, float4 *nbox,
                             uint2 *ninfo) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t r = idx[i];
    for (int k = 0; k < 4; ++k) nbox[4 * r + k] = rows[4 * (size_t)i + k];
    ninfo[r] = info[i];
}

// Set bits of n words, added to *out (the distinct records of an instrumented render).
__global__ void k_popcount(const uint32_t *bits, uint64_t n, unsigned long long *out) {
    uint32_t c = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        c += __popc(bits[i]);
    if (c) atomicAdd(out, (unsigned long long)c);
}

// Zeroes the per-frame queue counters (a kernel rather than a memset, so the frame's work can
// be captured in a graph).

__global__ void k_reset(uint32_t *counters, uint32_t n) {
    uint32_t i = threadIdx.x;
    if (i < n) counters[i] = 0u;
}

// The same for a render of one view launched directly, which also writes its frame (passed by
// value) to the device frame table: one launch instead of a copy and a launch.
__global__ void k_reset_frame(uint32_t *counters, uint32_t n, FrameWords f, uint32_t *dst) {
    uint32_t i = threadIdx.x;
    if (i < n) counters[i] = 0u;
    if (i < FrameWords::N) dst[i] = f.w[i];
}

// The same for a view rendered without any copy from the host: the view table comes as an
// argument too (specrel_render_noalloc).
__global__ void k_reset_view(uint32_t *counters, uint32_t n, FrameWords f, uint32_t *dst_frame, TableWords t,
                             uint32_t *dst_table) {
    uint32_t i = threadIdx.x;
    if (i < n) counters[i] = 0u;
    if (i < FrameWords::N) dst_frame[i] = f.w[i];
    if (i < 2 * TABLE_ROWS) dst_table[i] = t.w[i];
}

// Shading: the winner's record is intersected again (same inputs, same float operations, so the
// same lambda, event and barycentrics), interpolated and shaded with shadow rays, haze and the
// tile's stars; writes the HDR pixel and the optional per-pixel outputs.
//   PASS 0: shadow rays traced in place.
//   PASS 1: shadow queries queued (and the pixel's occlusion mask cleared); nothing else written.
//   PASS 2: shadow queries answered from the occlusion masks.
template <bool STATS, int SK, int PASS>
__global__ __launch_bounds__(256, (PASS == 2 && !STATS) ? 4 : 1) void k_shade(ShadeArgs a) {
    const uint32_t t = blockIdx.x, vi = find_view(a.V.level_start[0], a.V.n, t);
    const specrel_frame &f = a.V.frames[vi];
    const ViewDerived v = derive(f, a.num_stars);
    const uint32_t lt = t - a.V.level_start[0][vi], tiles_x = (v.w + TILE - 1) / TILE;
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const uint32_t x0 = (lt % tiles_x) * TILE, y0 = (lt / tiles_x) * TILE;
    const uint32_t px = x0 + (wave & 1) * 8 + (lane & 7);
    const uint32_t py = y0 + (wave >> 1) * 8 + (lane >> 3);
    __shared__ uint2 s_stack[PASS == 0 ? StackSel<SK>::lds_entries : 1];
    __shared__ uint32_t s_list[PASS == 1 ? 1 : TILE_CAP];
    __shared__ uint32_t s_wave[4], s_total;
    typename StackSel<SK>::type stk;
    typename StackSel<SK>::Private stk_priv;
    if constexpr (PASS == 0) StackSel<SK>::init(stk, s_stack, tid, stk_priv);
    Counters C = {};
    // The pixel's hit slot and occlusion bits, loaded before the star culling so that their
    // latency passes during it.
    const bool inside = px < v.w && py < v.h;
    const size_t pi = a.V.pix_start[vi] + (size_t)py * v.w + px;
    const uint32_t slot0 = inside ? a.slot[pi] : SPECREL_NO_HIT;
    const uint32_t mask0 = PASS == 2 && inside ? a.mask[pi] : 0u;
    if constexpr (PASS == 1) {  // queue the shadow queries of the hit pixels
        if (inside) {
            a.mask[pi] = 0u;
            if (slot0 != SPECREL_NO_HIT) {
                V4 dir = ray_dir(v, f, px, py);
                Hit hit;
                hit.slot = (int32_t)slot0;
                intersect(a.S.geoms + 5 * (size_t)slot0, v.origin, dir, 1e-3f, 180.0f, hit.lambda, hit.event, hit.bary);
                Rec rec;
                interpolate(a.S.attrs + 12 * (size_t)slot0, hit, rec);
                EnqueueShadow sh{a.queue, a.qcount, a.qcap, (uint32_t)pi};
                shade_hit(a.S, a.T, f, dir, rec, sh);
            }
        }
    } else {
        const StarList L = tile_stars<4>(a, vi, v, t, x0, y0, s_list, s_wave, &s_total);
        if (inside) {
            if constexpr (PASS == 0) {
                TraceShadow<STATS, typename StackSel<SK>::type> sh{a.S, C, stk};
                shade_pixel<STATS>(a, vi, f, v, px, py, slot0, sh, L, s_list, C);
            } else {
                LookupShadow sh{mask0};
                shade_pixel<STATS>(a, vi, f, v, px, py, slot0, sh, L, s_list, C);
            }
        }
    }
    if (STATS && PASS != 1) {
        __shared__ unsigned long long s_c[6];
        if (tid < 6) s_c[tid] = 0;
        __syncthreads();
        atomicAdd(&s_c[0], (unsigned long long)C.hits);
        atomicAdd(&s_c[1], (unsigned long long)C.shadow_rays);
        atomicAdd(&s_c[2], (unsigned long long)C.shadow_node_pops);
        atomicAdd(&s_c[3], (unsigned long long)C.shadow_box_tests);
        atomicAdd(&s_c[4], (unsigned long long)C.shadow_prim_tests);
        atomicAdd(&s_c[5], (unsigned long long)C.star_tests);
        __syncthreads();
        if (tid < 6) atomicAdd(&a.stats[4 + tid], s_c[tid]);
    }
}

#define SPECREL_INSTANTIATE(K)                                   \
    template __global__ void k_trace_p<false, K>(TraceArgs);     \
    template __global__ void k_trace_p<true, K>(TraceArgs);      \
    template __global__ void k_occlude_p<false, K>(OccludeArgs); \
    template __global__ void k_occlude_p<true, K>(OccludeArgs);  \
    template __global__ void k_shade<false, K, 0>(ShadeArgs);    \
    template __global__ void k_shade<true, K, 0>(ShadeArgs);     \
    template __global__ void k_shade<false, K, 1>(ShadeArgs);    \
    template __global__ void k_shade<true, K, 1>(ShadeArgs);     \
    template __global__ void k_shade<false, K, 2>(ShadeArgs);    \
    template __global__ void k_shade<true, K, 2>(ShadeArgs);
#define SPECREL_TRACE(K, TV, SH, W)                                    \
    template __global__ void k_trace<false, K, TV, SH, W>(TraceArgs); \
    template __global__ void k_trace<true, K, TV, SH, W>(TraceArgs);
#define SPECREL_TRACE_FAST(K, TV, SH, W) template __global__ void k_trace<false, K, TV, SH, W>(TraceArgs);
// LDS stacks: plain or merged load rounds (with instrumented twins), the fetch-width-4 and
// LDS-top variants for stacks of 16 and 8 (uninstrumented only); shadow queries traced here or
// not, and for stacks of 16 and 8 the fused shading; blocks of 1, 2 or 4 waves. Register stacks:
// plain rounds, blocks of 4 waves.
#define SPECREL_TRACE_LDS_W(K, W)                                                                    \
    SPECREL_TRACE(K, TV_PLAIN, 0, W) SPECREL_TRACE(K, TV_MERGED, 0, W)                               \
    SPECREL_TRACE(K, TV_PLAIN, 1, W) SPECREL_TRACE(K, TV_MERGED, 1, W)
#define SPECREL_TRACE_LDS_X(K, W)                                                                    \
    SPECREL_TRACE_FAST(K, TV_WIDE4, 0, W) SPECREL_TRACE_FAST(K, TV_WIDE4, 1, W)                      \
    SPECREL_TRACE_FAST(K, TV_LDSTOP, 0, W) SPECREL_TRACE_FAST(K, TV_LDSTOP, 1, W)                    \
    SPECREL_TRACE_FAST(K, TV_PLAIN, 2, W) SPECREL_TRACE_FAST(K, TV_WIDE4, 2, W)                      \
    SPECREL_TRACE_FAST(K, TV_LDSTOP, 2, W)
#define SPECREL_TRACE_LDS(K) SPECREL_TRACE_LDS_W(K, 1) SPECREL_TRACE_LDS_W(K, 2) SPECREL_TRACE_LDS_W(K, 4)
#define SPECREL_TRACE_LDS_XW(K) SPECREL_TRACE_LDS_X(K, 1) SPECREL_TRACE_LDS_X(K, 2) SPECREL_TRACE_LDS_X(K, 4)
SPECREL_INSTANTIATE(8)
SPECREL_INSTANTIATE(10)
SPECREL_INSTANTIATE(12)
SPECREL_INSTANTIATE(16)
SPECREL_TRACE_LDS(8)
SPECREL_TRACE_LDS(10)
SPECREL_TRACE_LDS(12)
SPECREL_TRACE_LDS(16)
SPECREL_TRACE_LDS_XW(8)
SPECREL_TRACE_LDS_XW(16)
SPECREL_INSTANTIATE(-4)
SPECREL_INSTANTIATE(-8)
SPECREL_INSTANTIATE(-12)
SPECREL_INSTANTIATE(-16)
SPECREL_TRACE(-4, TV_PLAIN, 0, 4)
SPECREL_TRACE(-8, TV_PLAIN, 0, 4)
SPECREL_TRACE(-12, TV_PLAIN, 0, 4)
SPECREL_TRACE(-16, TV_PLAIN, 0, 4)

// ---------------------------------------------------------------------------------------
// Bloom (bloom.rs:22-125): the pyramid lives after the full-resolution HDR image.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ V3 ld3(const float4 *p, size_t i) {
    float4 a = p[i];
    return V3{a.x, a.y, a.z};
}

// Pixel (x, y) of a 16x16 tile of pyramid level k, for the tile's global index over all views.
struct LevelPix {
    uint32_t vi, w, h, x, y;
};
__device__ __forceinline__ LevelPix level_pix(const Views &V, int k) {
    LevelPix p;
    const uint32_t t = blockIdx.x;
    p.vi = find_view(V.level_start[k], V.n, t);
    const specrel_frame &f = V.frames[p.vi];
    p.w = f.size[0];
    p.h = f.size[1];
    uint32_t lw, lh;
    level_size(p.w, p.h, k, lw, lh);
    const uint32_t lt = t - V.level_start[k][p.vi], tiles_x = (lw + 15) / 16;
    p.x = (lt % tiles_x) * 16 + (threadIdx.x & 15);
    p.y = (lt / tiles_x) * 16 + (threadIdx.x >> 4);
    return p;
}

// bloom::downsample (bloom.rs:54-75): pixel (x, y) of a level from the finer level src.
// The value of pixel (x, y) of a level from the finer level, whose pixel (sx, sy) src(sx, sy)
// returns (taps clamped into the finer level).
template <class F>
__device__ __forceinline__ V3 down_at(F src, uint32_t sw, uint32_t sh, uint32_t x, uint32_t y) {
    const float W[4] = {0.125f, 0.375f, 0.375f, 0.125f};
    V3 acc = V3{0.0f, 0.0f, 0.0f};
    for (int j = 0; j < 4; ++j) {
        int64_t syy = 2 * (int64_t)y - 1 + j;
        uint32_t sy = (uint32_t)(syy < 0 ? 0 : (syy > (int64_t)sh - 1 ? (int64_t)sh - 1 : syy));
        V3 row = V3{0.0f, 0.0f, 0.0f};
        for (int i = 0; i < 4; ++i) {
            int64_t sxx = 2 * (int64_t)x - 1 + i;
            uint32_t sx = (uint32_t)(sxx < 0 ? 0 : (sxx > (int64_t)sw - 1 ? (int64_t)sw - 1 : sxx));
            row += src(sx, sy) * W[i];
        }
        acc += row * W[j];
    }
    return acc;
}
__device__ __forceinline__ void downsample_px(const float4 *src, uint32_t sw, uint32_t sh, float4 *dst, uint32_t dw,
                                              uint32_t x, uint32_t y) {
    const V3 acc = down_at([&](uint32_t sx, uint32_t sy) { return ld3(src, (size_t)sy * sw + sx); }, sw, sh, x, y);
    dst[(size_t)y * dw + x] = make_float4(acc.x, acc.y, acc.z, 0.0f);
}

// bloom::upsample_at (bloom.rs:80-90), the coarse level's pixel (cx, cy) given by c(cx, cy).
template <class F>
__device__ __forceinline__ V3 up_at(F c, uint32_t cw, uint32_t ch, uint32_t x, uint32_t y) {
    uint32_t mx = x / 2, my = y / 2;
    int64_t ox = (x % 2 == 0) ? (int64_t)mx - 1 : (int64_t)mx + 1;
    int64_t oy = (y % 2 == 0) ? (int64_t)my - 1 : (int64_t)my + 1;
    uint32_t x1 = (uint32_t)(ox < 0 ? 0 : (ox > (int64_t)cw - 1 ? (int64_t)cw - 1 : ox));
    uint32_t y1 = (uint32_t)(oy < 0 ? 0 : (oy > (int64_t)ch - 1 ? (int64_t)ch - 1 : oy));
    V3 r0 = c(mx, my) * 0.75f + c(x1, my) * 0.25f;
    V3 r1 = c(mx, y1) * 0.75f + c(x1, y1) * 0.25f;
    return r0 * 0.75f + r1 * 0.25f;
}
__device__ __forceinline__ V3 upsample_at(const float4 *coarse, uint32_t cw, uint32_t ch, uint32_t x, uint32_t y) {
    return up_at([&](uint32_t cx, uint32_t cy) { return ld3(coarse, (size_t)cy * cw + cx); }, cw, ch, x, y);
}

// One step of the up pass (bloom.rs:110-114) at pixel (x, y): fine += upsample(coarse).
__device__ __forceinline__ void upsample_px(float4 *fine, uint32_t w, const float4 *coarse, uint32_t cw, uint32_t ch,
                                            uint32_t x, uint32_t y) {
    size_t i = (size_t)y * w + x;
    V3 c = ld3(fine, i) + upsample_at(coarse, cw, ch, x, y);
    fine[i] = make_float4(c.x, c.y, c.z, 0.0f);
}

// Level k of every view from level k-1.
__global__ __launch_bounds__(256) void k_bloom_down(Views V, int k, float4 *pyr) {
    const LevelPix p = level_pix(V, k);
    uint32_t sw, sh, dw, dh;
    level_size(p.w, p.h, k - 1, sw, sh);
    level_size(p.w, p.h, k, dw, dh);
    if (p.x >= dw || p.y >= dh) return;
    float4 *base = pyr + V.pyr_start[p.vi];
    downsample_px(base + level_offset(p.w, p.h, k - 1), sw, sh, base + level_offset(p.w, p.h, k), dw, p.x, p.y);
}

// Level k += upsample(level k+1), every view.
__global__ __launch_bounds__(256) void k_bloom_up(Views V, int k, float4 *pyr) {
    const LevelPix p = level_pix(V, k);
    uint32_t w, h, cw, ch;
    level_size(p.w, p.h, k, w, h);
    level_size(p.w, p.h, k + 1, cw, ch);
    if (p.x >= w || p.y >= h) return;
    float4 *base = pyr + V.pyr_start[p.vi];
    upsample_px(base + level_offset(p.w, p.h, k), w, base + level_offset(p.w, p.h, k + 1), cw, ch, p.x, p.y);
}

// The small levels of one view per block: down k0..SPECREL_BLOOM_LEVELS, then up
// SPECREL_BLOOM_LEVELS-1..k0, each level finished by the block before the next reads it (one
// launch instead of 2*(SPECREL_BLOOM_LEVELS-k0)+1).
__global__ __launch_bounds__(1024) void k_bloom_tail(Views V, int k0, float4 *pyr) {
    const uint32_t vi = blockIdx.x;
    const specrel_frame &f = V.frames[vi];
    const uint32_t w0 = f.size[0], h0 = f.size[1];
    float4 *base = pyr + V.pyr_start[vi];
    for (int k = k0; k <= SPECREL_BLOOM_LEVELS; ++k) {
        uint32_t sw, sh, dw, dh;
        level_size(w0, h0, k - 1, sw, sh);
        level_size(w0, h0, k, dw, dh);
        const float4 *src = base + level_offset(w0, h0, k - 1);
        float4 *dst = base + level_offset(w0, h0, k);
        for (uint32_t i = threadIdx.x; i < dw * dh; i += 1024) downsample_px(src, sw, sh, dst, dw, i % dw, i / dw);
        __syncthreads();
    }
    for (int k = SPECREL_BLOOM_LEVELS - 1; k >= k0; --k) {
        uint32_t w, h, cw, ch;
        level_size(w0, h0, k, w, h);
        level_size(w0, h0, k + 1, cw, ch);
        float4 *fine = base + level_offset(w0, h0, k);
        const float4 *coarse = base + level_offset(w0, h0, k + 1);
        for (uint32_t i = threadIdx.x; i < w * h; i += 1024) upsample_px(fine, w, coarse, cw, ch, i % w, i / w);
        __syncthreads();
    }
}

// Final mix (bloom.rs:116-124) and pack_rgb_with_lut (renderer.rs:1116-1129), every view.
// bloom::apply's composite (bloom.rs:93-125) of pixel (x, y) from the linear image and level 1
// (pixel (cx, cy) given by l1(cx, cy)), tone map and LUT pack (renderer.rs:1116-1129).
template <class F>
__device__ __forceinline__ uint32_t composite_at(const float4 *hdr, F l1, uint32_t w, uint32_t w1, uint32_t h1,
                                                 uint32_t x, uint32_t y, const uint8_t *lut) {
    size_t i = (size_t)y * w + x;
    float halo = 0.04f / 6.0f;
    V3 c = ld3(hdr, i) * (1.0f - 0.04f) + up_at(l1, w1, h1, x, y) * halo;
    V3 m = clamp3(V3{c.x / (1.0f + c.x * 0.35f), c.y / (1.0f + c.y * 0.35f), c.z / (1.0f + c.z * 0.35f)}, 0.0f, 1.0f);
    const float scale = 4095.0f;
    uint32_t r = lut[f2u_sat(m.x * scale)], g = lut[f2u_sat(m.y * scale)], b = lut[f2u_sat(m.z * scale)];
    return (r << 16) | (g << 8) | b;
}
__global__ __launch_bounds__(256) void k_composite(Views V, const float4 *pyr, const uint8_t *lut, uint32_t *out) {
    const LevelPix p = level_pix(V, 0);
    const uint32_t w = p.w, h = p.h, x = p.x, y = p.y;
    if (x >= w || y >= h) return;
    uint32_t w1, h1;
    level_size(w, h, 1, w1, h1);
    const float4 *hdr = pyr + V.pyr_start[p.vi];
    const float4 *l1 = hdr + level_offset(w, h, 1);
    out[V.pix_start[p.vi] + (size_t)y * w + x] = composite_at(
        hdr, [&](uint32_t cx, uint32_t cy) { return ld3(l1, (size_t)cy * w1 + cx); }, w, w1, h1, x, y, lut);
}

// Fused bloom passes. A block takes a 16x16 tile of one level and keeps in LDS the rectangle of
// a neighbouring level its pixels read (taps clamped into that level, so the rectangle spans at
// most 18 x 18 pixels); each pixel is computed with the same operations as by the one-level
// kernels, so the pyramid and the image are the same.
struct LdsRect {
    float *v;  // 3 floats per pixel, row-major
    uint32_t x0, y0, nx;
    __device__ V3 at(uint32_t x, uint32_t y) const {
        const float *p = v + 3 * ((size_t)(y - y0) * nx + (x - x0));
        return V3{p[0], p[1], p[2]};
    }
    __device__ void put(uint32_t i, V3 c) const {
        v[3 * i] = c.x, v[3 * i + 1] = c.y, v[3 * i + 2] = c.z;
    }
};
// The rectangle of a coarser level (cw x ch) that up_at reads for fine pixels x in [fx0, fx1].
__device__ __forceinline__ void up_span(uint32_t fx0, uint32_t fx1, uint32_t cw, uint32_t &a, uint32_t &b) {
    a = fx0 / 2 > 0 ? fx0 / 2 - 1 : 0;
    b = min(fx1 / 2 + 1, cw - 1);
}
struct TileOf {
    uint32_t vi, w0, h0, tx0, ty0;
};
__device__ __forceinline__ TileOf tile_of(const Views &V, int k) {
    TileOf o;
    const uint32_t t = blockIdx.x;
    o.vi = find_view(V.level_start[k], V.n, t);
    const specrel_frame &f = V.frames[o.vi];
    o.w0 = f.size[0];
    o.h0 = f.size[1];
    uint32_t lw, lh;
    level_size(o.w0, o.h0, k, lw, lh);
    const uint32_t lt = t - V.level_start[k][o.vi], tiles_x = (lw + 15) / 16;
    o.tx0 = (lt % tiles_x) * 16;
    o.ty0 = (lt / tiles_x) * 16;
    return o;
}

// Levels k and k + 1 of the down pass: the tile of level k (from level k - 1) and the level-(k+1)
// pixels X, Y with 2X, 2Y in the tile (from the tile and its one-pixel rim, in LDS).
// STAGE: the finer level's pixels the tile's taps read are first copied into LDS (one coalesced
// load each instead of 16 loads per pixel from the caches): faster on grids of many tiles (batches),
// slower on a single view's 1,000 tiles, where the extra phase lengthens every block.
template <bool STAGE>
__global__ __launch_bounds__(256) void k_bloom_down2(Views V, int k, float4 *pyr) {
    __shared__ float s_lv[18 * 18 * 3];
    __shared__ float s_src[STAGE ? 38 * 38 * 3 : 1];  // the finer level's pixels the rim tile's taps read
    const TileOf o = tile_of(V, k);
    uint32_t sw, sh, dw, dh, ew, eh;
    level_size(o.w0, o.h0, k - 1, sw, sh);
    level_size(o.w0, o.h0, k, dw, dh);
    level_size(o.w0, o.h0, k + 1, ew, eh);
    float4 *base = pyr + V.pyr_start[o.vi];
    const float4 *src = base + level_offset(o.w0, o.h0, k - 1);
    float4 *lk = base + level_offset(o.w0, o.h0, k);
    float4 *lk1 = base + level_offset(o.w0, o.h0, k + 1);
    const uint32_t rx0 = o.tx0 > 0 ? o.tx0 - 1 : 0, ry0 = o.ty0 > 0 ? o.ty0 - 1 : 0;
    const uint32_t rx1 = min(o.tx0 + 16, dw - 1), ry1 = min(o.ty0 + 16, dh - 1);
    const LdsRect R{s_lv, rx0, ry0, rx1 - rx0 + 1};
    const uint32_t n = R.nx * (ry1 - ry0 + 1);
    // The finer level's rectangle the taps of the rim tile read (clamped into that level, at most
    // 38 x 38 pixels).
    const uint32_t sx0 = 2 * rx0 > 0 ? 2 * rx0 - 1 : 0, sy0 = 2 * ry0 > 0 ? 2 * ry0 - 1 : 0;
    const uint32_t sx1 = min(2 * rx1 + 2, sw - 1), sy1 = min(2 * ry1 + 2, sh - 1);
    const LdsRect S{s_src, sx0, sy0, sx1 - sx0 + 1};
    if constexpr (STAGE) {
        const uint32_t ns = S.nx * (sy1 - sy0 + 1);
        for (uint32_t i = threadIdx.x; i < ns; i += 256)
            S.put(i, ld3(src, (size_t)(sy0 + i / S.nx) * sw + sx0 + i % S.nx));
        __syncthreads();
    }
    for (uint32_t i = threadIdx.x; i < n; i += 256) {
        const uint32_t x = rx0 + i % R.nx, y = ry0 + i / R.nx;
        const V3 v = STAGE ? down_at([&](uint32_t a, uint32_t b) { return S.at(a, b); }, sw, sh, x, y)
                           : down_at([&](uint32_t a, uint32_t b) { return ld3(src, (size_t)b * sw + a); }, sw, sh, x, y);
        R.put(i, v);
        if (x >= o.tx0 && x < o.tx0 + 16 && y >= o.ty0 && y < o.ty0 + 16)
            lk[(size_t)y * dw + x] = make_float4(v.x, v.y, v.z, 0.0f);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const uint32_t X = o.tx0 / 2 + (threadIdx.x & 7), Y = o.ty0 / 2 + (threadIdx.x >> 3);
        if (X < ew && Y < eh) {
            const V3 v = down_at([&](uint32_t a, uint32_t b) { return R.at(a, b); }, dw, dh, X, Y);
            lk1[(size_t)Y * ew + X] = make_float4(v.x, v.y, v.z, 0.0f);
        }
    }
}

// Levels k + 1 and k of the up pass: level k + 1 updated from level k + 2 (final) over the rectangle
// the tile of level k reads (in LDS only: nothing later reads it), then the tile of level k
// updated from it and written.
__global__ __launch_bounds__(256) void k_bloom_up2(Views V, int k, float4 *pyr) {
    __shared__ float s_lv[10 * 10 * 3];
    const TileOf o = tile_of(V, k);
    uint32_t fw, fh, cw, ch, gw, gh;
    level_size(o.w0, o.h0, k, fw, fh);
    level_size(o.w0, o.h0, k + 1, cw, ch);
    level_size(o.w0, o.h0, k + 2, gw, gh);
    float4 *base = pyr + V.pyr_start[o.vi];
    float4 *lk = base + level_offset(o.w0, o.h0, k);
    const float4 *lk1 = base + level_offset(o.w0, o.h0, k + 1);
    const float4 *lk2 = base + level_offset(o.w0, o.h0, k + 2);
    uint32_t rx0, rx1, ry0, ry1;
    up_span(o.tx0, min(o.tx0 + 15, fw - 1), cw, rx0, rx1);
    up_span(o.ty0, min(o.ty0 + 15, fh - 1), ch, ry0, ry1);
    const LdsRect R{s_lv, rx0, ry0, rx1 - rx0 + 1};
    const uint32_t n = R.nx * (ry1 - ry0 + 1);
    for (uint32_t i = threadIdx.x; i < n; i += 256) {
        const uint32_t X = rx0 + i % R.nx, Y = ry0 + i / R.nx;
        R.put(i, ld3(lk1, (size_t)Y * cw + X) +
                     up_at([&](uint32_t a, uint32_t b) { return ld3(lk2, (size_t)b * gw + a); }, gw, gh, X, Y));
    }
    __syncthreads();
    const uint32_t x = o.tx0 + (threadIdx.x & 15), y = o.ty0 + (threadIdx.x >> 4);
    if (x < fw && y < fh) {
        const size_t i = (size_t)y * fw + x;
        const V3 c = ld3(lk, i) + up_at([&](uint32_t a, uint32_t b) { return R.at(a, b); }, cw, ch, x, y);
        lk[i] = make_float4(c.x, c.y, c.z, 0.0f);
    }
}

// The last NUP levels of the up pass (2: levels 2 and 1; 1: level 1; 0: none) over the rectangles a
// tile of the image reads, in LDS, then the composite of the tile.
template <int NUP>
__global__ __launch_bounds__(256) void k_bloom_final(Views V, const float4 *pyr, const uint8_t *lut, uint32_t *out) {
    __shared__ float s_l1[10 * 10 * 3];
    __shared__ float s_l2[(NUP == 2 ? 8 * 8 : 1) * 3];
    const TileOf o = tile_of(V, 0);
    const uint32_t w = o.w0, h = o.h0;
    uint32_t w1, h1, w2, h2, w3, h3;
    level_size(w, h, 1, w1, h1);
    level_size(w, h, 2, w2, h2);
    level_size(w, h, 3, w3, h3);
    const float4 *hdr = pyr + V.pyr_start[o.vi];
    const float4 *l1 = hdr + level_offset(w, h, 1);
    const float4 *l2 = hdr + level_offset(w, h, 2);
    const float4 *l3 = hdr + level_offset(w, h, 3);
    uint32_t ax0, ax1, ay0, ay1;  // the level-1 rectangle the tile reads
    up_span(o.tx0, min(o.tx0 + 15, w - 1), w1, ax0, ax1);
    up_span(o.ty0, min(o.ty0 + 15, h - 1), h1, ay0, ay1);
    const LdsRect R1{s_l1, ax0, ay0, ax1 - ax0 + 1};
    if constexpr (NUP >= 1) {
        const uint32_t n1 = R1.nx * (ay1 - ay0 + 1);
        if constexpr (NUP == 2) {
            uint32_t bx0, bx1, by0, by1;  // the level-2 rectangle those read
            up_span(ax0, ax1, w2, bx0, bx1);
            up_span(ay0, ay1, h2, by0, by1);
            const LdsRect R2{s_l2, bx0, by0, bx1 - bx0 + 1};
            const uint32_t n2 = R2.nx * (by1 - by0 + 1);
            for (uint32_t i = threadIdx.x; i < n2; i += 256) {
                const uint32_t X = bx0 + i % R2.nx, Y = by0 + i / R2.nx;
                R2.put(i, ld3(l2, (size_t)Y * w2 + X) +
                              up_at([&](uint32_t a, uint32_t b) { return ld3(l3, (size_t)b * w3 + a); }, w3, h3, X, Y));
            }
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < n1; i += 256) {
                const uint32_t X = ax0 + i % R1.nx, Y = ay0 + i / R1.nx;
                R1.put(i, ld3(l1, (size_t)Y * w1 + X) + up_at([&](uint32_t a, uint32_t b) { return R2.at(a, b); }, w2, h2, X, Y));
            }
        } else {
            for (uint32_t i = threadIdx.x; i < n1; i += 256) {
                const uint32_t X = ax0 + i % R1.nx, Y = ay0 + i / R1.nx;
                R1.put(i, ld3(l1, (size_t)Y * w1 + X) +
                              up_at([&](uint32_t a, uint32_t b) { return ld3(l2, (size_t)b * w2 + a); }, w2, h2, X, Y));
            }
        }
        __syncthreads();
    }
    const uint32_t x = o.tx0 + (threadIdx.x & 15), y = o.ty0 + (threadIdx.x >> 4);
    if (x >= w || y >= h) return;
    uint32_t c;
    if constexpr (NUP >= 1)
        c = composite_at(hdr, [&](uint32_t a, uint32_t b) { return R1.at(a, b); }, w, w1, h1, x, y, lut);
    else
        c = composite_at(hdr, [&](uint32_t a, uint32_t b) { return ld3(l1, (size_t)b * w1 + a); }, w, w1, h1, x, y, lut);
    out[V.pix_start[o.vi] + (size_t)y * w + x] = c;
}
template __global__ void k_bloom_down2<false>(Views, int, float4 *);
template __global__ void k_bloom_down2<true>(Views, int, float4 *);
template __global__ void k_bloom_final<0>(Views, const float4 *, const uint8_t *, uint32_t *);
template __global__ void k_bloom_final<1>(Views, const float4 *, const uint8_t *, uint32_t *);
template __global__ void k_bloom_final<2>(Views, const float4 *, const uint8_t *, uint32_t *);

}  // namespace dev
