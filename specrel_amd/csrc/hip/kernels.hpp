// Kernel declarations and launch arguments shared by kernels.hip and capi.hip.
#pragma once
#include "device.hpp"

namespace dev {

constexpr int TILE = 16;          // pixels per side of a block's tile (4 wave64 of 8x8)
constexpr int SUPER = 64;         // pixels per side of a first-level star-culling tile
constexpr uint32_t SUPER_CAP = 4096;
constexpr uint32_t TILE_CAP = 1024;
// Tile ordering: one block sorts up to TILE_ORDER_ONE_BLOCK tiles; larger grids use blocks of
// TILE_ORDER_CHUNK tiles and the per-frame counters below.
constexpr uint32_t TILE_ORDER_ONE_BLOCK = 4096;
constexpr uint32_t TILE_ORDER_CHUNK = 1024;
// Per-frame counters (cleared by k_reset): 0 primary-ray queue, 1 shadow-query count, 2 next
// shadow query, 16..79 tile-order bucket totals, 80..143 tile-order bucket cursors.
constexpr uint32_t COUNTER_TILE_TOTAL = 16, COUNTER_TILE_CURSOR = 80, COUNTERS = 144;
constexpr int STACK_DEFAULT = 16; // traversal stack: > 0 positions in LDS, < 0 entries in registers

// The views of one launch (a single render is a batch of one view). Per-view ranges are prefix
// sums with n + 1 entries: level_start[k] counts 16x16 tiles of pyramid level k (level 0: the
// image tiles every pass but bloom works on), super_start 64x64 star super tiles, pix_start
// pixels (the offset of the view in every per-pixel buffer), pyr_start float4s of the view's
// HDR pyramid. The host keeps every total below 2^32.
struct Views {
    const specrel_frame *frames;
    const uint32_t *level_start[SPECREL_BLOOM_LEVELS + 1];
    const uint32_t *super_start;
    const uint32_t *pix_start;
    const uint32_t *pyr_start;
    uint32_t n;
};

// The view owning global item t of a prefix table (start[0] = 0 <= t < start[n]).
__device__ __forceinline__ uint32_t find_view(const uint32_t *start, uint32_t n, uint32_t t) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (start[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

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

// The pixel's ray direction (renderer.rs:369-384 with Tetrad::past/future_light_ray_direction,
// local_spatial_to_world; minkowski.rs:526-575).
__device__ __forceinline__ V4 ray_dir(const ViewDerived &v, const specrel_frame &f, uint32_t px, uint32_t py) {
    float sx = pix_sx(v, px), sy = pix_sy(v, py);
    V3 local_dir = normalized(V3{sx, sy, 1.0f});
    if (f.slice_mode == SPECREL_VIEW_PAST_CONE) {
        V3 n = normalized(local_dir);
        return -v.e0 + v.e1 * n.x + v.e2 * n.y + v.e3 * n.z;
    }
    if (f.slice_mode == SPECREL_VIEW_FRONTIER) {
        V3 n = normalized(local_dir);
        return v.e0 + v.e1 * n.x + v.e2 * n.y + v.e3 * n.z;
    }
    return v.e1 * local_dir.x + v.e2 * local_dir.y + v.e3 * local_dir.z;
}

// Traversal stack of a kernel: SK > 0 keeps SK positions per lane in LDS, SK < 0 keeps the top
// -SK entries in registers.
// Kernels declare `typename StackSel<SK>::Private priv;` next to the stack and pass it to init.
// LANES is the block size (the LDS array holds one column per lane).
template <int SK, int LANES = 256, bool = (SK > 0)>
struct StackSel {
    using type = LdsStack<SK>;
    static constexpr int lds_entries = SK * LANES;
    struct Private {
        uint2 deep[type::DEEP];
    };
    __device__ static void init(type &s, uint2 *lds, uint32_t tid, Private &p) {
        s.lds = (lds_uint2 *)(lds + tid);  // lds is a __shared__ array
        s.stride = LANES;
        s.n = 0;
        s.deep = p.deep;
    }
};
template <int SK, int LANES>
struct StackSel<SK, LANES, false> {
    using type = RegStack<-SK>;
    static constexpr int lds_entries = 1;
    struct Private {};
    __device__ static void init(type &s, uint2 *, uint32_t, Private &) { s.n = 0; }
};

struct ShadeArgs {
    Scene S;
    Tables T;
    Views V;
    const StarPrep *stars;  // num_stars per view
    const StarDir *dirs;    // num_stars per view
    const uint32_t *super_lists;
    const uint32_t *super_counts;
    const float4 *tile_cones;  // per tile: star cone axis and cosine bound (k_star_bin)
    uint32_t num_stars;
    int cull;
    int want_hits;
    const uint32_t *slot;
    float4 *hdr;  // pyramids (pyr_start); the rest are per pixel (pix_start)
    uint32_t *hit;
    float *lambda;
    float *tau;
    uint32_t *mask;         // per-pixel occlusion bits (bit 0 light, 1 + i flash i)
    ShadowQuery *queue;
    uint32_t *qcount;
    uint32_t qcap;
    unsigned long long *stats;
};

struct TraceArgs {
    Scene S;
    Views V;
    uint32_t num_tiles;
    uint32_t *slot;
    uint32_t *mask;   // with shadows traced here: per-pixel occlusion bits (bit 0 light, 1 + i flash i)
    uint32_t *queue;  // persistent variant: ray-queue counter (zeroed per frame)
    unsigned long long *stats;
    uint32_t wave_shape;  // pixels of a tile one wave takes: 0 an 8x8 quarter, 1 a 16x4 band, 2 a 4x16 band
    const uint32_t *order;  // tile of each scheduling slot (nullptr: slot = tile), heaviest first
    uint32_t *cost;         // per tile quarter: the longest traversal of its lanes, for the next order
    uint32_t prio_slots;    // tiles in the first slots of the order run at a raised wave priority
    uint32_t postpone;      // 0, or uninstrumented closest-hit rays postpone their leaves (trace_pp) and
                            // are traced again by trace() past this many stack entries
    specrel_wave_clock *clock = nullptr;  // or one record per wave (scheduling slot): start, end, where
    ShadeArgs shade;  // the shading pass's arguments, for the kernel that shades its own pixels
    bool shade_here = false;  // (host side) launch the kernel that shades its own pixels
};

struct OccludeArgs {
    Scene S;
    const ShadowQuery *queue;
    const uint32_t *count;  // queries enqueued
    uint32_t *next;         // grab counter (zeroed per frame)
    uint32_t cap;
    uint32_t *mask;
    unsigned long long *stats;
};


__global__ void k_star_prep(const float4 *words, uint32_t num_stars, Views V, StarPrep *out, StarDir *dirs);
__global__ void k_star_bin(Views V, const StarDir *dirs, uint32_t num_stars, uint32_t *lists, uint32_t *counts,
                           float4 *cones, const uint32_t *order_cost, uint32_t order_tiles, uint32_t *order);
// Adjoint of the shaded image (adjoint.hip): per-pixel floats, camera sums, pixel classes.
constexpr int ADJ_PIX = 12;     // d origin[4], d direction[4], d object boost[3], loss term
constexpr int ADJ_CAM = 21;     // loss, d event[4], d e0..e3[16]
constexpr int ADJ_CLASSES = 6;  // valid, miss, starlight, border, silhouette, shading edge
struct AdjArgs {
    Scene S;
    Tables T;
    Views V;
    const uint32_t *slot;
    const float4 *hdr;    // the pyramids (pyr_start); every other per-pixel array by pix_start
    const float *target;  // 3 floats per pixel
    const uint8_t *mask_in;
    uint8_t *mask_out;
    const uint32_t *obj_first;  // objects sorted by first input index, disjoint
    const uint32_t *obj_count;
    const float4 *obj_center;
    uint32_t n_obj;
    double *pix;     // ADJ_PIX per pixel
    uint32_t *sig;   // per pixel
    uint32_t *obj;   // per pixel
    uint8_t *inside; // per pixel: the hit lies inside its tetrahedron (every barycentric > 0)
    uint32_t strict; // validity: 0 silhouettes masked; 1 also every shading edge (SPECREL_OPT_ADJOINT_MASK)
    double *partials;       // per tile: ADJ_CAM + 3 n_obj
    uint32_t *block_counts; // per tile: ADJ_CLASSES
    uint32_t stride;        // ADJ_CAM + 3 n_obj
    double *result;         // per view: stride
    unsigned long long *result_counts;  // per view: ADJ_CLASSES
};
template <int SK>
__global__ void k_adj_pixel(AdjArgs a);
__global__ void k_adj_reduce(AdjArgs a);
__global__ void k_adj_final(AdjArgs a);

// TV: traversal variant (TV_PLAIN, TV_MERGED: merged load rounds for a wave's node and leaf
// lanes, TV_WIDE4: fetch width 4, TV_LDSTOP: top levels from LDS); SH: 1 the hit's shadow queries
// traced here, 2 also the pixels shaded here; W: waves per block (each wave one quarter of a
// 16x16 tile), 1, 2 or 4.
template <bool STATS, int SK, int TV, int SH, int W = 4>
__global__ void k_trace(TraceArgs a);
template <bool STATS, int SK>
__global__ void k_trace_p(TraceArgs a);
template <bool STATS, int SK>
__global__ void k_occlude_p(OccludeArgs a);
template <bool STATS, int SK, int PASS>
__global__ void k_shade(ShadeArgs a);
__global__ void k_popcount(const uint32_t *bits, uint64_t n, unsigned long long *out);
__global__ void k_patch_rows(const uint32_t *idx, const float4 *rows, const uint2 *info, uint32_t n, float4 *nbox,
                             uint2 *ninfo);
__global__ void k_reset(uint32_t *counters, uint32_t n);
struct FrameWords {
    static constexpr uint32_t N = sizeof(specrel_frame) / 4;
    uint32_t w[N];
};
static_assert(sizeof(specrel_frame) % 4 == 0 && FrameWords::N <= 256, "one frame per k_reset_frame thread block");
__global__ void k_reset_frame(uint32_t *counters, uint32_t n, FrameWords f, uint32_t *dst);
__global__ void k_star_prep_first(const float4 *words, uint32_t num_stars, StarPrep *out, StarDir *dirs,
                                  uint32_t *counters, uint32_t n, FrameWords fw, uint32_t *dst);
// The view table of one view (level tile prefix sums, super tiles, pixels, pyramid floats: each
// row's two entries), passed by value to k_reset_view.
constexpr int TABLE_ROWS = SPECREL_BLOOM_LEVELS + 4;
struct TableWords {
    uint32_t w[2 * TABLE_ROWS];
};
__global__ void k_reset_view(uint32_t *counters, uint32_t n, FrameWords f, uint32_t *dst_frame, TableWords t,
                             uint32_t *dst_table);
__global__ void k_tile_order(const uint32_t *cost, uint32_t n_tiles, uint32_t *order);
__global__ void k_tile_hist(const uint32_t *cost, uint32_t n_tiles, uint32_t *total);
__global__ void k_tile_scatter(const uint32_t *cost, uint32_t n_tiles, const uint32_t *total, uint32_t *cursor,
                               uint32_t *order);
__global__ void k_tile_init(uint32_t *order, uint32_t *cost, uint32_t n_tiles);
__global__ void k_patch_nodes(const uint32_t *pslot, const float4 *bounds, uint32_t n, float4 *nbox);
__global__ void k_bloom_down(Views V, int level, float4 *pyr);
__global__ void k_bloom_up(Views V, int level, float4 *pyr);
__global__ void k_bloom_tail(Views V, int k0, float4 *pyr);
// The bloom levels from which one block per view finishes the pyramid (k_bloom_tail): every
// view's level holds at most this many pixels.
constexpr uint32_t BLOOM_TAIL_PIXELS = 2560;
__global__ void k_composite(Views V, const float4 *pyr, const uint8_t *lut, uint32_t *out);
// Fused bloom passes (one block per 16x16 tile of the level named): levels k and k + 1 of the down
// pass; levels k + 1 and k of the up pass; the last NUP (0..2) levels of the up pass and the composite.
template <bool STAGE>
__global__ void k_bloom_down2(Views V, int k, float4 *pyr);
constexpr uint32_t BLOOM_STAGE_TILES = 4096;  // k_bloom_down2 stages its taps in LDS on grids above this
__global__ void k_bloom_up2(Views V, int k, float4 *pyr);
template <int NUP>
__global__ void k_bloom_final(Views V, const float4 *pyr, const uint8_t *lut, uint32_t *out);

// Pyramid geometry of bloom.rs:28-49, shared by host and device.
__host__ __device__ __forceinline__ void level_size(uint32_t w, uint32_t h, int k, uint32_t &ow, uint32_t &oh) {
    for (int i = 0; i < k; ++i) {
        w = (w + 1) / 2;
        h = (h + 1) / 2;
    }
    ow = w;
    oh = h;
}
__host__ __device__ __forceinline__ size_t level_offset(uint32_t w, uint32_t h, int k) {
    size_t off = 0;
    for (int i = 0; i < k; ++i) {
        off += (size_t)w * h;
        w = (w + 1) / 2;
        h = (h + 1) / 2;
    }
    return off;
}
__host__ __device__ __forceinline__ size_t pyramid_len(uint32_t w, uint32_t h) {
    return level_offset(w, h, SPECREL_BLOOM_LEVELS + 1);
}

}  // namespace dev
