// Kernel declarations and launch arguments shared by kernels.hip and capi.hip.
#pragma once
#include "device.hpp"

namespace dev {

constexpr int TILE = 16;          // pixels per side of a block's tile (4 wave64 of 8x8)
constexpr int SUPER = 64;         // pixels per side of a first-level star-culling tile
constexpr uint32_t SUPER_CAP = 4096;
constexpr uint32_t TILE_CAP = 1024;
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

struct TraceArgs {
    Scene S;
    Views V;
    uint32_t num_tiles;
    uint32_t *slot;
    uint32_t *queue;  // persistent variant: ray-queue counter (zeroed per frame)
    unsigned long long *stats;
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

struct ShadeArgs {
    Scene S;
    Tables T;
    Views V;
    const StarPrep *stars;  // num_stars per view
    const StarDir *dirs;    // num_stars per view
    const uint32_t *super_lists;
    const uint32_t *super_counts;
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

__global__ void k_star_prep(const float4 *words, uint32_t num_stars, Views V, StarPrep *out, StarDir *dirs);
__global__ void k_star_bin(Views V, const StarDir *dirs, uint32_t num_stars, uint32_t *lists, uint32_t *counts);
template <bool STATS, int SK>
__global__ void k_trace(TraceArgs a);
template <bool STATS, int SK>
__global__ void k_trace_p(TraceArgs a);
template <bool STATS, int SK>
__global__ void k_occlude_p(OccludeArgs a);
template <bool STATS, int SK, int PASS>
__global__ void k_shade(ShadeArgs a);
__global__ void k_reset(uint32_t *counters, uint32_t n);
__global__ void k_bloom_down(Views V, int level, float4 *pyr);
__global__ void k_bloom_up(Views V, int level, float4 *pyr);
__global__ void k_composite(Views V, const float4 *pyr, const uint8_t *lut, uint32_t *out);

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
