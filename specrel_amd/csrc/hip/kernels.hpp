// Kernel declarations and launch arguments shared by kernels.hip and capi.hip.
#pragma once
#include "device.hpp"

namespace dev {

constexpr int TILE = 16;          // pixels per side of a block's tile (4 wave64 of 8x8)
constexpr int SUPER = 64;         // pixels per side of a first-level star-culling tile
constexpr uint32_t SUPER_CAP = 4096;
constexpr uint32_t TILE_CAP = 1024;
constexpr int STACK_DEFAULT = 16; // traversal stack: > 0 positions in LDS, < 0 entries in registers

struct TraceArgs {
    Scene S;
    const specrel_frame *frame;
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
    const specrel_frame *frame;
    const StarPrep *stars;
    const StarDir *dirs;
    const uint32_t *super_lists;
    const uint32_t *super_counts;
    uint32_t num_stars;
    int cull;
    int want_hits;
    const uint32_t *slot;
    float4 *hdr;
    uint32_t *hit;
    float *lambda;
    float *tau;
    uint32_t *mask;         // per-pixel occlusion bits (bit 0 light, 1 + i flash i)
    ShadowQuery *queue;
    uint32_t *qcount;
    uint32_t qcap;
    unsigned long long *stats;
};

__global__ void k_star_prep(const float4 *words, uint32_t num_stars, const specrel_frame *frame, StarPrep *out,
                            StarDir *dirs);
__global__ void k_star_bin(const specrel_frame *frame, const StarDir *dirs, uint32_t num_stars, uint32_t *lists,
                           uint32_t *counts);
template <bool STATS, int SK>
__global__ void k_trace(TraceArgs a);
template <bool STATS, int SK>
__global__ void k_trace_p(TraceArgs a);
template <bool STATS, int SK>
__global__ void k_occlude_p(OccludeArgs a);
template <bool STATS, int SK, int PASS>
__global__ void k_shade(ShadeArgs a);
__global__ void k_reset(uint32_t *counters, uint32_t n);
__global__ void k_bloom_down(const float4 *src, uint32_t sw, uint32_t sh, float4 *dst, uint32_t dw, uint32_t dh);
__global__ void k_bloom_up(float4 *fine, uint32_t w, uint32_t h, const float4 *coarse, uint32_t cw, uint32_t ch);
__global__ void k_composite(const float4 *hdr, uint32_t w, uint32_t h, const float4 *l1, uint32_t w1, uint32_t h1,
                            const uint8_t *lut, uint32_t *out);

}  // namespace dev
