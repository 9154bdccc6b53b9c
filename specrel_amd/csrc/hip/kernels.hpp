// Kernel declarations and launch arguments shared by kernels.hip and capi.hip.
#pragma once
#include "device.hpp"

namespace dev {

struct RenderArgs {
    Scene S;
    Tables T;
    const specrel_frame *frame;
    const StarPrep *stars;
    uint32_t num_stars;
    int cull;
    int want_hits;
    float4 *hdr;
    uint32_t *hit;
    float *lambda;
    float *tau;
    unsigned long long *stats;
};

__global__ void k_star_prep(const float4 *words, uint32_t num_stars, const specrel_frame *frame, StarPrep *out);
template <bool STATS>
__global__ void k_render(RenderArgs a);
__global__ void k_bloom_down(const float4 *src, uint32_t sw, uint32_t sh, float4 *dst, uint32_t dw, uint32_t dh);
__global__ void k_bloom_up(float4 *fine, uint32_t w, uint32_t h, const float4 *coarse, uint32_t cw, uint32_t ch);
__global__ void k_composite(const float4 *hdr, uint32_t w, uint32_t h, const float4 *l1, uint32_t w1, uint32_t h1,
                            const uint8_t *lut, uint32_t *out);

}  // namespace dev
