// Device-side extrusion: argument block and kernels (extrude.hip), launched by
// specrel_extrude_tubes (capi.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "../core/extrude.h"
#include "specrel_types.h"

namespace dev {

struct ExTube {
    uint64_t prism_first;  // first prism of the tube in the launch
    uint32_t tri_first, n_tris, sample_first, n_samples;
    uint32_t decided_first;  // ~0u: DECIDED_AT_STATIC
    uint32_t material;
};

struct ExArgs {
    const ExTube *tubes;
    uint32_t n_tubes;
    uint64_t n_prisms;
    const sr::TubeTri *tris;
    const sr::Sample *samples;
    const float *decided;
};

__global__ void k_extrude_count(ExArgs a, uint32_t *block_counts);
__global__ void k_extrude_direct(ExArgs a, uint32_t *block_counts, specrel_geom *geoms, specrel_attr *attrs,
                                 float4 *boxes);
__global__ void k_extrude_scan(uint32_t *counts, uint32_t n, uint64_t *offsets);
__global__ void k_extrude_emit(ExArgs a, const uint64_t *offsets, specrel_geom *geoms, specrel_attr *attrs,
                               float4 *boxes, uint64_t *tube_first);

}  // namespace dev
