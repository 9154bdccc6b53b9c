// Worldtube extrusion on the device (extend_worldtube_simplices, mesh4d.rs:721-854): one thread
// per prism (a triangle swept over one worldline step), the prism's three tetrahedra computed by
// the same source the scene producer runs on the host (core/extrude.h), so the records are
// bit-identical. Degenerate tetrahedra are dropped in order: a counting pass, a scan of the
// per-block counts, and an emitting pass that writes each record at its final index.
#include <hip/hip_runtime.h>

#include "extrude_dev.hpp"

namespace dev {

// Tube of a prism: the last tube whose first prism is <= p (tubes with no prisms are skipped
// because their prism_first equals the next tube's).
__device__ __forceinline__ uint32_t find_tube(const ExTube *tubes, uint32_t n, uint64_t p) {
    uint32_t lo = 0, hi = n;  // invariant: tubes[lo].prism_first <= p < tubes[hi].prism_first
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (tubes[mid].prism_first <= p) lo = mid;
        else hi = mid;
    }
    return lo;
}

// The prism p: its tube and index in the tube; sink(ok, geom, attr, box) per tetrahedron.
template <class Sink>
__device__ __forceinline__ void prism_of(const ExArgs &a, uint64_t p, uint32_t &tube, uint64_t &local, Sink &&sink) {
    tube = find_tube(a.tubes, a.n_tubes, p);
    const ExTube &t = a.tubes[tube];
    local = p - t.prism_first;
    const uint32_t step = (uint32_t)(local / t.n_tris), tri = (uint32_t)(local % t.n_tris);
    const sr::Sample sc = a.samples[t.sample_first + step], sn = a.samples[t.sample_first + step + 1];
    const float dec0 = t.decided_first == ~0u ? sr::DECIDED_AT_STATIC : a.decided[t.decided_first + step];
    const float dec1 = t.decided_first == ~0u ? sr::DECIDED_AT_STATIC : a.decided[t.decided_first + step + 1];
    sr::prism(a.tris[t.tri_first + tri], sc, sn, dec0, dec1, t.material, sink);
}

// Proper tetrahedra of prism p (0..3).
__device__ __forceinline__ uint32_t prism_count(const ExArgs &a, uint64_t p, uint32_t &tube, uint64_t &local) {
    uint32_t c = 0;
    prism_of(a, p, tube, local, [&](bool ok, const specrel_geom &, const specrel_attr &, const sr::Box4 &) {
        c += ok ? 1u : 0u;
    });
    return c;
}

// Exclusive prefix of v over the block's threads in thread order (256 threads, 4 waves).
__device__ __forceinline__ uint32_t block_prefix(uint32_t v, uint32_t *s_wave, uint32_t &total) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d);
        if (lane >= (uint32_t)d) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; ++w) before += s_wave[w];
    total = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    return before + x - v;
}

__global__ __launch_bounds__(256) void k_extrude_count(ExArgs a, uint32_t *block_counts) {
    const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    uint32_t c = 0, tube;
    uint64_t local;
    if (p < a.n_prisms) c = prism_count(a, p, tube, local);
    __shared__ uint32_t s_wave[4];
    uint32_t total;
    block_prefix(c, s_wave, total);
    if (threadIdx.x == 0) block_counts[blockIdx.x] = total;
}

// Exclusive scan of the n block counts into offsets, one block of 1024 threads; offsets[n]
// receives the total.
__global__ __launch_bounds__(1024) void k_extrude_scan(uint32_t *counts, uint32_t n, uint64_t *offsets) {
    __shared__ uint64_t s[1024];
    const uint32_t tid = threadIdx.x, per = (n + 1023) / 1024;
    const uint32_t lo = min(n, tid * per), hi = min(n, lo + per);
    uint64_t sum = 0;
    for (uint32_t i = lo; i < hi; ++i) sum += counts[i];
    s[tid] = sum;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {  // Hillis-Steele over the 1024 partial sums
        uint64_t y = tid >= d ? s[tid - d] : 0;
        __syncthreads();
        s[tid] += y;
        __syncthreads();
    }
    uint64_t run = s[tid] - sum;
    for (uint32_t i = lo; i < hi; ++i) {
        offsets[i] = run;
        run += counts[i];
    }
    if (tid == 1023) offsets[n] = s[1023];
}

__global__ __launch_bounds__(256) void k_extrude_emit(ExArgs a, const uint64_t *offsets, specrel_geom *geoms,
                                                      specrel_attr *attrs, float4 *boxes, uint64_t *tube_first) {
    const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    // Count first (the block prefix needs every thread's count), then recompute and write each
    // record straight to its final index.
    uint32_t c = 0, tube = 0;
    uint64_t local = 1;
    if (p < a.n_prisms) c = prism_count(a, p, tube, local);
    __shared__ uint32_t s_wave[4];
    uint32_t total;
    uint64_t o = offsets[blockIdx.x] + block_prefix(c, s_wave, total);
    if (p >= a.n_prisms) return;
    if (local == 0) tube_first[tube] = o;
    if (c == 0) return;
    prism_of(a, p, tube, local, [&](bool ok, const specrel_geom &g, const specrel_attr &at, const sr::Box4 &box) {
        if (!ok) return;
        geoms[o] = g;
        attrs[o] = at;
        boxes[2 * o] = make_float4(box.mn.t, box.mn.x, box.mn.y, box.mn.z);
        boxes[2 * o + 1] = make_float4(box.mx.t, box.mx.x, box.mx.y, box.mx.z);
        ++o;
    });
}

}  // namespace dev
