// Worldtube extrusion on the device (extend_worldtube_simplices, mesh4d.rs:721-854): one thread
// per prism (a triangle swept over one worldline step), the prism's three tetrahedra computed by
// the same source the scene producer runs on the host (core/extrude.h), so the records are
// bit-identical. A first pass writes each tetrahedron's record at its slot 3p + k and counts the
// proper ones; when a degenerate tetrahedron was dropped, a counting pass per block of prisms, a
// scan of those counts and an emitting pass write every record again at its index after the
// drops (in order).
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

// Proper tetrahedra per block of 256 prisms (the compaction's first pass).
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

// Optimistic single pass, one thread per tetrahedron slot q = 3p + k: the record goes to slot q
// and each block counts its proper ones. When every tetrahedron is proper (the usual case) the
// slots are the final indices and no compaction is needed. The block's records are staged in LDS
// and stored as contiguous runs (one component array at a time), so every store instruction
// writes whole cache lines.
__global__ __launch_bounds__(256) void k_extrude_direct(ExArgs a, uint32_t *block_counts, specrel_geom *geoms,
                                                        specrel_attr *attrs, float4 *boxes) {
    const uint64_t slot0 = (uint64_t)blockIdx.x * 256, q = slot0 + threadIdx.x, n_slots = 3 * a.n_prisms;
    const uint32_t tid = threadIdx.x;
    bool ok = false;
    specrel_geom g;
    specrel_attr at;
    sr::Box4 box;
    if (q < n_slots) {
        const uint64_t p = q / 3;
        const int k = (int)(q - 3 * p);
        const uint32_t tube = find_tube(a.tubes, a.n_tubes, p);
        const ExTube &t = a.tubes[tube];
        const uint64_t local = p - t.prism_first;
        const uint32_t step = (uint32_t)(local / t.n_tris), tri = (uint32_t)(local % t.n_tris);
        const sr::Sample sc = a.samples[t.sample_first + step], sn = a.samples[t.sample_first + step + 1];
        const float dec0 = t.decided_first == ~0u ? sr::DECIDED_AT_STATIC : a.decided[t.decided_first + step];
        const float dec1 = t.decided_first == ~0u ? sr::DECIDED_AT_STATIC : a.decided[t.decided_first + step + 1];
        const sr::TubeTri &tt = a.tris[t.tri_first + tri];
        const sr::PrismShared sh = sr::prism_shared(tt, sc, sn);
        ok = sr::prism_tet(sh, tt, dec0, dec1, t.material, k, g, at, box);
    }
    __shared__ uint32_t s_wave[4];
    uint32_t total;
    block_prefix(ok ? 1u : 0u, s_wave, total);
    if (tid == 0) block_counts[blockIdx.x] = total;
    // Staged stores: geoms (5 float4 per record), attrs (12), boxes (2).
    __shared__ float4 s_buf[256 * 12];
    const uint32_t n_here = (uint32_t)min((uint64_t)256, n_slots - slot0);
    const float4 *gv = reinterpret_cast<const float4 *>(&g);
    const float4 *av = reinterpret_cast<const float4 *>(&at);
#pragma unroll
    for (int j = 0; j < 5; ++j) s_buf[tid * 5 + j] = gv[j];
    __syncthreads();
    float4 *go = reinterpret_cast<float4 *>(geoms) + slot0 * 5;
    for (uint32_t i = tid; i < n_here * 5; i += 256) go[i] = s_buf[i];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 12; ++j) s_buf[tid * 12 + j] = av[j];
    __syncthreads();
    float4 *ao = reinterpret_cast<float4 *>(attrs) + slot0 * 12;
    for (uint32_t i = tid; i < n_here * 12; i += 256) ao[i] = s_buf[i];
    __syncthreads();
    s_buf[tid * 2] = make_float4(box.mn.t, box.mn.x, box.mn.y, box.mn.z);
    s_buf[tid * 2 + 1] = make_float4(box.mx.t, box.mx.x, box.mx.y, box.mx.z);
    __syncthreads();
    float4 *bo = boxes + slot0 * 2;
    for (uint32_t i = tid; i < n_here * 2; i += 256) bo[i] = s_buf[i];
}

}  // namespace dev
