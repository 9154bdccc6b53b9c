// Device-side renderer for gfx950: ray generation, 4D BVH traversal, affine ray-tetrahedron
// intersection, proper-time shading, null / camera-slice shadow rays and star light.
//
// Semantics follow the reference CPU renderer exactly (src/renderer.rs, src/bvh4d.rs,
// src/mesh4d.rs, src/star.rs) so that RGB8, hit index, lambda and tau are bit-identical;
// the code is compiled with -ffp-contract=off and without fast-math. The WGSL port's
// deviations (render.wgsl: dividing camera beta, truncated `%`, round_away) are not followed.
//
// Memory layout on the device (built by the upload):
//   nbox[i]  : 4 x float4 = left child min, left max, right min, right max, for each interior
//              node i (reference node numbering); 64 B, 64-B aligned.
//   ninfo[i] : uint2 = left entry, right entry (see encode below).
//   geoms    : 5 x float4 per leaf slot (GpuGeom, 80 B), attrs 12 x float4 (GpuAttr, 192 B).
// A stack entry is (entry, t_enter); an entry names either an interior node or a leaf with
// its first slot and count, so popping a leaf needs no node load.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../core/minkowski.h"
#include "specrel_types.h"

namespace dev {
using namespace sr;

constexpr uint32_t LEAF = 0x80000000u;
constexpr uint32_t COUNT_SHIFT = 26;
constexpr uint32_t FIRST_MASK = (1u << COUNT_SHIFT) - 1u;  // small leaf: first < 2^26
constexpr uint32_t COUNT_MAX = 31u;                         // small leaf: count in 5 bits
// An interior entry with TOP_BIT names a node of the block's LDS copy of the top levels by its
// position there (uploads keep node indices below 2^30).
constexpr uint32_t TOP_BIT = 0x40000000u;
constexpr int TOP_LEVELS_MAX = 8;
constexpr uint32_t TOP_MAX = (1u << TOP_LEVELS_MAX) - 1u;  // interior nodes of 8 levels
// Traversal variants of the primary-ray kernel (the visited sequence is the same in all):
//   TV_PLAIN  one round of loads per node step;
//   TV_MERGED a step merges the loads of a wave's node lanes and leaf lanes (closest-hit trace());
//   TV_WIDE4  fetch width 4: a node step also fetches both children's box pairs, so a continued
//             child's boxes are in registers at the next step (postponed-leaf traversal);
//   TV_LDSTOP the top levels of the tree are read from a per-block copy in LDS.
constexpr int TV_PLAIN = 0, TV_MERGED = 1, TV_WIDE4 = 2, TV_LDSTOP = 3;
// Big leaf: LEAF | index into big_leaves (count field 0).

struct Scene {
    const float4 *nbox;
    const uint2 *ninfo;
    const uint2 *big_leaves;
    const float4 *geoms;
    const float4 *attrs;
    const uint32_t *order;
    const float4 *materials;
    uint32_t root_entry;
    uint32_t num_prims;
    uint32_t empty;  // no nodes: every ray misses
    float4 root_min, root_max;
    float4 light;  // (0, x, y, z)
    // Instrumented renders only (NULL otherwise): one bit per interior node whose box pair was
    // loaded, per primitive record tested and per primitive whose attributes a hit read, so the
    // distinct records a frame touches can be counted.
    uint32_t *touch_nodes, *touch_prims, *touch_hits;
    // The interior nodes of the tree's top levels in breadth-first order (node index, and child
    // entries naming top nodes as TOP_BIT | position), which a block of the primary-ray kernel
    // may copy into LDS: the first top_n of them are used, entered from top_root.
    const uint32_t *top_idx;
    const uint2 *top_info;
    uint32_t top_n, top_root;
};

__device__ __forceinline__ void touch(uint32_t *bits, uint32_t i) {
    if (bits) atomicOr(bits + (i >> 5), 1u << (i & 31u));
}

struct Tables {
    const uint8_t *lut;     // 4096
    const uint8_t *glyphs;  // 14 x 7
};

// Per-star values that depend only on the view (star.rs:251-283): r = d - tau e0, the
// blackbody colour at the Doppler-shifted temperature, the colour norm, and the direction
// of r in the camera's rest frame (for tile culling).
struct StarPrep {
    float4 r;
    float4 bb_cn;      // blackbody rgb, colour_norm
    uint32_t flags;    // bit0: exists on the cone; bit1: non-finite, never culled
    uint32_t pad[3];
};
// Culling record of a star: unit direction of r in the camera frame, flags in w.
struct StarDir {
    float x, y, z;
    uint32_t flags;
};

SR_HD V4 f4v(float4 a) { return V4{a.x, a.y, a.z, a.w}; }
SR_HD V4 lv(const float *p) { return V4{p[0], p[1], p[2], p[3]}; }

struct Counters {
    uint64_t node_pops, box_tests, prim_tests, hits, shadow_rays, shadow_node_pops, shadow_box_tests,
        shadow_prim_tests, star_tests;
    uint32_t steps;  // traversal loop iterations of the lane (always counted: the tile's cost estimate)
};

// Aabb4D::intersect_ray (mesh4d.rs:113-146)
__device__ __forceinline__ bool box_hit(float4 mn, float4 mx, V4 o, V4 inv, float tmin, float tmax, float &enter) {
    float t0 = (mn.x - o.t) * inv.t, t1 = (mx.x - o.t) * inv.t;
    float x0 = (mn.y - o.x) * inv.x, x1 = (mx.y - o.x) * inv.x;
    float y0 = (mn.z - o.y) * inv.y, y1 = (mx.z - o.y) * inv.y;
    float z0 = (mn.w - o.z) * inv.z, z1 = (mx.w - o.z) * inv.z;
    float en = fmaxf(fmaxf(fmaxf(fmaxf(tmin, fminf(t0, t1)), fminf(x0, x1)), fminf(y0, y1)), fminf(z0, z1));
    float ex = fminf(fminf(fminf(fminf(tmax, fmaxf(t0, t1)), fmaxf(x0, x1)), fmaxf(y0, y1)), fmaxf(z0, z1));
    enter = en;
    return ex >= en;
}

// Bvh4D::safe_inv_dir (bvh4d.rs:430-441)
__device__ __forceinline__ float safe_inv(float v) {
    if (fabsf(v) > 1e-12f) return 1.0f / v;
    return v >= 0.0f ? 1e12f : -1e12f;
}

// SimplexGeom4D::intersect_fast (mesh4d.rs:175-211), in two parts: the values (everything but the
// t range test), and the reference's early-out conditions on them, combined without short-circuit.
// A record's five vectors are loaded at once and every value is computed before the conditions are
// tested, so a test costs one round trip to memory instead of one per vector; the arithmetic, and
// with it every accepted hit, is the reference's.
struct GeomRec {
    float4 v0, n, b1, b2, b3;
};
__device__ __forceinline__ GeomRec load_geom(const float4 *g) { return GeomRec{g[0], g[1], g[2], g[3], g[4]}; }
struct Isect {
    float denom, lambda, u, v, w;
    V4 event;
};
__device__ __forceinline__ Isect isect_values(const GeomRec &g, V4 o, V4 d) {
    Isect x;
    V4 n = f4v(g.n);
    x.denom = edot(n, d);
    V4 v0 = f4v(g.v0);
    x.lambda = edot(n, v0 - o) / x.denom;
    x.event = o + d * x.lambda;
    V4 q = x.event - v0;
    x.u = edot(f4v(g.b1), q);
    x.v = edot(f4v(g.b2), q);
    x.w = edot(f4v(g.b3), q);
    return x;
}
__device__ __forceinline__ bool isect_accept(const Isect &x, float tmin, float tmax) {
    const float eps = -1e-5f;
    const bool reject = (fabsf(x.denom) < 1e-9f) | (x.lambda <= tmin) | (x.lambda >= tmax) | (x.u < eps) |
                        (x.u > 1.0f - eps) | (x.v < eps) | (x.u + x.v > 1.0f - eps) | (x.w < eps) |
                        (x.u + x.v + x.w > 1.0f - eps);
    return !reject;
}
__device__ __forceinline__ void isect_bary(const Isect &x, float bary[4]) {
    bary[0] = 1.0f - x.u - x.v - x.w;
    bary[1] = x.u;
    bary[2] = x.v;
    bary[3] = x.w;
}
__device__ __forceinline__ bool intersect(const float4 *g, V4 o, V4 d, float tmin, float tmax, float &lam,
                                          V4 &ev, float bary[4]) {
    const Isect x = isect_values(load_geom(g), o, d);
    if (!isect_accept(x, tmin, tmax)) return false;
    lam = x.lambda;
    ev = x.event;
    isect_bary(x, bary);
    return true;
}

__device__ __forceinline__ void leaf_range(const Scene &S, uint32_t e, uint32_t &first, uint32_t &count) {
    uint32_t c = (e >> COUNT_SHIFT) & COUNT_MAX;
    if (c != 0) {
        first = e & FIRST_MASK;
        count = c;
    } else {
        uint2 b = S.big_leaves[e & FIRST_MASK];
        first = b.x;
        count = b.y;
    }
}

struct Hit {
    int32_t slot;
    float lambda;
    V4 event;
    float bary[4];
};

constexpr int STACK = 64;  // the reference's stack size (bvh4d.rs:454, 537), overflow semantics kept

// The logical 64-entry traversal stack of one lane, two implementations with one interface:
//   LdsStack<K>: positions below K in LDS (slot-major, so the 64 lanes of a wave touch
//                consecutive words whatever their depth), deeper ones in private memory.
//   RegStack<N>: the top N entries in registers (static-index shifts, no memory on push/pop),
//                older ones spilled to private memory; a spilled entry is restored when it
//                re-enters the window, N pops before it is needed, so its load latency hides.
// A pointer into LDS as such (address space 3): accesses compile to ds_read/ds_write. Through a
// generic pointer, a choice between an LDS slot and a private slot becomes one flat access.
typedef __attribute__((address_space(3))) uint2 lds_uint2;

// The private part is an array the kernel declares (`deep`), so that the stack object itself
// holds only scalars and lives in registers; the LDS read is volatile so that it is not merged
// with the private read into one access through a generic pointer.
template <int K>
struct LdsStack {
    static constexpr int DEEP = STACK - K;
    lds_uint2 *lds;   // this lane's column: lds[pos * stride]
    uint32_t stride;  // lanes per block
    int n;
    uint2 *deep;      // DEEP entries in private memory
    __device__ __forceinline__ void reset() { n = 0; }
    __device__ __forceinline__ int size() const { return n; }
    // While no active lane of the wave reaches the private part (the common case), push and pop
    // take a wave-uniform branch straight to the LDS access.
    __device__ __forceinline__ void push(uint32_t e, float t) {
        if (__builtin_expect(__builtin_amdgcn_ballot_w64(n >= K) == 0, 1) || n < K) {
            lds_uint2 *p = lds + (uint32_t)n * stride;
            p->x = e;
            p->y = __float_as_uint(t);
        } else {
            deep[n - K] = make_uint2(e, __float_as_uint(t));
        }
        n += 1;
    }
    __device__ __forceinline__ void pop(uint32_t &e, float &t) {
        n -= 1;
        uint2 v;
        if (__builtin_expect(__builtin_amdgcn_ballot_w64(n >= K) == 0, 1) || n < K) {
            const volatile lds_uint2 *p = lds + (uint32_t)n * stride;
            v.x = p->x;
            v.y = p->y;
        } else {
            v = deep[n - K];
        }
        e = v.x;
        t = __uint_as_float(v.y);
    }
    // Branch-free forms for loops that keep the lane's state in registers: put() writes (e, t)
    // at position n without pushing it, grow(b) then pushes it when b; pop_if(b) pops when b and
    // leaves e, t as they were otherwise. Every lane touches a position that exists (n < STACK
    // for put, 0 for lanes that do not pop).
    __device__ __forceinline__ void put(uint32_t e, float t) {
        if (__builtin_expect(__builtin_amdgcn_ballot_w64(n >= K) == 0, 1)) {  // every lane in LDS
            lds_uint2 *p = lds + (uint32_t)n * stride;
            p->x = e;
            p->y = __float_as_uint(t);
        } else if (n < K) {
            lds_uint2 *p = lds + (uint32_t)n * stride;
            p->x = e;
            p->y = __float_as_uint(t);
        } else {
            deep[n - K] = make_uint2(e, __float_as_uint(t));
        }
    }
    __device__ __forceinline__ void grow(bool b) { n += b ? 1 : 0; }
    __device__ __forceinline__ void pop_if(bool b, uint32_t &e, float &t) {
        const int i = b ? n - 1 : 0;
        uint2 v;
        if (__builtin_expect(__builtin_amdgcn_ballot_w64(i >= K) == 0, 1)) {  // every lane in LDS
            const volatile lds_uint2 *p = lds + (uint32_t)i * stride;
            v.x = p->x;
            v.y = p->y;
        } else if (i < K) {
            const volatile lds_uint2 *p = lds + (uint32_t)i * stride;
            v.x = p->x;
            v.y = p->y;
        } else {
            v = deep[i - K];
        }
        e = b ? v.x : e;
        t = b ? __uint_as_float(v.y) : t;
        n -= b ? 1 : 0;
    }
};

template <int N>
struct RegStack {
    uint32_t e_[N];
    float t_[N];
    int n;
    uint2 deep[STACK - N];
    __device__ __forceinline__ void reset() { n = 0; }
    __device__ __forceinline__ int size() const { return n; }
    __device__ __forceinline__ void push(uint32_t e, float t) {
        if (n >= N) deep[n - N] = make_uint2(e_[N - 1], __float_as_uint(t_[N - 1]));
#pragma unroll
        for (int i = N - 1; i > 0; --i) {
            e_[i] = e_[i - 1];
            t_[i] = t_[i - 1];
        }
        e_[0] = e;
        t_[0] = t;
        n += 1;
    }
    __device__ __forceinline__ void pop(uint32_t &e, float &t) {
        e = e_[0];
        t = t_[0];
#pragma unroll
        for (int i = 0; i < N - 1; ++i) {
            e_[i] = e_[i + 1];
            t_[i] = t_[i + 1];
        }
        n -= 1;
        if (n >= N) {
            uint2 v = deep[n - N];
            e_[N - 1] = v.x;
            t_[N - 1] = __uint_as_float(v.y);
        }
    }
    uint32_t put_e_;
    float put_t_;
    __device__ __forceinline__ void put(uint32_t e, float t) { put_e_ = e, put_t_ = t; }
    __device__ __forceinline__ void grow(bool b) {
        if (b) push(put_e_, put_t_);
    }
    __device__ __forceinline__ void pop_if(bool b, uint32_t &e, float &t) {
        if (b) pop(e, t);
    }
};

// Bvh4D::trace (bvh4d.rs:445-520). The near child of a node whose children both hit is not
// pushed and popped again but traversed at once: the reference pops it right after pushing it,
// with nothing in between, so the visit order, the t_enter >= t_max skips and the stack-size
// limit (both children only when sp + 2 <= 64) are the same.
// An interior node's two child boxes and child entries (the 64 + 8 bytes one step loads).
struct NodeData {
    float4 lmn, lmx, rmn, rmx;
    uint2 info;
};
__device__ __forceinline__ NodeData load_node(const Scene &S, uint32_t e) {
    const float4 *bx = S.nbox + 4 * (size_t)e;
    return NodeData{bx[0], bx[1], bx[2], bx[3], S.ninfo[e]};
}

// The block's LDS copy of the top levels (TV_LDSTOP kernels only: a kernel that never names
// these arrays is given no LDS for them).
typedef __attribute__((address_space(3))) float4 lds_float4;
static __shared__ float4 g_top_box[4 * TOP_MAX];
static __shared__ uint2 g_top_info[TOP_MAX];
// Copies the first S.top_n top rows into LDS (every thread of the block takes part): the box
// pairs from the node array as they stand, the child entries with those naming top nodes past
// top_n turned back into node indices.
__device__ __forceinline__ void top_fill(const Scene &S) {
    const uint32_t n = S.top_n;
    for (uint32_t i = threadIdx.x; i < 5 * n; i += blockDim.x) {
        const uint32_t r = i / 5, part = i - 5 * r;
        if (part < 4) {
            g_top_box[4 * r + part] = S.nbox[4 * (size_t)S.top_idx[r] + part];
        } else {
            uint2 inf = S.top_info[r];
            if ((inf.x & (LEAF | TOP_BIT)) == TOP_BIT && (inf.x & ~TOP_BIT) >= n) inf.x = S.top_idx[inf.x & ~TOP_BIT];
            if ((inf.y & (LEAF | TOP_BIT)) == TOP_BIT && (inf.y & ~TOP_BIT) >= n) inf.y = S.top_idx[inf.y & ~TOP_BIT];
            g_top_info[r] = inf;
        }
    }
    __syncthreads();
}
template <int TV>
__device__ __forceinline__ uint32_t root_of(const Scene &S) {
    if constexpr (TV == TV_LDSTOP) return S.top_n ? S.top_root : S.root_entry;
    else return S.root_entry;
}
// An interior node's row: from LDS for a top entry of a TV_LDSTOP kernel, else from the node array.
template <int TV>
__device__ __forceinline__ NodeData load_node_v(const Scene &S, uint32_t e) {
    if constexpr (TV == TV_LDSTOP) {
        if (e & TOP_BIT) {
            const uint32_t i = e & ~TOP_BIT;
            const lds_float4 *b = (const lds_float4 *)g_top_box + 4 * i;
            const lds_uint2 *f = (const lds_uint2 *)g_top_info + i;
            auto rd = [](const lds_float4 &v) { return make_float4(v.x, v.y, v.z, v.w); };
            return NodeData{rd(b[0]), rd(b[1]), rd(b[2]), rd(b[3]), make_uint2(f->x, f->y)};
        }
    }
    return load_node(S, e);
}

// Marks a loaded value as needed here, so that the loads of one step are all issued before any
// of them is waited for (an empty asm statement: no instruction).
typedef float f32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void needed(float4 a) {
    const f32x4 v = {a.x, a.y, a.z, a.w};
    asm volatile("" ::"v"(v));
}
__device__ __forceinline__ void needed(uint2 a) { asm volatile("" ::"v"(a.x), "v"(a.y)); }
__device__ __forceinline__ void needed(const GeomRec &g) {
    needed(g.v0), needed(g.n), needed(g.b1), needed(g.b2), needed(g.b3);
}

// One leaf's primitives first .. end-1 from slot i on, tested in slot order, each with the t_max
// the previous ones left (x0..x2: slots i, i+1, i+2, those past the leaf untested).
template <bool STATS>
__device__ __forceinline__ void leaf_tests(const Isect &x0, const Isect &x1, const Isect &x2, uint32_t i,
                                           uint32_t end, float tmin, float &tmax, Hit &hit, Counters &C) {
    const Isect *xs[3] = {&x0, &x1, &x2};
#pragma unroll
    for (uint32_t k = 0; k < 3; ++k) {
        if (i + k >= end) break;
        if (STATS) C.prim_tests++;
        if (isect_accept(*xs[k], tmin, tmax)) {
            tmax = xs[k]->lambda;
            hit.slot = (int32_t)(i + k);
            hit.lambda = xs[k]->lambda;
            hit.event = xs[k]->event;
            isect_bary(*xs[k], hit.bary);
        }
    }
}

// Bvh4D::trace (bvh4d.rs:443-520). A leaf's primitives are loaded three at a time, all five vectors
// of a record in one round. With M, a step also merges the rounds of a wave's two kinds of lanes:
// the lanes at an interior node load its pair and the lanes at a leaf its first three primitives
// together (each kind only when some lane of the wave needs it; a lane of the other kind reads a
// valid record it ignores), and all of them are waited for at once, instead of the leaf lanes and
// the node lanes each waiting for their own loads in turn (measured slower on the large scene:
// the extra loads cost more than the round they save). The visiting order, and with it every
// result and counter, is the same either way.
template <bool STATS, class ST, bool M = false, int TV = TV_PLAIN>
__device__ bool trace(const Scene &S, V4 o, V4 d, float tmin, float tmax, Hit &hit, Counters &C, ST &stk) {
    static_assert(!(STATS && TV == TV_LDSTOP), "instrumented traversals name nodes by their index");
    hit.slot = -1;
    if (S.empty) return false;
    V4 inv = V4{safe_inv(d.t), safe_inv(d.x), safe_inv(d.y), safe_inv(d.z)};
    float t_enter;
    if (!box_hit(S.root_min, S.root_max, o, inv, tmin, tmax, t_enter)) return false;
    uint32_t e = root_of<TV>(S);
    stk.reset();
    bool have = true;
    while (true) {
        if (!have) {
            if (stk.size() == 0) break;
            stk.pop(e, t_enter);
        }
        have = false;
        C.steps++;
        if (t_enter >= tmax) continue;
        if (STATS) C.node_pops++;
        const bool leaf = (e & LEAF) != 0;
        uint32_t first = 0, count = 0;
        if (leaf) leaf_range(S, e, first, count);
        NodeData D;
        Isect x0, x1, x2;
        uint32_t done = 0;  // primitives of the leaf tested in the merged round
        if (M) {
            const bool prims = leaf && count > 0;
            const uint32_t last = first + count - 1;
            const uint32_t j0 = prims ? first : 0u, j1 = prims ? min(first + 1, last) : 0u,
                           j2 = prims ? min(first + 2, last) : 0u;
            const bool any_prims = __builtin_amdgcn_ballot_w64(prims) != 0;
            const bool any_nodes = __builtin_amdgcn_ballot_w64(!leaf) != 0;
            GeomRec g0{}, g1{}, g2{};
            D = NodeData{};
            if (any_nodes) D = load_node(S, leaf ? 0u : e);  // record 0 exists whenever a node does
            if (any_prims) {  // slot 0 exists whenever a leaf of the scene has a primitive
                g0 = load_geom(S.geoms + 5 * (size_t)j0);
                g1 = load_geom(S.geoms + 5 * (size_t)j1);
                g2 = load_geom(S.geoms + 5 * (size_t)j2);
            }
            needed(D.lmn), needed(D.lmx), needed(D.rmn), needed(D.rmx), needed(D.info);
            needed(g0), needed(g1), needed(g2);
            if (prims) {
                x0 = isect_values(g0, o, d);
                x1 = isect_values(g1, o, d);
                x2 = isect_values(g2, o, d);
                leaf_tests<STATS>(x0, x1, x2, first, first + count, tmin, tmax, hit, C);
                done = min(count, 3u);
            }
        }
        if (leaf) {
            const uint32_t end = first + count;
            if (STATS)
                for (uint32_t j = first; j < end; ++j) touch(S.touch_prims, j);
            for (uint32_t i = first + done; i < end; i += 3) {
                const uint32_t last = end - 1, i1 = min(i + 1, last), i2 = min(i + 2, last);
                x0 = isect_values(load_geom(S.geoms + 5 * (size_t)i), o, d);
                x1 = isect_values(load_geom(S.geoms + 5 * (size_t)i1), o, d);
                x2 = isect_values(load_geom(S.geoms + 5 * (size_t)i2), o, d);
                leaf_tests<STATS>(x0, x1, x2, i, end, tmin, tmax, hit, C);
            }
        } else {
            if (!M) D = load_node_v<TV>(S, e);
            const uint2 info = D.info;
            float tl, tr;
            if (STATS) C.box_tests += 2;
            if (STATS) touch(S.touch_nodes, e);
            bool hl = box_hit(D.lmn, D.lmx, o, inv, tmin, tmax, tl);
            bool hr = box_hit(D.rmn, D.rmx, o, inv, tmin, tmax, tr);
            // Both hit: push the far one, continue with the near one (ties to the left), if two
            // more entries fit; one hit: continue with it if one more fits (bvh4d.rs:488-512).
            // One push site with selects instead of two branches.
            const int sz = stk.size();
            const bool both = hl && hr && sz + 2 <= STACK;
            const bool near_left = tl <= tr;
            if (both) stk.push(near_left ? info.y : info.x, near_left ? tr : tl);
            const bool go_left = both ? near_left : (hl && !hr && sz < STACK);
            const bool go_right = both ? !near_left : (hr && !hl && sz < STACK);
            if (go_left || go_right) {
                e = go_left ? info.x : info.y;
                t_enter = go_left ? tl : tr;
                have = true;
            }
        }
    }
    return hit.slot >= 0;
}

// Bvh4D::trace with postponed leaves: a lane that reaches a leaf keeps it pending and goes on
// traversing up to its next leaf; the pending leaves are tested once 3/8 of the wave's lanes still
// tracing hold one (the lanes without one wait meanwhile), so a wave's leaf rounds serve many of
// its lanes at once instead of the lanes at interior nodes and the lanes at leaves waiting for
// each other's loads step after step (1280x800 large frame: primary rays 0.290 -> 0.277 ms; with
// every lane required to hold a leaf, 0.51 ms). Leaves are still tested in the reference's order, each after the
// reference's `t_enter >= t_max` check with the t_max of the leaves before it. Between a leaf and
// its test the lane traverses with the older, larger t_max, so it can visit nodes the reference
// prunes; every leaf under such a node fails that check when its turn comes (its t_enter is at
// least that of the pruned node, which reached the t_max the reference had there), so the tested
// leaves, their order and the hit are the reference's. The extra entries can deepen the stack:
// when a step would leave more than `limit` entries (at most 62, two below the 64-entry limit,
// whose drop rule depends on the exact depth), `redo` is set and the caller traces the ray with
// trace(). Counters are not the reference's (more nodes are visited); instrumented renders use
// trace().
//
// The lane's state lives in registers as values, not as flags: the entry at the cursor (e, or
// NO_ENTRY), the pending leaf (pe, or 0: a leaf entry always has its top bit set) and the stack
// size. Each step recomputes its conditions from them, pops and pushes by selects around one
// unconditional stack access, and a lane past the stack limit ends by emptying its state. Only
// the leaf round and the interior-node step are branches, so the wave's exec-mask bookkeeping
// per step stays small.
constexpr int POSTPONE_EIGHTHS = 3;  // leaf rounds start once 3/8 of a wave's tracing lanes hold a leaf
constexpr uint32_t NO_ENTRY = 0x7FFFFFFFu;  // neither a leaf (top bit clear) nor a node (uploads keep
                                            // node indices below 2^30)
// One leaf's primitives i .. end-1 (up to three loaded), as leaf_tests but keeping only the winning
// slot and the running t_max; the winner's event and barycentrics are recomputed once at the end.
__device__ __forceinline__ void leaf_tests_slot(const Isect &x0, const Isect &x1, const Isect &x2, uint32_t i,
                                                uint32_t end, float tmin, float &tmax, int32_t &slot) {
    const Isect *xs[3] = {&x0, &x1, &x2};
#pragma unroll
    for (uint32_t k = 0; k < 3; ++k) {
        const bool acc = i + k < end && isect_accept(*xs[k], tmin, tmax);
        tmax = acc ? xs[k]->lambda : tmax;
        slot = acc ? (int32_t)(i + k) : slot;
    }
}

template <int TV, class ST>
__device__ bool trace_pp(const Scene &S, V4 o, V4 d, float tmin, float tmax, Hit &hit, Counters &C, ST &stk,
                         int limit, bool &redo) {
    hit.slot = -1;
    redo = false;
    if (S.empty) return false;
    V4 inv = V4{safe_inv(d.t), safe_inv(d.x), safe_inv(d.y), safe_inv(d.z)};
    float t_enter;
    if (!box_hit(S.root_min, S.root_max, o, inv, tmin, tmax, t_enter)) return false;
    uint32_t e = root_of<TV>(S), pe = 0u;
    float pt = 0.0f;
    int32_t slot = -1;
    // TV_WIDE4: the row of the node at the cursor, when the previous step fetched it.
    bool have_row = false;
    NodeData row{};
    stk.reset();
    while (true) {
        stk.pop_if(e == NO_ENTRY && stk.size() > 0, e, t_enter);
        if (e == NO_ENTRY && pe == 0u) break;  // stack empty, nothing at the cursor or pending
        C.steps++;
        // The next leaf in order becomes the pending one.
        const bool to_pending = (e & LEAF) != 0u && pe == 0u;
        pe = to_pending ? e : pe;
        pt = to_pending ? t_enter : pt;
        e = to_pending ? NO_ENTRY : e;
        const int tracing = __popcll(__builtin_amdgcn_ballot_w64(true));
        const int holding = __popcll(__builtin_amdgcn_ballot_w64(pe != 0u));
        if (8 * holding >= POSTPONE_EIGHTHS * tracing) {  // a leaf round: the lanes holding a leaf test it
            if (pe != 0u && !(pt >= tmax)) {
                uint32_t first, count;
                leaf_range(S, pe, first, count);
                const uint32_t end = first + count;
                for (uint32_t i = first; i < end; i += 3) {
                    const uint32_t last = end - 1, i1 = min(i + 1, last), i2 = min(i + 2, last);
                    const Isect x0 = isect_values(load_geom(S.geoms + 5 * (size_t)i), o, d);
                    const Isect x1 = isect_values(load_geom(S.geoms + 5 * (size_t)i1), o, d);
                    const Isect x2 = isect_values(load_geom(S.geoms + 5 * (size_t)i2), o, d);
                    leaf_tests_slot(x0, x1, x2, i, end, tmin, tmax, slot);
                }
            }
            pe = 0u;
        } else {
            // An interior node at the cursor, unless the running t_max prunes it: its two children.
            const uint32_t cur = e;
            const bool node = (cur & LEAF) == 0u && cur != NO_ENTRY;
            e = node ? NO_ENTRY : e;
            const bool had_row = have_row;
            have_row = node ? false : have_row;
            if (node && !(t_enter >= tmax)) {
                const bool over = stk.size() + 2 > limit;  // past the limit: traced again by trace()
                NodeData D, DL, DR;
                if constexpr (TV == TV_WIDE4) {
                    D = row;
                    if (!had_row) D = load_node(S, cur);
                    // Both children's box pairs (a leaf child's stands in with this node's own
                    // row): their loads run under this step's box tests.
                    DL = load_node(S, (D.info.x & LEAF) ? cur : D.info.x);
                    DR = load_node(S, (D.info.y & LEAF) ? cur : D.info.y);
                } else {
                    D = load_node_v<TV>(S, cur);
                }
                float tl, tr;
                uint32_t hits = (box_hit(D.lmn, D.lmx, o, inv, tmin, tmax, tl) ? 1u : 0u) |
                                (box_hit(D.rmn, D.rmx, o, inv, tmin, tmax, tr) ? 2u : 0u);
                // Both tests settled before the stack write's branch (a min/max operand defined in
                // another block is canonicalized again there), and kept as an integer so that the
                // choice below stays selects.
                asm volatile("" : "+v"(hits), "+v"(tl), "+v"(tr));
                // Both hit: the far one pushed, the near one (ties to the left) continued; one
                // hit: continued. go: 1 left, 2 right, 0 neither.
                const bool near_left = tl <= tr;
                const uint32_t go = hits == 3u ? (near_left ? 1u : 2u) : hits;
                stk.put(near_left ? D.info.y : D.info.x, near_left ? tr : tl);
                stk.grow(hits == 3u && !over);
                e = go != 0u && !over ? (go == 1u ? D.info.x : D.info.y) : NO_ENTRY;
                t_enter = go == 1u ? tl : tr;
                if constexpr (TV == TV_WIDE4) {
                    have_row = e != NO_ENTRY && (e & LEAF) == 0u;
                    row = go == 1u ? DL : DR;
                }
                if (over) {
                    redo = true;
                    pe = 0u;
                    stk.reset();
                }
            }
        }
    }
    hit.slot = slot;
    if (slot >= 0) {  // the winner's values, recomputed by the same operations
        const Isect x = isect_values(load_geom(S.geoms + 5 * (size_t)slot), o, d);
        hit.lambda = x.lambda;
        hit.event = x.event;
        isect_bary(x, hit.bary);
    }
    return slot >= 0;
}

// Bvh4D::is_occluded (bvh4d.rs:523-578): right pushed, then left pushed and popped at once.
template <bool STATS, class ST, int TV = TV_PLAIN>
__device__ bool occluded(const Scene &S, V4 o, V4 d, float tmin, float tmax, Counters &C, ST &stk) {
    static_assert(!(STATS && TV == TV_LDSTOP), "instrumented traversals name nodes by their index");
    if (S.empty) return false;
    V4 inv = V4{safe_inv(d.t), safe_inv(d.x), safe_inv(d.y), safe_inv(d.z)};
    float en;
    if (!box_hit(S.root_min, S.root_max, o, inv, tmin, tmax, en)) return false;
    // The lane's state as values (see trace_pp): the entry at the cursor or NO_ENTRY, and whether
    // a primitive was hit, after which the lane ends.
    uint32_t e = root_of<TV>(S);
    bool occ = false;
    stk.reset();
    while (true) {
        float unused = 0.0f;
        stk.pop_if(e == NO_ENTRY && stk.size() > 0, e, unused);
        if (e == NO_ENTRY) break;
        C.steps++;
        if (STATS) C.shadow_node_pops++;
        const uint32_t cur = e;
        e = NO_ENTRY;
        if (cur & LEAF) {
            uint32_t first, count;
            leaf_range(S, cur, first, count);
            // Up to three primitives per round of loads, tested in slot order (as in trace()).
            const uint32_t end = first + count;
            for (uint32_t i = first; i < end && !occ; i += 3) {
                const uint32_t last = end - 1, i1 = min(i + 1, last), i2 = min(i + 2, last);
                const bool h0 = isect_accept(isect_values(load_geom(S.geoms + 5 * (size_t)i), o, d), tmin, tmax);
                const bool h1 = isect_accept(isect_values(load_geom(S.geoms + 5 * (size_t)i1), o, d), tmin, tmax);
                const bool h2 = isect_accept(isect_values(load_geom(S.geoms + 5 * (size_t)i2), o, d), tmin, tmax);
                const bool hs[3] = {h0, h1, h2};
#pragma unroll
                for (uint32_t k = 0; k < 3; ++k) {
                    if (i + k >= end || occ) break;
                    if (STATS) C.shadow_prim_tests++;
                    if (STATS) touch(S.touch_prims, i + k);
                    occ = hs[k];
                }
            }
            if (occ) stk.reset();
        } else {
            const NodeData D = load_node_v<TV>(S, cur);
            float t;
            if (STATS) C.shadow_box_tests += 2;
            if (STATS) touch(S.touch_nodes, cur);
            const bool hl = box_hit(D.lmn, D.lmx, o, inv, tmin, tmax, t);
            const bool hr = box_hit(D.rmn, D.rmx, o, inv, tmin, tmax, t);
            if (hr && stk.size() < STACK) stk.push(D.info.y, 0.0f);
            e = hl && stk.size() < STACK ? D.info.x : NO_ENTRY;
        }
    }
    return occ;
}

// HitRecord4D via Simplex4D::interpolate_hit (mesh4d.rs:382-425)
struct Rec {
    float lambda;
    V4 event;
    float tet_bary[4];
    V3 tri_bary;
    float uv_u, uv_v, tau, grad_u, grad_v, uvws_u, uvws_v, decided_at;
    V4 normal, four_vel;
    uint32_t material;
};

__device__ __forceinline__ float blend(float4 a, const float b[4]) {
    return a.x * b[0] + a.y * b[1] + a.z * b[2] + a.w * b[3];
}

__device__ __forceinline__ void interpolate(const float4 *A, const Hit &h, Rec &r) {
    r.lambda = h.lambda;
    r.event = h.event;
    for (int i = 0; i < 4; ++i) r.tet_bary[i] = h.bary[i];
    r.uv_u = blend(A[0], h.bary);
    r.uv_v = blend(A[1], h.bary);
    r.tri_bary = V3{blend(A[2], h.bary), blend(A[3], h.bary), blend(A[4], h.bary)};
    r.tau = blend(A[5], h.bary);
    r.grad_u = blend(A[6], h.bary);
    r.grad_v = blend(A[7], h.bary);
    r.normal = f4v(A[8]);
    r.four_vel = f4v(A[9]);
    float4 misc = A[10];
    r.uvws_u = misc.x;
    r.uvws_v = misc.y;
    r.material = __float_as_uint(misc.z);
    r.decided_at = blend(A[11], h.bary);
}

struct Mat {
    V3 base, checker;
    float emission, grid, wire, cavity;
    uint32_t pattern;
};
__device__ __forceinline__ Mat load_mat(const Scene &S, uint32_t id) {
    const float4 *m = S.materials + 3 * (size_t)id;
    float4 a = m[0], b = m[1], c = m[2];
    return Mat{V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z}, a.w, b.w, c.x, c.z, __float_as_uint(c.y)};
}

constexpr float TAU_F = 6.28318530717958647692f;
constexpr float PI_F = 3.14159265358979323846f;

__device__ V3 hsv_to_rgb(float h_deg, float s, float v) {  // renderer.rs:1081-1100
    float c = v * s;
    float h = rem_euclid_(h_deg / 60.0f, 6.0f);
    float x = c * (1.0f - fabsf(rem_euclid_(h, 2.0f) - 1.0f));
    float m = v - c;
    float r, g, b;
    if (h < 1.0f) { r = c; g = x; b = 0.0f; }
    else if (h < 2.0f) { r = x; g = c; b = 0.0f; }
    else if (h < 3.0f) { r = 0.0f; g = c; b = x; }
    else if (h < 4.0f) { r = 0.0f; g = x; b = c; }
    else if (h < 5.0f) { r = x; g = 0.0f; b = c; }
    else { r = c; g = 0.0f; b = x; }
    return V3{r + m, g + m, b + m};
}

__device__ V3 sky_gradient(V4 dir) {  // renderer.rs:1056-1062
    V3 s = normalized(spatial(dir));
    float horizon = clamp_(s.y * 0.5f + 0.5f, 0.0f, 1.0f);
    return V3{0.12f, 0.16f, 0.25f} * (1.0f - horizon) + V3{0.04f, 0.06f, 0.13f} * horizon;
}

// Whether rem_euclid(x, 1) < 0.025 is certain for every x within m of xa: 1 certainly below, 0
// certainly not, -1 undecided (within m of 0, 0.025 or 1, or NaN).
__device__ __forceinline__ int grid_side(float xa, float m) {
    const float f = xa - floorf(xa);  // exact: in [0, 1)
    if (f >= 0.025f + m && f <= 1.0f - m) return 0;
    if (f >= m && f <= 0.025f - m) return 1;
    return -1;
}

// shade_sky's grid lines depend on acos and atan2 only through those comparisons, so the f32
// library functions decide them first: the f32 acosf / atan2f are within a few ulp of the true
// angle and acosf_ / atan2f_ within half an ulp, so the reference's gu and gv (the same operations
// on their results) lie within 1e-5 of the values computed from the fast angles, far inside the
// margin 2e-3 below which a comparison is left undecided. Only undecided directions (about 1% of
// the sky) evaluate acosf_ and atan2f_; the line, and with it the colour, is the reference's.
__device__ V3 shade_sky(V4 dir) {  // renderer.rs:1064-1079
    V3 s = normalized(spatial(dir));
    V3 sky = sky_gradient(dir);
    const float cy = clamp_(s.y, -1.0f, 1.0f);
    constexpr float M = 2e-3f;
    const int su = grid_side(atan2f(s.z, s.x) / PI_F * 12.0f, M);
    const int sv = grid_side(acosf(cy) / PI_F * 6.0f, M);
    bool line;
    if (su == 1 || sv == 1) {
        line = true;
    } else if (su == 0 && sv == 0) {
        line = false;
    } else {
        float theta = acosf_(cy);
        float phi = atan2f_(s.z, s.x);
        float gu = rem_euclid_(phi / PI_F * 12.0f, 1.0f);
        float gv = rem_euclid_(theta / PI_F * 6.0f, 1.0f);
        line = gu < 0.025f || gv < 0.025f;
    }
    if (line) sky += V3{0.06f, 0.09f, 0.14f};
    return sky;
}

__device__ V3 distance_fog(V3 c, V4 dir, float lambda) {  // renderer.rs:691-695
    float fog = expf_(-0.008f * lambda);
    V3 sky = sky_gradient(dir);
    return c * fog + sky * (1.0f - fog);
}

__device__ float decision_age(V4 event, V4 origin, V4 u) {  // renderer.rs:319-325
    V4 d = event - origin;
    float time_ago = mdot(d, u);
    float dist = sqrtf(fmax_(time_ago * time_ago + mdot(d, d), 0.0f));
    return time_ago + dist;
}

__device__ V3 decision_age_color(float age) {  // renderer.rs:494-507
    if (age < 0.0f) return V3{0.14f, 0.14f, 0.18f};
    V3 col = hsv_to_rgb(fmin_(age * 4.0f, 240.0f), 0.82f, 0.88f);
    if (age < 0.45f) {
        float f = 1.0f - age / 0.45f;
        col = col * (1.0f - f * 0.7f) + V3{1.0f, 0.96f, 0.65f} * (f * 0.9f);
    }
    if (fabsf(rem_euclid_(age * 0.5f, 1.0f) - 0.5f) > 0.465f) col = V3{0.95f, 0.98f, 1.0f};
    return col;
}

__device__ V3 haze(V3 base, V4 event, const specrel_frame &f) {  // renderer.rs:514-541
    const float SHELL = 0.18f;
    float inside = -__builtin_inff();
    for (uint32_t i = 0; i < f.num_haze; ++i) {
        V4 d = event - lv(f.haze_origin[i]);
        inside = fmax_(inside, d.t - length(spatial(d)));
    }
    if (inside <= -SHELL) return base;
    V3 col = base;
    if (fabsf(inside) < SHELL) {
        float k = 1.0f - fabsf(inside) / SHELL;
        col += V3{0.45f, 0.92f, 1.0f} * (0.85f * k * k);
    }
    if (inside > 0.0f) {
        float ripple = 0.5f + 0.5f * sinf_(event.x * 3.5f + event.z * 3.5f - event.t * 4.0f) * cosf_(inside * 5.0f);
        V3 hz = V3{0.32f, 0.18f, 0.62f} * (1.0f - 0.35f * ripple) + V3{0.18f, 0.72f, 0.92f} * (0.35f * ripple);
        float alpha = 0.55f + 0.25f * fmin_(inside * 0.5f, 1.0f);
        col = col * (1.0f - alpha) + hz * alpha;
    }
    return col;
}

__device__ bool point_on_hand(float dx, float dy, float angle_cw, float len, float half_w, float back) {
    float hx = sinf_(angle_cw), hy = cosf_(angle_cw);  // renderer.rs:930-943
    float along = dx * hx + dy * hy;
    float perp = fabsf(dx * hy - dy * hx);
    return along >= back && along <= len && perp <= half_w;
}

__device__ V3 mini_clock(float mdx, float mdy, float mr, float mini_tau, const Mat &m) {  // renderer.rs:880-927
    if (mr > 0.080f) return m.base * 1.15f;
    float angle = rem_euclid_(atan2f_(mdx, mdy), TAU_F);
    float fast = rem_euclid_((mini_tau / 1.0f) * TAU_F, TAU_F);
    float main_a = rem_euclid_((mini_tau / 10.0f) * TAU_F, TAU_F);
    V3 col = V3{0.04f, 0.06f, 0.10f};
    if (angle <= fast && mr < 0.066f) col += m.base * 0.25f;
    if (0.062f <= mr && mr <= 0.078f) {
        float sector = angle / (TAU_F / 10.0f);
        float dist = fabsf(sector - round_(sector));
        int32_t tick = f2i_sat(round_(sector)) % 10;
        bool card = tick == 0 || tick == 5;
        float hw = card ? 0.13f : 0.07f;
        if (dist < hw) return tick == 0 ? V3{1.0f, 0.88f, 0.15f} : V3{0.80f, 0.85f, 0.92f};
    }
    if (point_on_hand(mdx, mdy, fast, 0.068f, 0.0065f, -0.008f)) return V3{0.25f, 0.98f, 1.0f};
    if (point_on_hand(mdx, mdy, main_a, 0.052f, 0.0085f, -0.008f)) return V3{1.0f, 0.95f, 0.30f};
    if (mr < 0.012f) return V3{1.0f, 1.0f, 1.0f};
    return col;
}

__device__ bool digital_tau(const Tables &T, float u, float v, float tau) {  // renderer.rs:946-981
    float rel_u = (u - 0.28f) / (0.72f - 0.28f);
    float rel_v = (0.18f - v) / (0.18f - 0.06f);
    uint64_t col = f2u_sat(floorf(rel_u * 35.0f));
    uint64_t row = f2u_sat(floorf(rel_v * 7.0f));
    if (col >= 35 || row >= 7) return false;
    uint32_t ci = (uint32_t)(col / 6), cc = (uint32_t)(col % 6);
    if (ci >= 6 || cc >= 5) return false;
    int32_t t_tenths = f2i_sat(round_(tau * 10.0f));
    uint32_t abs_t = t_tenths < 0 ? 0u - (uint32_t)t_tenths : (uint32_t)t_tenths;
    uint32_t tenths = abs_t % 10, ones = (abs_t / 10) % 10, tens = (abs_t / 100) % 10;
    uint32_t g;
    switch (ci) {
        case 0: g = t_tenths < 0 ? 11 : 10; break;
        case 1: g = tens; break;
        case 2: g = ones; break;
        case 3: g = 12; break;
        case 4: g = tenths; break;
        default: g = 13; break;
    }
    return ((T.glyphs[g * 7 + row] >> (4 - cc)) & 1) != 0;
}

__device__ V3 clock_face(const Tables &T, const Rec &h, const Mat &m) {  // renderer.rs:737-874
    float grid_scale = fmax_(m.grid, 1e-4f);
    float inv_grid = 1.0f / grid_scale;
    float u = rem_euclid_(h.uv_u * grid_scale, 1.0f);
    float v = rem_euclid_(h.uv_v * grid_scale, 1.0f);
    // clock_tile_central_proper_time (renderer.rs:704-715)
    float ig2 = 1.0f / fmax_(grid_scale, 1e-4f);
    float ctau = h.tau - (((u - 0.5f) * ig2) * h.grad_u + ((v - 0.5f) * ig2) * h.grad_v);
    const float cx[4] = {0.135f, 0.865f, 0.135f, 0.865f};
    const float cy[4] = {0.865f, 0.865f, 0.135f, 0.135f};
    for (int k = 0; k < 4; ++k) {
        float mdx = u - cx[k], mdy = v - cy[k];
        float mr = sqrtf(mdx * mdx + mdy * mdy);
        if (mr <= 0.095f) {
            float osc = ctau + ((cx[k] - 0.5f) * inv_grid) * h.grad_u + ((cy[k] - 0.5f) * inv_grid) * h.grad_v;
            float px = (mdx * inv_grid) * h.uvws_u;
            float py = (mdy * inv_grid) * h.uvws_v;
            float mr_phys = sqrtf(px * px + py * py);
            return mini_clock(mdx, mdy, mr, osc - mr_phys, m);
        }
    }
    if ((0.035f <= v && v <= 0.205f) && (0.25f <= u && u <= 0.75f)) {
        if (!(0.050f <= v && v <= 0.190f) || !(0.265f <= u && u <= 0.735f)) return m.base * 0.9f;
        if ((0.06f <= v && v <= 0.18f) && (0.28f <= u && u <= 0.72f) && digital_tau(T, u, v, ctau))
            return V3{1.0f, 0.96f, 0.30f};
        return V3{0.03f, 0.04f, 0.07f};
    }
    float dx = u - 0.50f, dy = v - 0.57f;
    float r = sqrtf(dx * dx + dy * dy);
    if (r > 0.330f) {
        int32_t gu = f2i_sat(floorf(u * 4.0f)), gv = f2i_sat(floorf(v * 4.0f));
        int32_t sum = (int32_t)((uint32_t)gu + (uint32_t)gv);
        return (((sum % 2) + 2) % 2) == 0 ? m.base * 0.78f : m.checker * 0.78f;
    }
    if (r > 0.292f) return m.base * 1.15f;
    float dxp = (dx * inv_grid) * h.uvws_u, dyp = (dy * inv_grid) * h.uvws_v;
    float r_phys = sqrtf(dxp * dxp + dyp * dyp);
    float dial_tau = ctau - r_phys;
    float angle = rem_euclid_(atan2f_(dx, dy), TAU_F);
    float main_a = rem_euclid_((dial_tau / 10.0f) * TAU_F, TAU_F);
    V3 dial = V3{0.04f, 0.06f, 0.10f};
    if (angle <= main_a && r < 0.236f) dial += m.base * 0.22f;
    if (0.220f <= r && r <= 0.286f) {
        float sector = angle / (TAU_F / 10.0f);
        float dist = fabsf(sector - round_(sector));
        int32_t tick = f2i_sat(round_(sector)) % 10;
        bool card = tick == 0 || tick == 5;
        float hw = card ? 0.105f : 0.055f;
        float inner = card ? 0.218f : 0.242f;
        if (dist < hw && r >= inner) {
            if (tick == 0) return V3{1.0f, 0.88f, 0.15f};
            if (card) return V3{1.0f, 1.0f, 1.0f};
            return V3{0.75f, 0.80f, 0.88f};
        }
    }
    float fast = rem_euclid_((dial_tau / 1.0f) * TAU_F, TAU_F);
    if (point_on_hand(dx, dy, fast, 0.254f, 0.012f, -0.025f)) return V3{0.25f, 0.98f, 1.0f};
    if (point_on_hand(dx, dy, main_a, 0.228f, 0.022f, -0.025f)) return V3{1.0f, 0.95f, 0.30f};
    if (point_on_hand(dx, dy, main_a, 0.236f, 0.030f, -0.025f)) return V3{0.02f, 0.02f, 0.04f};
    float slow = rem_euclid_((dial_tau / 100.0f) * TAU_F, TAU_F);
    if (point_on_hand(dx, dy, slow, 0.162f, 0.026f, -0.025f)) return V3{1.0f, 0.35f, 0.25f};
    if (r < 0.034f) return V3{1.0f, 1.0f, 1.0f};
    return dial;
}

__device__ V3 light_clock(float uvu, float uvv, float tau, const Mat &m) {  // renderer.rs:988-1052
    float u = rem_euclid_(uvu * m.grid, 1.0f);
    float v = rem_euclid_(uvv * m.grid, 1.0f);
    if (!(0.12f <= u && u <= 0.88f)) return m.base * 0.65f;
    if ((0.04f <= v && v <= 0.10f) || (0.90f <= v && v <= 0.96f)) return V3{0.75f, 0.90f, 0.98f};
    if (!(0.10f <= v && v <= 0.90f)) return V3{0.08f, 0.12f, 0.16f};
    float period = fmax_(2.0f * m.cavity, 1e-4f);
    float cycle = rem_euclid_(tau / period, 1.0f);
    float bounce = cycle < 0.5f ? cycle * 2.0f : (1.0f - cycle) * 2.0f;
    float photon_v = 0.10f + 0.80f * bounce;
    float dv = fabsf(v - photon_v), du = fabsf(u - 0.5f);
    float dist = sqrtf(du * du * 0.7f + dv * dv);
    V3 col = V3{0.05f, 0.09f, 0.13f};
    float ruler = rem_euclid_((v - 0.10f) / 0.80f * 8.0f, 1.0f);
    if ((u < 0.22f || u > 0.78f) && ruler < 0.08f) col += V3{0.20f, 0.35f, 0.35f};
    if (du < 0.015f) col += V3{0.06f, 0.16f, 0.16f};
    if (dv < 0.032f) {
        float a = 1.0f - dv / 0.032f;
        col += m.base * (0.9f * (a * a));
    }
    if (dist < 0.11f) {
        float a = 1.0f - dist / 0.11f;
        col += V3{1.0f, 1.0f, 0.85f} * (1.4f * (a * a));
    }
    return col;
}

__device__ V3 albedo(const Tables &T, const Rec &h, const Mat &m) {  // renderer.rs:544-580
    V3 a;
    if (m.pattern == SPECREL_PATTERN_CLOCK) {
        a = clock_face(T, h, m);
    } else if (m.pattern == SPECREL_PATTERN_LIGHT_CLOCK) {
        a = light_clock(h.uv_u, h.uv_v, h.tau, m);
    } else {
        int32_t gu = f2i_sat(floorf(h.uv_u * m.grid)), gv = f2i_sat(floorf(h.uv_v * m.grid));
        int32_t sum = (int32_t)((uint32_t)gu + (uint32_t)gv);
        V3 col = (((sum % 2) + 2) % 2) == 0 ? m.base : m.checker;
        float mtb = fmin_(fmin_(h.tri_bary.x, h.tri_bary.y), h.tri_bary.z);
        if (mtb < m.wire) {
            float et = clamp_(mtb / m.wire, 0.0f, 1.0f);
            col = V3{0.07f, 0.07f, 0.07f} * (1.0f - et) + col * et;
        }
        a = col;
    }
    if (m.pattern != SPECREL_PATTERN_CHECKER) {
        float ut = rem_euclid_(h.uv_u * m.grid, 1.0f), vt = rem_euclid_(h.uv_v * m.grid, 1.0f);
        float bd = fmin_(fmin_(ut, 1.0f - ut), fmin_(vt, 1.0f - vt));
        if (bd < 0.025f) a = a * 0.35f;
    }
    return a;
}

// compute_illumination + light_emission_event (renderer.rs:585-653)
// `sh.query(id, origin, dir, tmin, tmax)` answers Bvh4D::is_occluded for shadow query `id`
// of this pixel (0: the light, 1 + i: flash i): traced at once, queued for a later pass, or
// looked up from that pass's result.
template <class SH>
__device__ float illumination(const Scene &S, const specrel_frame &f, V4 dir, const Rec &h, const Mat &m, SH &sh) {
    V3 light = V3{S.light.y, S.light.z, S.light.w};
    V3 s2l = light - spatial(h.event);
    float dist = fmax_(length(s2l), 1e-4f);
    float t_emit;
    if (f.shadow_mode == SPECREL_SHADOW_CAMERA_SLICE) {
        V3 beta = beta_from_four_velocity(lv(f.e[0]));
        t_emit = h.event.t + dot(beta, s2l);
    } else {
        t_emit = h.event.t - dist;
    }
    V4 to_light = (from_ts(t_emit, light) - h.event) * (1.0f / dist);
    float cl = fmax_(rest_frame_cos(to_light, h.normal, h.four_vel), 0.0f);
    float cv = fmax_(-rest_frame_cos(dir, h.normal, h.four_vel), 0.0f);
    if (m.pattern == SPECREL_PATTERN_CHECKER) {
        bool in_shadow = false;
        if (cl > 0.0f && f.shadow_mode != SPECREL_SHADOW_OFF)
            in_shadow = sh.query(0u, h.event, to_light, 2e-3f, dist - 2e-3f);
        float sf = in_shadow ? 0.32f : 1.0f;
        float atten = 1.0f / (1.0f + 0.006f * dist * dist);
        return m.emission + 0.30f * cv + 0.80f * sf * cl * atten;
    }
    return 0.72f + 0.38f * cv + 0.22f * cl;
}

// compute_flash_illumination (renderer.rs:658-688)
template <class SH>
__device__ V3 flashes(const Scene &S, const specrel_frame &f, const Rec &h, SH &sh) {
    V3 total = V3{0.0f, 0.0f, 0.0f};
    for (uint32_t i = 0; i < f.num_flashes; ++i) {
        V4 at = lv(f.flash_at[i]);
        V3 color = V3{f.flash_color_width[i][0], f.flash_color_width[i][1], f.flash_color_width[i][2]};
        float width = f.flash_color_width[i][3], intensity = f.flash_intensity[i];
        float dt = h.event.t - at.t;
        if (dt <= 0.0f) continue;
        float dist = fmax_(length(spatial(at) - spatial(h.event)), 1e-4f);
        float shell_err = fabsf(dt - dist);
        float half_w = fmax_(0.5f * width, 1e-3f);
        if (shell_err >= half_w) continue;
        float k = 1.0f - shell_err / half_w;
        float pulse = k * k;
        V4 to_flash = (at - h.event) * (1.0f / dist);
        float cf = fmax_(rest_frame_cos(to_flash, h.normal, h.four_vel), 0.0f);
        bool in_shadow = false;
        if (cf > 0.0f && f.shadow_mode != SPECREL_SHADOW_OFF)
            in_shadow = sh.query(1u + i, h.event, to_flash, 2e-3f, dist - 2e-3f);
        float sf = in_shadow ? 0.25f : 1.0f;
        float atten = 1.0f / (1.0f + 0.015f * dist * dist);
        total += color * (intensity * pulse * atten * (0.25f + 0.75f * sf * cf));
    }
    return total;
}

// shade_hit (renderer.rs:414-489)
template <class SH>
__device__ V3 shade_hit(const Scene &S, const Tables &T, const specrel_frame &f, V4 dir, const Rec &h, SH &sh) {
    Mat m = load_mat(S, h.material);
    if (f.mode == SPECREL_MODE_SHADED) {
        V3 a = albedo(T, h, m);
        float lighting = illumination(S, f, dir, h, m, sh);
        V3 lit = a * lighting;
        if (f.num_flashes > 0) {
            V3 fl = flashes(S, f, h, sh);
            lit += (a * 0.65f + V3{0.35f, 0.35f, 0.35f}) * fl;
        }
        return distance_fog(lit, dir, h.lambda);
    }
    if (f.mode == SPECREL_MODE_ISOCHRONES) {
        float phase = rem_euclid_(h.lambda * 0.12f, 1.0f);
        V3 base = hsv_to_rgb(phase * 300.0f, 0.8f, 0.85f);
        float tc = fabsf(rem_euclid_(h.event.t, 1.0f) - 0.5f);
        float uc = fabsf(rem_euclid_(h.tau, 1.0f) - 0.5f);
        if (tc > 0.45f) return V3{1.0f, 1.0f, 1.0f};
        if (uc > 0.46f) return V3{0.1f, 0.1f, 0.1f};
        return base;
    }
    if (f.mode == SPECREL_MODE_INSPECTOR) {
        const float *b = h.tet_bary;
        float min_b = fmin_(fmin_(fmin_(b[0], b[1]), b[2]), b[3]);
        float edge = min_b < 0.035f ? 0.12f : 1.0f;
        float sb = clamp_((h.four_vel.t - 1.0f) * 0.8f, 0.0f, 1.0f);
        V3 tint = V3{0.25f + 0.65f * b[1] + 0.3f * sb, 0.25f + 0.65f * b[2], 0.35f + 0.55f * b[3] * (1.0f - 0.5f * sb)};
        return tint * edge;
    }
    float age;
    if (h.decided_at > -1.0e9f * 0.5f) age = f.frontier_tau - h.decided_at;
    else age = decision_age(h.event, lv(f.frontier_origin), lv(f.e[0]));
    float cv = fmax_(-rest_frame_cos(dir, h.normal, h.four_vel), 0.0f);
    return distance_fog(decision_age_color(age) * (0.55f + 0.45f * cv), dir, h.lambda);
}

// Shadow queries traced where they are asked (the direct path).
template <bool STATS, class ST>
struct TraceShadow {
    const Scene &S;
    Counters &C;
    ST &stk;
    __device__ bool query(uint32_t, V4 o, V4 d, float tmin, float tmax) {
        if (STATS) C.shadow_rays++;
        return occluded<STATS>(S, o, d, tmin, tmax, C, stk);
    }
};

// One queued shadow query: the occlusion test of Bvh4D::is_occluded for (pixel, id).
struct ShadowQuery {
    float4 o, d;
    float tmin, tmax;
    uint32_t pixel, id;
};

// Shadow queries appended to a queue (the value returned is not used by the caller's pass).
struct EnqueueShadow {
    ShadowQuery *queue;
    uint32_t *count;
    uint32_t cap;
    uint32_t pixel;
    __device__ bool query(uint32_t id, V4 o, V4 d, float tmin, float tmax) {
        uint32_t k = atomicAdd(count, 1u);
        if (k < cap)
            queue[k] = ShadowQuery{make_float4(o.t, o.x, o.y, o.z), make_float4(d.t, d.x, d.y, d.z), tmin, tmax, pixel, id};
        return false;
    }
};

// Shadow queries traced where they are asked, their answers kept as bit `id` of `mask` for a
// later shading pass that looks them up (LookupShadow).
template <bool STATS, class ST, int TV = TV_PLAIN>
struct RecordShadow {
    const Scene &S;
    Counters &C;
    ST &stk;
    uint32_t mask;
    __device__ bool query(uint32_t id, V4 o, V4 d, float tmin, float tmax) {
        if (STATS) C.shadow_rays++;
        bool occ = occluded<STATS, ST, TV>(S, o, d, tmin, tmax, C, stk);
        if (occ) mask |= 1u << id;
        return occ;
    }
};

// Shadow queries answered from the occlusion pass: bit `id` of the pixel's mask.
struct LookupShadow {
    uint32_t mask;
    __device__ bool query(uint32_t id, V4, V4, float, float) { return (mask >> id) & 1u; }
};

// Closest-hit traversal of one ray as a resumable state machine: step() processes one stack
// entry (the loop body of trace() above) so that a lane can take a new ray when its ray ends.
struct RayState {
    V4 o, d, inv;
    float tmax, t_enter;
    int32_t best;
    uint32_t e, pix;
    bool have;
};

template <class ST>
__device__ __forceinline__ bool trace_begin(const Scene &S, RayState &r, float tmin, ST &stk) {
    r.inv = V4{safe_inv(r.d.t), safe_inv(r.d.x), safe_inv(r.d.y), safe_inv(r.d.z)};
    r.best = -1;
    stk.reset();
    r.have = false;
    if (S.empty) return false;
    float te;
    if (!box_hit(S.root_min, S.root_max, r.o, r.inv, tmin, r.tmax, te)) return false;
    r.e = S.root_entry;
    r.t_enter = te;
    r.have = true;
    return true;
}

// Returns false when the ray is finished.
template <bool STATS, class ST>
__device__ __forceinline__ bool trace_step(const Scene &S, RayState &r, float tmin, Counters &C, ST &stk) {
    if (!r.have) {
        if (stk.size() == 0) return false;
        stk.pop(r.e, r.t_enter);
    }
    r.have = false;
    if (r.t_enter >= r.tmax) return true;
    if (STATS) C.node_pops++;
    if (r.e & LEAF) {
        uint32_t first, count;
        leaf_range(S, r.e, first, count);
        for (uint32_t i = first; i < first + count; ++i) {
            float lam, b[4];
            V4 ev;
            if (STATS) C.prim_tests++;
            if (STATS) touch(S.touch_prims, i);
            if (intersect(S.geoms + 5 * (size_t)i, r.o, r.d, tmin, r.tmax, lam, ev, b)) {
                r.tmax = lam;
                r.best = (int32_t)i;
            }
        }
    } else {
        const float4 *bx = S.nbox + 4 * (size_t)r.e;
        float4 lmn = bx[0], lmx = bx[1], rmn = bx[2], rmx = bx[3];
        uint2 info = S.ninfo[r.e];
        float tl, tr;
        if (STATS) C.box_tests += 2;
        if (STATS) touch(S.touch_nodes, r.e);
        bool hl = box_hit(lmn, lmx, r.o, r.inv, tmin, r.tmax, tl);
        bool hr = box_hit(rmn, rmx, r.o, r.inv, tmin, r.tmax, tr);
        if (hl && hr) {
            if (stk.size() + 2 <= STACK) {
                if (tl <= tr) {
                    stk.push(info.y, tr);
                    r.e = info.x;
                    r.t_enter = tl;
                } else {
                    stk.push(info.x, tl);
                    r.e = info.y;
                    r.t_enter = tr;
                }
                r.have = true;
            }
        } else if (hl) {
            if (stk.size() < STACK) {
                r.e = info.x;
                r.t_enter = tl;
                r.have = true;
            }
        } else if (hr) {
            if (stk.size() < STACK) {
                r.e = info.y;
                r.t_enter = tr;
                r.have = true;
            }
        }
    }
    return true;
}

// Any-hit (is_occluded) as a state machine. step(): 0 = running, 1 = occluded, 2 = clear.
template <bool STATS, class ST>
__device__ __forceinline__ int occlude_step(const Scene &S, RayState &r, float tmin, Counters &C, ST &stk) {
    if (!r.have) {
        if (stk.size() == 0) return 2;
        float unused;
        stk.pop(r.e, unused);
    }
    r.have = false;
    if (STATS) C.shadow_node_pops++;
    if (r.e & LEAF) {
        uint32_t first, count;
        leaf_range(S, r.e, first, count);
        for (uint32_t i = first; i < first + count; ++i) {
            float lam, b[4];
            V4 ev;
            if (STATS) C.shadow_prim_tests++;
            if (STATS) touch(S.touch_prims, i);
            if (intersect(S.geoms + 5 * (size_t)i, r.o, r.d, tmin, r.tmax, lam, ev, b)) return 1;
        }
    } else {
        const float4 *bx = S.nbox + 4 * (size_t)r.e;
        uint2 info = S.ninfo[r.e];
        float t;
        if (STATS) C.shadow_box_tests += 2;
        if (STATS) touch(S.touch_nodes, r.e);
        bool hl = box_hit(bx[0], bx[1], r.o, r.inv, tmin, r.tmax, t);
        bool hr = box_hit(bx[2], bx[3], r.o, r.inv, tmin, r.tmax, t);
        if (hr && stk.size() < STACK) stk.push(info.y, 0.0f);
        if (hl && stk.size() < STACK) {
            r.e = info.x;
            r.have = true;
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------------------
// Stars (star.rs:114-286)
// ---------------------------------------------------------------------------------------
__device__ float planck(float lambda_um, float temperature) {  // star.rs:162-166
    float x = fmin_(14388.0f / (lambda_um * fmax_(temperature, 1.0f)), 80.0f);
    float l5 = (float)pow((double)lambda_um, 5.0);
    return 1.0f / (l5 * (expf_(x) - 1.0f));
}
__device__ V3 blackbody(float T) {  // star.rs:171-180
    return V3{planck(0.610f, T) / planck(0.610f, 6504.0f), planck(0.550f, T) / planck(0.550f, 6504.0f),
              planck(0.465f, T) / planck(0.465f, 6504.0f)};
}

__device__ float past_the_cone(V4 event, V4 origin, V4 e0, float cone) {  // star.rs:215-223
    V4 d = event - origin;
    float tau = -mdot(d, e0);
    if (cone == 0.0f) return tau;
    float rho = sqrtf(fmax_(mdot(d, d) + tau * tau, 0.0f));
    return tau - cone * rho;
}

__device__ float cone_crossing(V4 ev, V4 u, V4 origin, V4 e0, float cone) {  // star.rs:185-210
    V4 d0 = ev - origin;
    if (cone == 0.0f) {
        float tau0 = -mdot(d0, e0);
        float gamma = -mdot(u, e0);
        return -tau0 / gamma;
    }
    float a = mdot(d0, d0);
    float b = mdot(d0, u);
    float sq = sqrtf(fmax_(b * b + a, 0.0f));
    if (cone < 0.0f) return b > 0.0f ? -a / (b + sq) : b - sq;
    return b < 0.0f ? -a / (b - sq) : b + sq;
}

// worldline_cone_event (star.rs:228-235) with slice::partition_point's binary search.
__device__ bool cone_event(const float4 *words, uint32_t num_stars, uint32_t star, V4 origin, V4 e0, float cone,
                           V4 &d_out, V4 &u_out) {
    float4 hdr = words[star];
    uint32_t first = __float_as_uint(hdr.z), count = __float_as_uint(hdr.w);
    const float4 *segs = words + num_stars;
    if (count == 0) return false;
    uint32_t size = count, base = 0;
    while (size > 1) {
        uint32_t half = size / 2, mid = base + half;
        bool less = past_the_cone(f4v(segs[3 * (first + mid)]), origin, e0, cone) <= 0.0f;
        base = less ? mid : base;
        size -= half;
    }
    bool less = past_the_cone(f4v(segs[3 * (first + base)]), origin, e0, cone) <= 0.0f;
    uint32_t before = base + (less ? 1u : 0u);
    uint32_t idx = (before > 1 ? before : 1) - 1;
    if (idx >= count) return false;
    const float4 *sg = segs + 3 * (first + idx);
    V4 ev = f4v(sg[0]), u = f4v(sg[1]);
    float4 lim = sg[2];
    float s = cone_crossing(ev, u, origin, e0, cone);
    if (!(s >= lim.x && s <= lim.y)) return false;
    d_out = (ev - origin) + u * s;
    u_out = u;
    return true;
}

// Per-pixel part of star_light for one star (star.rs:255-283 with its per-view values
// precomputed): adds to total exactly as the reference's loop body does.
template <bool STATS>
__device__ __forceinline__ void star_pixel(const StarPrep &p, V4 n, float spread, float lambda_max, V3 &total,
                                           Counters &C) {
    if (!(p.flags & 1u)) return;
    if (STATS) C.star_tests++;
    V4 r = f4v(p.r);
    float along = mdot(r, n);
    if (along <= 0.0f || along >= lambda_max) return;
    V4 perp = r - n * along;
    float d2 = fmax_(mdot(perp, perp), 0.0f);
    float sigma = 0.6f * (0.0f + spread * along);
    float cut = 4.0f * sigma;
    if (d2 >= cut * cut) return;
    float inv_2s2 = 1.0f / (2.0f * sigma * sigma);
    float kernel = expf_(-d2 * inv_2s2) * inv_2s2 * 0.318309886183790671538f;
    total += V3{p.bb_cn.x, p.bb_cn.y, p.bb_cn.z} * (p.bb_cn.w * kernel);
}

}  // namespace dev
