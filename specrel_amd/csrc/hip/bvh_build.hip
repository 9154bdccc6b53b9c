// Device BVH build emitting the host builder's tree bit for bit (Bvh4D::build, bvh4d.rs:122-427,
// as restated in scene.cpp, sharing core/bvh_split.h for every split decision).
//
// Large segments (more than SMALL references) are processed breadth first, all segments of a
// level together: bounds and centroid bounds per chunk of CHUNK references, merged per segment;
// 8 bins per axis per chunk (one wave per axis), merged per segment; the split chosen by one
// thread per segment. The partition must be the host's in-place swap partition (bvh4d.rs:409-416):
// references going left keep their order; those going right end up in the order of a queue in
// which each right-going reference is appended as it is met and each left-going reference met
// after the first right-going one moves the queue's front to its back. The final queue is the last
// F entries of the stream of appends, and each re-append copies an earlier stream entry, so the
// stream is filled by pointer jumping (a few passes) after a scan of the left flags. Segments at
// or below SMALL references are then built one wave each from LDS, every decision the host's
// (k_bb_small). Node indices are the host's preorder: subtree sizes are summed bottom-up over the
// large nodes (few) on the host, the small subtrees are copied into place with their offsets.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../core/bvh_split.h"
#include "bvh_build.hpp"
#include "extrude_dev.hpp"

namespace dev {
namespace {

constexpr uint32_t CHUNK = 4096;  // references per chunk of a large segment
constexpr uint32_t SMALL = 256;   // subtrees at or below this size are built by one wave each
constexpr uint32_t MERGE = 8;     // segments of more chunks have their bins merged by a block first
constexpr uint32_t NONE = 0xFFFFFFFFu;
enum { K_SAH = sr::SPLIT_SAH, K_MID = sr::SPLIT_MIDPOINT, K_LEAF = sr::SPLIT_LEAF, K_FALLBACK = 3 };
enum { ERR_E = 1, ERR_JUMP = 2, ERR_SCATTER = 4, ERR_SMALL_STACK = 8, ERR_SMALL_NODES = 16, ERR_PLACE = 32 };

struct Refs {
    float4 *mn, *mx, *c;
    uint32_t *idx;
};
struct Seg {
    uint32_t begin, count, chunk_first, n_chunks;
};
struct Chunk {
    uint32_t seg, begin, end, pad;
};
struct SegInfo {
    float4 bmn, bmx, cmn, cmx;
    float axis_min[4], inv_ext[4], ext[4];
    uint32_t live;  // bit a: axis a has extent
    int32_t kind, axis;
    float pos;
    uint32_t first_false, nT, pad[2];
};
struct BinPart {
    float4 mn[32], mx[32];
    uint32_t n[32];
};
struct SmallTask {
    uint32_t begin, count;
};
struct Place {
    uint32_t scratch, count, dest, pad;
};
struct PutNode {
    uint32_t dest, pad[3];
    specrel_node node;
};

__device__ __host__ __forceinline__ sr::V4 v(float4 f) { return sr::v4(f.x, f.y, f.z, f.w); }
__device__ __host__ __forceinline__ float4 f4(sr::V4 a) { return make_float4(a.t, a.x, a.y, a.z); }
__device__ __forceinline__ float comp(float4 f, int a) { return a == 0 ? f.x : a == 1 ? f.y : a == 2 ? f.z : f.w; }

__device__ __forceinline__ sr::BinFrame frame_of(const SegInfo &s) {
    sr::BinFrame f;
    for (int a = 0; a < 4; ++a) {
        f.live[a] = (s.live >> a) & 1u;
        f.axis_min[a] = s.axis_min[a];
        f.inv_ext[a] = s.inv_ext[a];
        f.ext[a] = s.ext[a];
    }
    return f;
}

__global__ __launch_bounds__(256) void k_bb_init(const float4 *boxes, uint32_t n, Refs r) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 a = boxes[2 * (size_t)i], b = boxes[2 * (size_t)i + 1];
    r.mn[i] = a;
    r.mx[i] = b;
    r.c[i] = f4((v(a) + v(b)) * 0.5f);  // BuildRef centroid (scene.cpp finish_into)
    r.idx[i] = i;
}

__global__ __launch_bounds__(256) void k_bb_bounds(const Chunk *chunks, Refs r, float4 *part) {
    const Chunk ch = chunks[blockIdx.x];
    const uint32_t tid = threadIdx.x;
    sr::Box4 b = sr::box_empty(), cb = sr::box_empty();
    for (uint32_t i = ch.begin + tid; i < ch.end; i += 256) {
        b = sr::box_union(b, sr::Box4{v(r.mn[i]), v(r.mx[i])});
        sr::box_add_point(cb, v(r.c[i]));
    }
    __shared__ float4 s[4][256];
    s[0][tid] = f4(b.mn);
    s[1][tid] = f4(b.mx);
    s[2][tid] = f4(cb.mn);
    s[3][tid] = f4(cb.mx);
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (tid < w) {
            s[0][tid] = f4(sr::vmin(v(s[0][tid]), v(s[0][tid + w])));
            s[1][tid] = f4(sr::vmax(v(s[1][tid]), v(s[1][tid + w])));
            s[2][tid] = f4(sr::vmin(v(s[2][tid]), v(s[2][tid + w])));
            s[3][tid] = f4(sr::vmax(v(s[3][tid]), v(s[3][tid + w])));
        }
        __syncthreads();
    }
    if (tid < 4) part[4 * (size_t)blockIdx.x + tid] = s[tid][0];
}

__global__ __launch_bounds__(256) void k_bb_frame(const Seg *segs, uint32_t nsegs, const float4 *part,
                                                  SegInfo *info) {
    const uint32_t si = blockIdx.x * 256 + threadIdx.x;
    if (si >= nsegs) return;
    const Seg sg = segs[si];
    sr::Box4 b = sr::box_empty(), cb = sr::box_empty();
    for (uint32_t c = sg.chunk_first; c < sg.chunk_first + sg.n_chunks; ++c) {
        b = sr::box_union(b, sr::Box4{v(part[4 * (size_t)c]), v(part[4 * (size_t)c + 1])});
        cb = sr::box_union(cb, sr::Box4{v(part[4 * (size_t)c + 2]), v(part[4 * (size_t)c + 3])});
    }
    SegInfo o;
    o.bmn = f4(b.mn);
    o.bmx = f4(b.mx);
    o.cmn = f4(cb.mn);
    o.cmx = f4(cb.mx);
    const sr::BinFrame f = sr::bin_frame(cb);
    o.live = 0;
    for (int a = 0; a < 4; ++a) {
        o.axis_min[a] = f.axis_min[a];
        o.inv_ext[a] = f.inv_ext[a];
        o.ext[a] = f.ext[a];
        o.live |= (f.live[a] ? 1u : 0u) << a;
    }
    o.kind = -1;
    o.axis = 0;
    o.pos = 0.0f;
    o.first_false = NONE;
    o.nT = 0;
    o.pad[0] = o.pad[1] = 0;
    info[si] = o;
}

// Bins of one chunk: wave a bins axis a; each lane keeps 8 bins in registers, then the wave
// reduces them (min / max / sum: the order does not matter).
__global__ __launch_bounds__(256) void k_bb_bins(const Chunk *chunks, Refs r, const SegInfo *info, BinPart *bp) {
    const Chunk ch = chunks[blockIdx.x];
    const SegInfo &si = info[ch.seg];
    const int a = threadIdx.x >> 6;
    const uint32_t lane = threadIdx.x & 63;
    const float inf = __builtin_inff();
    float4 bmn[8], bmx[8];
    uint32_t bn[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        bmn[k] = make_float4(inf, inf, inf, inf);
        bmx[k] = make_float4(-inf, -inf, -inf, -inf);
        bn[k] = 0;
    }
    if ((si.live >> a) & 1u) {
        const sr::BinFrame f = frame_of(si);
        for (uint32_t i = ch.begin + lane; i < ch.end; i += 64) {
            const uint32_t k = sr::bin_of(f, a, comp(r.c[i], a));
            const float4 mn = r.mn[i], mx = r.mx[i];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (k == (uint32_t)q) {
                    bmn[q] = f4(sr::vmin(v(bmn[q]), v(mn)));
                    bmx[q] = f4(sr::vmax(v(bmx[q]), v(mx)));
                    bn[q] += 1;
                }
        }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        for (int off = 32; off > 0; off >>= 1) {
            float4 m = make_float4(__shfl_xor(bmn[q].x, off), __shfl_xor(bmn[q].y, off), __shfl_xor(bmn[q].z, off),
                                   __shfl_xor(bmn[q].w, off));
            float4 M = make_float4(__shfl_xor(bmx[q].x, off), __shfl_xor(bmx[q].y, off), __shfl_xor(bmx[q].z, off),
                                   __shfl_xor(bmx[q].w, off));
            bmn[q] = f4(sr::vmin(v(bmn[q]), v(m)));
            bmx[q] = f4(sr::vmax(v(bmx[q]), v(M)));
            bn[q] += __shfl_xor(bn[q], off);
        }
    }
    if (lane == 0) {
        BinPart &o = bp[blockIdx.x];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            o.mn[a * 8 + q] = bmn[q];
            o.mx[a * 8 + q] = bmx[q];
            o.n[a * 8 + q] = bn[q];
        }
    }
}

// Bins of a segment of more than MERGE chunks, merged into its first chunk's partial by one block:
// lane q of each 32-lane group takes bin q of every 8th chunk, then the 8 groups are merged.
// Unions are min / max and counts sums, so the order changes nothing.
__global__ __launch_bounds__(256) void k_bb_merge(const uint32_t *big, const Seg *segs, BinPart *bp) {
    const Seg sg = segs[big[blockIdx.x]];
    const uint32_t q = threadIdx.x & 31, g = threadIdx.x >> 5;
    const float inf = __builtin_inff();
    float4 mn = make_float4(inf, inf, inf, inf), mx = make_float4(-inf, -inf, -inf, -inf);
    uint32_t cnt = 0;
    for (uint32_t c = sg.chunk_first + g; c < sg.chunk_first + sg.n_chunks; c += 8) {
        mn = f4(sr::vmin(v(mn), v(bp[c].mn[q])));
        mx = f4(sr::vmax(v(mx), v(bp[c].mx[q])));
        cnt += bp[c].n[q];
    }
    __shared__ float4 smn[8][32], smx[8][32];
    __shared__ uint32_t sn[8][32];
    smn[g][q] = mn;
    smx[g][q] = mx;
    sn[g][q] = cnt;
    __syncthreads();
    if (g == 0) {
        for (int k = 1; k < 8; ++k) {
            mn = f4(sr::vmin(v(mn), v(smn[k][q])));
            mx = f4(sr::vmax(v(mx), v(smx[k][q])));
            cnt += sn[k][q];
        }
        BinPart &o = bp[sg.chunk_first];
        o.mn[q] = mn;
        o.mx[q] = mx;
        o.n[q] = cnt;
    }
}

__global__ __launch_bounds__(64) void k_bb_split(const Seg *segs, uint32_t nsegs, const BinPart *bp, SegInfo *info) {
    const uint32_t si = blockIdx.x * 64 + threadIdx.x;
    if (si >= nsegs) return;
    const Seg sg = segs[si];
    SegInfo &o = info[si];
    sr::Bins B;
    sr::bins_clear(B);
    const uint32_t nch = sg.n_chunks > MERGE ? 1 : sg.n_chunks;  // merged by k_bb_merge
    for (uint32_t c = sg.chunk_first; c < sg.chunk_first + nch; ++c)
        for (int a = 0; a < 4; ++a)
            for (int q = 0; q < 8; ++q) {
                B.b[a][q] = sr::box_union(B.b[a][q], sr::Box4{v(bp[c].mn[a * 8 + q]), v(bp[c].mx[a * 8 + q])});
                B.n[a][q] += bp[c].n[a * 8 + q];
            }
    int axis;
    float pos;
    const sr::Box4 cb{v(o.cmn), v(o.cmx)};
    o.kind = sg.count <= 3 ? K_LEAF : sr::choose_split(cb, frame_of(o), B, axis, pos);
    o.axis = o.kind == K_LEAF ? 0 : axis;
    o.pos = o.kind == K_LEAF ? 0.0f : pos;
}

__global__ __launch_bounds__(256) void k_bb_flags(const Chunk *chunks, const Seg *segs, Refs r, SegInfo *info,
                                                  uint32_t *flag) {
    const Chunk ch = chunks[blockIdx.x];
    SegInfo &si = info[ch.seg];
    const uint32_t begin = segs[ch.seg].begin;
    const bool part = si.kind == K_SAH || si.kind == K_MID;
    const int a = si.axis;
    const float pos = si.pos;
    uint32_t first = NONE;
    for (uint32_t i = ch.begin + threadIdx.x; i < ch.end; i += 256) {
        const bool f = part && comp(r.c[i], a) < pos;
        flag[i] = f ? 1u : 0u;
        if (part && !f) first = min(first, i - begin);
    }
    if (first != NONE) atomicMin(&si.first_false, first);
}

// Exclusive prefix of v over the block (256 threads).
__device__ __forceinline__ uint32_t block_excl(uint32_t x, uint32_t *s_wave, uint32_t &total) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, v0 = x;
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
    return before + x - v0;
}

// Block-local exclusive scan of 1024 flags per block (4 per thread) and the block totals.
__global__ __launch_bounds__(256) void k_bb_scan_local(const uint32_t *flag, uint32_t n, uint32_t *scanv,
                                                       uint32_t *block_sums) {
    const uint32_t base = blockIdx.x * 1024 + threadIdx.x * 4;
    uint32_t f[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        f[k] = base + k < n ? flag[base + k] : 0u;
        s += f[k];
    }
    __shared__ uint32_t s_wave[4];
    uint32_t total;
    uint32_t run = block_excl(s, s_wave, total);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) scanv[base + k] = run;
        run += f[k];
    }
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void k_bb_scan_add(uint32_t *scanv, uint32_t n, const uint64_t *offsets,
                                                     uint32_t nblocks) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) scanv[i] += (uint32_t)offsets[i / 1024];
    if (i == 0) scanv[n] = (uint32_t)offsets[nblocks];
}

__global__ __launch_bounds__(256) void k_bb_counts(const Seg *segs, uint32_t nsegs, const uint32_t *scanv,
                                                   SegInfo *info) {
    const uint32_t si = blockIdx.x * 256 + threadIdx.x;
    if (si >= nsegs) return;
    const Seg sg = segs[si];
    SegInfo &o = info[si];
    const uint32_t nT = scanv[sg.begin + sg.count] - scanv[sg.begin];
    o.nT = nT;
    if ((o.kind == K_SAH || o.kind == K_MID) && (nT == 0 || nT == sg.count)) o.kind = K_FALLBACK;
}

// The stream of queue appends of each partitioned segment, at [begin, begin + L):
// right-going reference with m right-goers and R queue rotations before it -> entry m + R holds
// its position; the j-th rotation, met after A right-goers -> entry A + j copies entry j.
__global__ __launch_bounds__(256) void k_bb_stream(const Chunk *chunks, const Seg *segs, const SegInfo *info,
                                                   const uint32_t *flag, const uint32_t *scanv, uint32_t *src,
                                                   uint32_t *val, uint32_t *err) {
    const Chunk ch = chunks[blockIdx.x];
    const SegInfo &si = info[ch.seg];
    if (si.kind != K_SAH && si.kind != K_MID) return;
    const Seg sg = segs[ch.seg];
    const uint32_t base = scanv[sg.begin], t0 = si.first_false, nT = si.nT;
    const uint32_t L = (sg.count - nT) + (nT - t0);
    for (uint32_t p = ch.begin + threadIdx.x; p < ch.end; p += 256) {
        const uint32_t i = p - sg.begin, tb = scanv[p] - base, fb = i - tb;
        uint32_t e, s, x;
        if (!flag[p]) {
            e = fb + (tb - t0);
            s = NONE;
            x = p;
        } else if (fb > 0) {
            e = fb + (tb - t0);
            s = sg.begin + (tb - t0);
            x = NONE;
        } else {
            continue;
        }
        if (e >= L || tb < t0) {
            atomicOr(err, ERR_E);
            continue;
        }
        src[sg.begin + e] = s;
        val[sg.begin + e] = x;
    }
}

__global__ __launch_bounds__(256) void k_bb_jump(const Chunk *chunks, const Seg *segs, const SegInfo *info,
                                                 uint32_t *src, uint32_t *val, uint32_t *unresolved, uint32_t *err) {
    const Chunk ch = chunks[blockIdx.x];
    const SegInfo &si = info[ch.seg];
    if (si.kind != K_SAH && si.kind != K_MID) return;
    const Seg sg = segs[ch.seg];
    const uint32_t L = (sg.count - si.nT) + (si.nT - si.first_false);
    uint32_t left = 0;
    for (uint32_t p = ch.begin + threadIdx.x; p < ch.end && p - sg.begin < L; p += 256) {
        if (val[p] != NONE) continue;
        const uint32_t s = src[p];
        if (s < sg.begin || s >= p) {
            atomicOr(err, ERR_JUMP);
            continue;
        }
        const uint32_t x = val[s];
        if (x != NONE) {
            val[p] = x;
        } else {
            src[p] = src[s];
            ++left;
        }
    }
    if (left) atomicAdd(unresolved, left);
}

__device__ __forceinline__ void move_ref(const Refs &from, uint32_t s, const Refs &to, uint32_t d) {
    to.mn[d] = from.mn[s];
    to.mx[d] = from.mx[s];
    to.c[d] = from.c[s];
    to.idx[d] = from.idx[s];
}

__global__ __launch_bounds__(256) void k_bb_scatter(const Chunk *chunks, const Seg *segs, const SegInfo *info,
                                                    const uint32_t *flag, const uint32_t *scanv, const uint32_t *val,
                                                    Refs cur, Refs nxt, Refs fin, uint32_t *err) {
    const Chunk ch = chunks[blockIdx.x];
    const SegInfo &si = info[ch.seg];
    const Seg sg = segs[ch.seg];
    if (si.kind == K_LEAF) {
        for (uint32_t p = ch.begin + threadIdx.x; p < ch.end; p += 256) move_ref(cur, p, fin, p);
        return;
    }
    if (si.kind != K_SAH && si.kind != K_MID) return;
    const uint32_t base = scanv[sg.begin], nT = si.nT, t0 = si.first_false;
    const uint32_t r = nT - t0, nF = sg.count - nT, end = sg.begin + sg.count;
    const Refs &lo = nT <= SMALL ? fin : nxt, &hi = nF <= SMALL ? fin : nxt;
    for (uint32_t p = ch.begin + threadIdx.x; p < ch.end; p += 256) {
        const uint32_t i = p - sg.begin;
        if (flag[p]) {
            const uint32_t d = sg.begin + (scanv[p] - base);
            if (d >= sg.begin + nT) atomicOr(err, ERR_SCATTER);
            else move_ref(cur, p, lo, d);
        }
        if (i >= r && i < r + nF) {
            const uint32_t q = val[p], d = sg.begin + nT + (i - r);
            if (q < sg.begin || q >= end || d >= end) atomicOr(err, ERR_SCATTER);
            else move_ref(cur, q, hi, d);
        }
    }
}

// Fallback partition of one segment: dest position begin + k takes the reference at begin + perm[k].
__global__ __launch_bounds__(256) void k_bb_permute(uint32_t begin, uint32_t count, uint32_t mid,
                                                    const uint32_t *perm, Refs cur, Refs nxt, Refs fin,
                                                    uint32_t *err) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= count) return;
    const uint32_t s = perm[k];
    if (s >= count) {
        atomicOr(err, ERR_SCATTER);
        return;
    }
    const bool small = k < mid ? mid <= SMALL : count - mid <= SMALL;
    move_ref(cur, begin + s, small ? fin : nxt, begin + k);
}

// A subtree of at most SMALL references, built by one wave exactly as the host builds it
// (scene.cpp split / build_recursive), its references staged in LDS:
//  - a node's bounds and centroid bounds: 16 lanes, one component each, fold the references in
//    order (the host's sequence of min / max operations per component);
//  - the SAH bins: lane (axis, bin, min or max half) folds, in order, the references of its bin
//    (the host's sequence per bin); every lane then sweeps the bins from LDS (the same sweep);
//  - the host's in-place swap partition (bvh4d.rs:409-416): left-going references keep their
//    order (prefix counts); right-going ones end in the order of a queue in which each one is
//    appended as met and each left-going reference met after the first right-going one moves the
//    queue's front to its back. After the first right-going reference f, step k appends stream
//    entry k - f: reference k itself if it goes right, else the entry (number of right-going
//    references before k) places earlier; the rights' final positions p hold entries p - f,
//    resolved by pointer jumping;
//  - the stable-sort fallback is run by lane 0 as the host runs it.
// Nodes in preorder to scratch[2 * begin ...] with subtree-local child indices; the references
// are written back in their final order. One wave per subtree (one block of 64 threads).
__device__ __forceinline__ float comp_lds(const float4 *a, uint32_t i, int c) { return ((const float *)(a + i))[c]; }
// arr[eb + p] = arr[eb + src[p]] for p < ec <= 256 (one wave: every lane gathers four positions,
// then writes them back).
template <class T>
__device__ __forceinline__ void permute4(T *arr, const uint32_t *src, uint32_t eb, uint32_t ec, uint32_t lane) {
    const uint32_t p0 = lane, p1 = lane + 64u, p2 = lane + 128u, p3 = lane + 192u;
    const T v0 = arr[eb + src[p0 < ec ? p0 : 0u]], v1 = arr[eb + src[p1 < ec ? p1 : 0u]];
    const T v2 = arr[eb + src[p2 < ec ? p2 : 0u]], v3 = arr[eb + src[p3 < ec ? p3 : 0u]];
    __syncthreads();
    if (p0 < ec) arr[eb + p0] = v0;
    if (p1 < ec) arr[eb + p1] = v1;
    if (p2 < ec) arr[eb + p2] = v2;
    if (p3 < ec) arr[eb + p3] = v3;
    __syncthreads();
}

__global__ __launch_bounds__(64) void k_bb_small(const SmallTask *tasks, uint32_t ntasks, Refs f,
                                                 specrel_node *scratch, uint32_t *task_nodes, uint32_t *err) {
    const uint32_t t = blockIdx.x, lane = threadIdx.x;
    if (t >= ntasks) return;
    const SmallTask task = tasks[t];
    const uint32_t n0 = task.count, g0 = task.begin;
    if (n0 > SMALL) {
        if (lane == 0) atomicOr(err, ERR_SMALL_STACK);
        return;
    }
    specrel_node *out = scratch + 2 * (size_t)g0;
    const uint32_t cap = 2 * n0;
    // A pending range (first reference, local to the task; count <= SMALL) and the node whose right
    // child it becomes (a subtree-local index < 2 * SMALL, or NO_PARENT).
    struct Ent {
        uint32_t b;
        uint16_t c, parent;
    };
    static_assert(SMALL < 0xFFFFu && 2 * SMALL < 0xFFFFu, "Ent fields");
    static_assert(sr::BVH_BINS == 8, "one lane per axis, bin and half");
    constexpr uint16_t NO_PARENT = 0xFFFFu;
    constexpr uint32_t RES = 0x80000000u;  // a resolved stream entry: the reference it holds
    __shared__ float4 s_mn[SMALL], s_mx[SMALL], s_c[SMALL];
    __shared__ uint32_t s_idx[SMALL], s_src[SMALL], s_ptr[SMALL];
    __shared__ Ent s_stack[SMALL + 2];
    __shared__ float4 s_bb[4][8][2];
    __shared__ uint32_t s_bn[4][8];
    __shared__ float s_nb[16];
    for (uint32_t i = lane; i < n0; i += 64) {
        s_mn[i] = f.mn[g0 + i];
        s_mx[i] = f.mx[g0 + i];
        s_c[i] = f.c[g0 + i];
        s_idx[i] = f.idx[g0 + i];
    }
    if (lane == 0) s_stack[0] = Ent{0u, (uint16_t)n0, NO_PARENT};
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    const float inf = __builtin_inff();
    int sp = 1;
    uint32_t nn = 0;
    while (sp > 0) {
        const Ent e = s_stack[--sp];
        const uint32_t me = nn++;
        if (me >= cap) {
            if (lane == 0) atomicOr(err, ERR_SMALL_NODES);
            return;
        }
        if (lane == 0 && e.parent != NO_PARENT) out[e.parent].right_or_first = me;
        const uint32_t eb = e.b, ec = e.c, ee = eb + ec;
        // Bounds (components 0-7: the references' box union) and centroid bounds (8-15).
        if (lane < 16) {
            const int c = lane & 3;
            const float4 *arr = lane < 4 ? s_mn : lane < 8 ? s_mx : s_c;
            const bool is_min = lane < 4 || (lane >= 8 && lane < 12);
            float acc = is_min ? inf : -inf;
#pragma unroll 8
            for (uint32_t i = eb; i < ee; ++i) {
                const float x = comp_lds(arr, i, c);
                acc = is_min ? sr::fmin_(acc, x) : sr::fmax_(acc, x);
            }
            s_nb[lane] = acc;
        }
        __syncthreads();
        const sr::Box4 bounds = sr::Box4{sr::v4(s_nb[0], s_nb[1], s_nb[2], s_nb[3]), sr::v4(s_nb[4], s_nb[5], s_nb[6], s_nb[7])};
        const sr::Box4 cb = sr::Box4{sr::v4(s_nb[8], s_nb[9], s_nb[10], s_nb[11]), sr::v4(s_nb[12], s_nb[13], s_nb[14], s_nb[15])};
        specrel_node nd;
        sr::put4(nd.min, bounds.mn);
        sr::put4(nd.max, bounds.mx);
        nd.pad = 0;
        int kind = K_LEAF, axis = 0;
        float pos = 0.0f;
        if (ec > 3) {
            const sr::BinFrame fr = sr::bin_frame(cb);
            {
                const int a = (int)(lane >> 4), q = (int)((lane >> 1) & 7u), h = (int)(lane & 1u);
                const float4 *arr = h ? s_mx : s_mn;
                float4 acc = h ? make_float4(-inf, -inf, -inf, -inf) : make_float4(inf, inf, inf, inf);
                uint32_t cnt = 0;
                // the lane's axis of the frame by selects (no dynamically indexed copy in private
                // memory); the bin is sr::bin_of's expression
                const bool live = a == 0 ? fr.live[0] : a == 1 ? fr.live[1] : a == 2 ? fr.live[2] : fr.live[3];
                const float am = a == 0 ? fr.axis_min[0] : a == 1 ? fr.axis_min[1] : a == 2 ? fr.axis_min[2] : fr.axis_min[3];
                const float ie = a == 0 ? fr.inv_ext[0] : a == 1 ? fr.inv_ext[1] : a == 2 ? fr.inv_ext[2] : fr.inv_ext[3];
                if (live) {
#pragma unroll 4
                    for (uint32_t i = eb; i < ee; ++i) {
                        const uint64_t k64 = sr::f2u_sat((comp_lds(s_c, i, a) - am) * ie);
                        const uint32_t k = k64 > (uint64_t)(sr::BVH_BINS - 1) ? (uint32_t)(sr::BVH_BINS - 1) : (uint32_t)k64;
                        const float4 x = arr[i];
                        if (k == (uint32_t)q) {
                            acc = h ? f4(sr::vmax(v(acc), v(x))) : f4(sr::vmin(v(acc), v(x)));
                            cnt += 1;
                        }
                    }
                }
                s_bb[a][q][h] = acc;
                if (!h) s_bn[a][q] = cnt;
            }
            __syncthreads();
            // The SAH sweep (sr::sah_axis over the live axes in order), one lane per candidate:
            // lane 7a + j is axis a's split after bin i = 6 - j, the host's visiting order. Each
            // lane forms its left box from bins 0..i and its right box from bins 7..i+1 in the
            // host's order of unions, and its cost with the host's expression; the host keeps the
            // first candidate strictly cheaper than every one before it, which is the first in
            // that order with the least cost (candidates without both sides, or with a cost that
            // is not below +inf, never win).
            float cost = inf, cpos = 0.0f;
            bool elig = false;
            if (lane < 28) {
                const int a = (int)(lane / 7u), i = 6 - (int)(lane % 7u);
                const bool live = a == 0 ? fr.live[0] : a == 1 ? fr.live[1] : a == 2 ? fr.live[2] : fr.live[3];
                if (live) {
                    sr::Box4 lb = sr::box_empty(), rb = sr::box_empty();
                    uint32_t lc = 0, rc = 0;
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        if (q <= i) {
                            lb = sr::box_union(lb, sr::Box4{v(s_bb[a][q][0]), v(s_bb[a][q][1])});
                            lc += s_bn[a][q];
                        }
                    }
#pragma unroll
                    for (int q = 7; q >= 0; --q) {
                        if (q > i) {
                            rb = sr::box_union(rb, sr::Box4{v(s_bb[a][q][0]), v(s_bb[a][q][1])});
                            rc += s_bn[a][q];
                        }
                    }
                    if (lc > 0 && rc > 0) {
                        cost = sr::sah_cost(lb, lc, rb, rc);
                        elig = cost < inf;
                        const float am = a == 0 ? fr.axis_min[0] : a == 1 ? fr.axis_min[1] : a == 2 ? fr.axis_min[2] : fr.axis_min[3];
                        const float ex = a == 0 ? fr.ext[0] : a == 1 ? fr.ext[1] : a == 2 ? fr.ext[2] : fr.ext[3];
                        cpos = sr::sah_split_pos(am, ex, i);
                    }
                }
            }
            float m = elig ? cost : inf;
#pragma unroll
            for (int off = 16; off > 0; off >>= 1) m = fminf(m, __shfl_xor(m, off, 64));
            m = __shfl(m, 0, 64);  // lanes 0..31 hold every candidate
            const uint64_t win = __builtin_amdgcn_ballot_w64(elig && cost == m);
            float best_cost = inf;
            if (win) {
                const int w = __builtin_ctzll(win);
                best_cost = m;
                axis = w / 7;
                pos = __shfl(cpos, w, 64);
            }
            kind = sr::split_after_sweeps(cb, fr, best_cost, axis, pos);
        }
        if (kind == K_LEAF) {
            if (lane == 0) {
                nd.right_or_first = g0 + eb;
                nd.count = ec;
                nd.left = 0;
                out[me] = nd;
            }
            __syncthreads();
            continue;
        }
        // Partition: flags, prefix counts of right-going references, the first right-going one.
        bool goes_left[4];
        uint32_t r_before[4];
        uint32_t n_left = 0, n_right = 0, first_r = 0xFFFFFFFFu;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t k = 64u * j + lane;
            const bool in = k < ec;
            goes_left[j] = in && comp_lds(s_c, eb + k, axis) < pos;
            const uint64_t mr = __builtin_amdgcn_ballot_w64(in && !goes_left[j]);
            const uint64_t ml = __builtin_amdgcn_ballot_w64(goes_left[j]);
            r_before[j] = n_right + (uint32_t)__popcll(mr & below);
            if (first_r == 0xFFFFFFFFu && mr) first_r = 64u * j + (uint32_t)__builtin_ctzll(mr);
            n_right += (uint32_t)__popcll(mr);
            n_left += (uint32_t)__popcll(ml);
        }
        uint32_t mid = n_left;
        if (mid == 0 || mid == ec) {
            // stable insertion sort by the centroid on the axis, as the host does
            if (lane == 0) {
                for (uint32_t i = eb + 1; i < ee; ++i) {
                    const float4 k0 = s_mn[i], k1 = s_mx[i], k2 = s_c[i];
                    const uint32_t k3 = s_idx[i];
                    const float key = comp(k2, axis);
                    uint32_t j = i;
                    while (j > eb && key < comp_lds(s_c, j - 1, axis)) {
                        s_mn[j] = s_mn[j - 1];
                        s_mx[j] = s_mx[j - 1];
                        s_c[j] = s_c[j - 1];
                        s_idx[j] = s_idx[j - 1];
                        --j;
                    }
                    s_mn[j] = k0;
                    s_mx[j] = k1;
                    s_c[j] = k2;
                    s_idx[j] = k3;
                }
            }
            mid = ec / 2;
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t k = 64u * j + lane;
                if (k >= ec) continue;
                if (goes_left[j]) s_src[k - r_before[j]] = k;
                if (k >= first_r) {
                    const uint32_t st = k - first_r;
                    s_ptr[st] = goes_left[j] ? st - r_before[j] : (RES | k);
                }
            }
            __syncthreads();
            const uint32_t ns = ec - first_r;
            bool pending = true;
            for (int round = 0; round < 10 && pending; ++round) {
                bool mine = false;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t st = 64u * j + lane;
                    if (st >= ns) continue;
                    const uint32_t p = s_ptr[st];
                    if (!(p & RES)) {
                        const uint32_t q = s_ptr[p];
                        s_ptr[st] = q;
                        mine = mine || !(q & RES);
                    }
                }
                __syncthreads();
                pending = __builtin_amdgcn_ballot_w64(mine) != 0;
            }
            if (pending) {
                if (lane == 0) atomicOr(err, ERR_SMALL_STACK);
                return;
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t p = 64u * j + lane;
                if (p >= mid && p < ec) s_src[p] = s_ptr[p - first_r] & ~RES;
            }
            __syncthreads();
            permute4(s_mn, s_src, eb, ec, lane);
            permute4(s_mx, s_src, eb, ec, lane);
            permute4(s_c, s_src, eb, ec, lane);
            permute4(s_idx, s_src, eb, ec, lane);
        }
        if (sp + 2 > (int)SMALL + 2) {
            if (lane == 0) atomicOr(err, ERR_SMALL_STACK);
            return;
        }
        if (lane == 0) {
            nd.right_or_first = 0;
            nd.count = 0;
            nd.left = me + 1;
            out[me] = nd;
            s_stack[sp] = Ent{eb + mid, (uint16_t)(ec - mid), (uint16_t)me};
            s_stack[sp + 1] = Ent{eb, (uint16_t)mid, NO_PARENT};
        }
        sp += 2;
        __syncthreads();
    }
    for (uint32_t i = lane; i < n0; i += 64) {
        f.mn[g0 + i] = s_mn[i];
        f.mx[g0 + i] = s_mx[i];
        f.c[g0 + i] = s_c[i];
        f.idx[g0 + i] = s_idx[i];
    }
    if (lane == 0) task_nodes[t] = nn;
}

__global__ __launch_bounds__(256) void k_bb_place(const Place *places, uint32_t nplaces, const specrel_node *scratch,
                                                  specrel_node *nodes, uint32_t total, uint32_t *err) {
    const Place pl = places[blockIdx.x];
    for (uint32_t j = threadIdx.x; j < pl.count; j += 256) {
        specrel_node nd = scratch[pl.scratch + j];
        if (nd.count == 0) {
            nd.left += pl.dest;
            nd.right_or_first += pl.dest;
        }
        if (pl.dest + j >= total) atomicOr(err, ERR_PLACE);
        else nodes[pl.dest + j] = nd;
    }
}

__global__ __launch_bounds__(256) void k_bb_put(const PutNode *puts, uint32_t n, specrel_node *nodes, uint32_t total,
                                                uint32_t *err) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (puts[i].dest >= total) atomicOr(err, ERR_PLACE);
    else nodes[puts[i].dest] = puts[i].node;
}

// Device allocations of one build, freed on every exit.
struct Mem {
    std::vector<void *> ptrs;
    ~Mem() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <class T>
    T *get(size_t count) {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
            (void)hipGetLastError();
            throw std::string("hipMalloc failed in the BVH build");
        }
        ptrs.push_back(p);
        return (T *)p;
    }
};

void ok(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::string(what) + ": " + hipGetErrorString(e);
}

template <class T>
void put(T *d, const std::vector<T> &h, hipStream_t s) {
    if (!h.empty()) ok(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
}

struct BigNode {
    int kind;  // 0 interior, 1 leaf (large), 2 small subtree
    uint32_t left, right, first, count, task;
    float4 mn, mx;
};

}  // namespace

std::string build_bvh(const float4 *d_boxes, uint32_t n, specrel_node *d_nodes, uint32_t *d_order, uint32_t *n_nodes,
                      hipStream_t s, BvhBuildStats *stats) {
    BvhBuildStats st{};
    auto t_start = std::chrono::steady_clock::now();
    try {
        if (n == 0) {
            *n_nodes = 0;
            return "";
        }
        if (n > 0x7FFFFFFFu) return "too many records for the device build";
        Mem m;
        Refs A{m.get<float4>(n), m.get<float4>(n), m.get<float4>(n), m.get<uint32_t>(n)};
        Refs B{m.get<float4>(n), m.get<float4>(n), m.get<float4>(n), m.get<uint32_t>(n)};
        Refs Fn{m.get<float4>(n), m.get<float4>(n), m.get<float4>(n), d_order};
        uint32_t *flag = m.get<uint32_t>(n), *scanv = m.get<uint32_t>(n + 1);
        uint32_t *src = m.get<uint32_t>(n), *val = m.get<uint32_t>(n);
        const uint32_t nsb = (n + 1023) / 1024;
        uint32_t *bsum = m.get<uint32_t>(nsb);
        uint64_t *boff = m.get<uint64_t>(nsb + 1);
        uint32_t *d_err = m.get<uint32_t>(2);  // [0] error bits, [1] unresolved count
        specrel_node *scratch = m.get<specrel_node>(2 * (size_t)n);
        ok(hipMemsetAsync(d_err, 0, 8, s), "clear");
        k_bb_init<<<(n + 255) / 256, 256, 0, s>>>(d_boxes, n, A);
        ok(hipGetLastError(), "k_bb_init");

        std::vector<BigNode> nodes;  // breadth-first ids
        std::vector<SmallTask> small;
        struct Pending {
            uint32_t id, begin, count;
        };
        std::vector<Pending> segs;
        nodes.push_back(BigNode{});
        if (n <= SMALL) {
            nodes[0].kind = 2;
            nodes[0].task = 0;
            small.push_back(SmallTask{0, n});
            ok(hipMemcpyAsync(Fn.mn, A.mn, (size_t)n * 16, hipMemcpyDeviceToDevice, s), "copy");
            ok(hipMemcpyAsync(Fn.mx, A.mx, (size_t)n * 16, hipMemcpyDeviceToDevice, s), "copy");
            ok(hipMemcpyAsync(Fn.c, A.c, (size_t)n * 16, hipMemcpyDeviceToDevice, s), "copy");
            ok(hipMemcpyAsync(Fn.idx, A.idx, (size_t)n * 4, hipMemcpyDeviceToDevice, s), "copy");
        } else {
            segs.push_back(Pending{0, 0, n});
        }
        Refs cur = A, nxt = B;
        // Per-level device tables, sized for the worst level.
        const size_t max_chunks = n / CHUNK + n / SMALL + 2, max_segs = n / SMALL + 1;
        Seg *d_segs = m.get<Seg>(max_segs);
        Chunk *d_chunks = m.get<Chunk>(max_chunks);
        float4 *d_part = m.get<float4>(4 * max_chunks);
        BinPart *d_bins = m.get<BinPart>(max_chunks);
        SegInfo *d_info = m.get<SegInfo>(max_segs);
        uint32_t *d_perm = m.get<uint32_t>(n);
        uint32_t *d_big = m.get<uint32_t>(max_segs);
        std::vector<Seg> hs;
        std::vector<Chunk> hc;
        std::vector<SegInfo> hi;
        std::vector<uint32_t> hbig;
        while (!segs.empty()) {
            st.levels += 1;
            st.large_segments += (uint32_t)segs.size();
            hs.clear();
            hc.clear();
            hbig.clear();
            for (uint32_t k = 0; k < segs.size(); ++k) {
                const Pending &p = segs[k];
                Seg sg{p.begin, p.count, (uint32_t)hc.size(), (p.count + CHUNK - 1) / CHUNK};
                if (sg.n_chunks > MERGE) hbig.push_back(k);
                for (uint32_t c = 0; c < sg.n_chunks; ++c)
                    hc.push_back(Chunk{k, p.begin + c * CHUNK, std::min(p.begin + p.count, p.begin + (c + 1) * CHUNK), 0});
                hs.push_back(sg);
            }
            const uint32_t ns = (uint32_t)hs.size(), nc = (uint32_t)hc.size();
            put(d_segs, hs, s);
            put(d_chunks, hc, s);
            k_bb_bounds<<<nc, 256, 0, s>>>(d_chunks, cur, d_part);
            k_bb_frame<<<(ns + 255) / 256, 256, 0, s>>>(d_segs, ns, d_part, d_info);
            k_bb_bins<<<nc, 256, 0, s>>>(d_chunks, cur, d_info, d_bins);
            if (!hbig.empty()) {
                put(d_big, hbig, s);
                k_bb_merge<<<(uint32_t)hbig.size(), 256, 0, s>>>(d_big, d_segs, d_bins);
            }
            k_bb_split<<<(ns + 63) / 64, 64, 0, s>>>(d_segs, ns, d_bins, d_info);
            k_bb_flags<<<nc, 256, 0, s>>>(d_chunks, d_segs, cur, d_info, flag);
            k_bb_scan_local<<<nsb, 256, 0, s>>>(flag, n, scanv, bsum);
            k_extrude_scan<<<1, 1024, 0, s>>>(bsum, nsb, boff);
            k_bb_scan_add<<<(n + 255) / 256, 256, 0, s>>>(scanv, n, boff, nsb);
            k_bb_counts<<<(ns + 255) / 256, 256, 0, s>>>(d_segs, ns, scanv, d_info);
            k_bb_stream<<<nc, 256, 0, s>>>(d_chunks, d_segs, d_info, flag, scanv, src, val, d_err);
            ok(hipGetLastError(), "level kernels");
            // Pointer jumping until every stream entry is resolved.
            for (int round = 0;; ++round) {
                if (round >= 16) throw std::string("partition streams did not resolve");
                for (int k = 0; k < 4; ++k) {
                    if (k == 3) ok(hipMemsetAsync(d_err + 1, 0, 4, s), "clear");  // count the last pass only
                    k_bb_jump<<<nc, 256, 0, s>>>(d_chunks, d_segs, d_info, src, val, d_err + 1, d_err);
                    st.jump_passes += 1;
                }
                uint32_t left = 0;
                ok(hipMemcpyAsync(&left, d_err + 1, 4, hipMemcpyDeviceToHost, s), "readback");
                ok(hipStreamSynchronize(s), "jump");
                if (left == 0) break;
            }
            k_bb_scatter<<<nc, 256, 0, s>>>(d_chunks, d_segs, d_info, flag, scanv, val, cur, nxt, Fn, d_err);
            ok(hipGetLastError(), "k_bb_scatter");
            hi.resize(ns);
            ok(hipMemcpyAsync(hi.data(), d_info, ns * sizeof(SegInfo), hipMemcpyDeviceToHost, s), "readback");
            ok(hipStreamSynchronize(s), "level");
            std::vector<Pending> next;
            for (uint32_t k = 0; k < ns; ++k) {
                const Pending &p = segs[k];
                const SegInfo &o = hi[k];
                BigNode &bn = nodes[p.id];
                bn.mn = o.bmn;
                bn.mx = o.bmx;
                if (o.kind == K_LEAF) {
                    bn.kind = 1;
                    bn.first = p.begin;
                    bn.count = p.count;
                    continue;
                }
                uint32_t mid = o.nT;
                if (o.kind == K_FALLBACK) {
                    // Stable sort by the centroid on the axis (scene.cpp split), on the host.
                    st.fallback_segments += 1;
                    std::vector<float4> c(p.count);
                    ok(hipMemcpy(c.data(), cur.c + p.begin, p.count * sizeof(float4), hipMemcpyDeviceToHost), "fallback");
                    std::vector<uint32_t> perm(p.count);
                    std::iota(perm.begin(), perm.end(), 0u);
                    const int ax = o.axis;
                    auto key = [&](uint32_t i) { return sr::axis(v(c[i]), ax); };
                    std::stable_sort(perm.begin(), perm.end(), [&](uint32_t x, uint32_t y) { return key(x) < key(y); });
                    mid = p.count / 2;
                    ok(hipMemcpyAsync(d_perm, perm.data(), p.count * 4, hipMemcpyHostToDevice, s), "fallback");
                    k_bb_permute<<<(p.count + 255) / 256, 256, 0, s>>>(p.begin, p.count, mid, d_perm, cur, nxt, Fn,
                                                                       d_err);
                    ok(hipStreamSynchronize(s), "fallback");
                }
                bn.kind = 0;
                bn.left = (uint32_t)nodes.size();
                bn.right = bn.left + 1;
                const uint32_t lb = p.begin, lc = mid, rb = p.begin + mid, rc = p.count - mid, lid = bn.left,
                               rid = bn.right;
                nodes.push_back(BigNode{});
                nodes.push_back(BigNode{});
                for (int side = 0; side < 2; ++side) {
                    const uint32_t id = side ? rid : lid, b = side ? rb : lb, c = side ? rc : lc;
                    if (c <= SMALL) {
                        nodes[id].kind = 2;
                        nodes[id].task = (uint32_t)small.size();
                        small.push_back(SmallTask{b, c});
                    } else {
                        next.push_back(Pending{id, b, c});
                    }
                }
            }
            segs.swap(next);
            std::swap(cur, nxt);
        }
        // Small subtrees.
        const uint32_t nt = (uint32_t)small.size();
        st.small_subtrees = nt;
        SmallTask *d_tasks = m.get<SmallTask>(nt);
        uint32_t *d_task_nodes = m.get<uint32_t>(nt);
        put(d_tasks, small, s);
        if (nt) {  // every segment may have ended as a leaf of the level phase
            k_bb_small<<<nt, 64, 0, s>>>(d_tasks, nt, Fn, scratch, d_task_nodes, d_err);
            ok(hipGetLastError(), "k_bb_small");
        }
        std::vector<uint32_t> tn(nt);
        ok(hipMemcpyAsync(tn.data(), d_task_nodes, nt * 4, hipMemcpyDeviceToHost, s), "readback");
        uint32_t herr[2] = {0, 0};
        ok(hipMemcpyAsync(herr, d_err, 8, hipMemcpyDeviceToHost, s), "readback");
        ok(hipStreamSynchronize(s), "small subtrees");
        if (herr[0]) return "device BVH build consistency check failed (code " + std::to_string(herr[0]) + ")";
        // Preorder numbering: children have larger breadth-first ids than their parents.
        const size_t nb = nodes.size();
        std::vector<uint64_t> size(nb), pre(nb);
        for (size_t k = nb; k-- > 0;) {
            const BigNode &b = nodes[k];
            size[k] = b.kind == 0 ? 1 + size[b.left] + size[b.right] : b.kind == 1 ? 1 : tn[b.task];
        }
        const uint64_t total = size[0];
        if (total > 2 * (uint64_t)n) return "device BVH build produced too many nodes";
        pre[0] = 0;
        std::vector<PutNode> puts;
        std::vector<Place> places;
        for (size_t k = 0; k < nb; ++k) {
            const BigNode &b = nodes[k];
            if (b.kind == 0) {
                pre[b.left] = pre[k] + 1;
                pre[b.right] = pre[k] + 1 + size[b.left];
            }
            if (b.kind == 2) {
                places.push_back(Place{2 * small[b.task].begin, tn[b.task], (uint32_t)pre[k], 0});
                continue;
            }
            PutNode pn;
            std::memset(&pn, 0, sizeof pn);
            pn.dest = (uint32_t)pre[k];
            sr::put4(pn.node.min, v(b.mn));
            sr::put4(pn.node.max, v(b.mx));
            if (b.kind == 0) {
                pn.node.left = (uint32_t)pre[b.left];
                pn.node.right_or_first = (uint32_t)pre[b.right];
                pn.node.count = 0;
            } else {
                pn.node.right_or_first = b.first;
                pn.node.count = b.count;
                pn.node.left = 0;
            }
            puts.push_back(pn);
        }
        PutNode *d_puts = m.get<PutNode>(puts.size());
        Place *d_places = m.get<Place>(places.size());
        put(d_puts, puts, s);
        put(d_places, places, s);
        if (!puts.empty())
            k_bb_put<<<(uint32_t)((puts.size() + 255) / 256), 256, 0, s>>>(d_puts, (uint32_t)puts.size(), d_nodes,
                                                                            (uint32_t)total, d_err);
        if (!places.empty())
            k_bb_place<<<(uint32_t)places.size(), 256, 0, s>>>(d_places, (uint32_t)places.size(), scratch, d_nodes,
                                                               (uint32_t)total, d_err);
        ok(hipGetLastError(), "placement");
        ok(hipMemcpyAsync(herr, d_err, 8, hipMemcpyDeviceToHost, s), "readback");
        ok(hipStreamSynchronize(s), "placement");
        if (herr[0]) return "device BVH build consistency check failed (code " + std::to_string(herr[0]) + ")";
        *n_nodes = (uint32_t)total;
    } catch (const std::string &e) {
        return e;
    }
    st.ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    if (stats) *stats = st;
    return "";
}

}  // namespace dev
