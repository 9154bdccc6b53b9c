// C ABI of the gfx950 renderer (include/specrel.h): context, scene upload with re-packing
// into the device layout, render, timings. The counterpart of GpuRenderer (src/gpu.rs:432-901).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../core/scene_file.hpp"
#include "../core/tables.h"
#include "bvh_build.hpp"
#include "extrude_dev.hpp"
#include "kernels.hpp"
#include "specrel.h"

using namespace dev;

// The fixed scratch area of specrel_render_noalloc: every intermediate buffer of a one-view render
// (frame and view table, per-frame counters, hit slots, occlusion bits, the HDR pyramid, star
// lists and cones, the stars' per-view values) is carved from it, so that render allocates
// nothing. Each GPU holds its own copy for the life of the process; host code reaches it only
// through the address hipGetSymbolAddress returns for that GPU (g_scratch below).
static __device__ __align__(256) unsigned char scratch[SPECREL_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int SCRATCH_DEVICES = 64;
std::mutex g_scratch_lock;
void *g_scratch[SCRATCH_DEVICES] = {};  // device address of `scratch` on each GPU, looked up once

// Device allocation owned by one holder: freed when the holder goes (also on the throw paths of
// the entry points), moved but never copied.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr, o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            if (p) (void)hipFree(p);
            p = o.p, bytes = o.bytes;
            o.p = nullptr, o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

struct SceneSlot {
    bool used = false;
    DevBuf nbox, ninfo, big, geoms, attrs, order, mats, stars;
    Scene S{};
    uint32_t num_stars = 0, num_star_words = 0, num_materials = 0, num_nodes = 0, num_prims = 0;
    // Host copies for scene patching: each node's parent slot (parent << 1 | right side; ~0u for
    // the root and for nodes the tree does not reach) and the encoded child entries and large
    // leaves as uploaded.
    std::vector<uint32_t> parent;
    std::vector<uint2> ninfo_host, big_host;
    // The BVH version (bvh4d.rs:47-61) the device arrays hold, when the caller gave one.
    bool versioned = false;
    uint64_t layout = 0, edit = 0;
    // The tree's top levels for kernels that copy them into LDS (Scene::top_idx / top_info):
    // top_count[l] = interior nodes of depth below l, in breadth-first order.
    DevBuf top_idx, top_info;
    uint32_t top_count[TOP_LEVELS_MAX + 1] = {};
};

struct Fail {
    int code;
    std::string msg;
};

}  // namespace

struct specrel_ctx {
    int device = 0;
    std::string name, err;
    hipStream_t stream = nullptr;
    std::vector<SceneSlot> scenes;
    // Per-launch view tables (prefix sums, see dev::Views) and the buffers sized for them.
    std::vector<uint32_t> tab_host;
    std::vector<uint32_t> sizes;  // w, h per view of the last render
    uint32_t n_views = 0, total_tiles = 0, total_supers = 0, total_pix = 0, total_pyr = 0;
    uint32_t stars_for = 0;  // num_stars the star buffers are sized for (per view)
    DevBuf tab, frames, hdr, rgb, hit, lam, tau, slot, prep, dirs, super_lists, super_counts, tables, stats;
    DevBuf tile_cost, tile_order;  // per tile quarter: last traversal cost; tiles by decreasing cost
    DevBuf tile_cones;             // per tile: the star cone of the shading pass's culling
    DevBuf counters, mask, squeue;
    DevBuf touch;  // instrumented renders: bitmaps of the distinct nodes, primitives and hits touched
    DevBuf wave_clock;          // SPECREL_OPT_WAVE_CLOCK: the last render's per-wave records
    uint64_t wave_clock_n = 0;  // records the last render wrote
    int clock_khz = 0;          // rate of the device clock the records use
    uint32_t touch_words[3] = {};  // words of each bitmap of the last instrumented render
    // Pinned staging of the frame uploads, a ring of slots so that renders can be enqueued ahead of
    // the device: a slot is refilled once its previous copy has completed.
    struct Stage {
        void *host = nullptr;
        size_t cap = 0;
        hipEvent_t done = nullptr;
        bool used = false;
    };
    static constexpr uint32_t STAGES = 4;
    Stage stage[STAGES];
    uint32_t stage_next = 0;
    // Pinned host memory the readbacks land in (grow-only): the device-to-host copies are enqueued
    // behind the render on its stream and the caller's buffers filled from here by memcpy, as
    // GpuRenderer::render copies its mapped readback buffer (gpu.rs:853-886).
    void *readback = nullptr;
    size_t readback_cap = 0;
    uint64_t ex_records = 0;  // records the last specrel_extrude_tubes left on the device
    void *scratch = nullptr;  // this device's scratch area (specrel_render_noalloc)
    bool last_noalloc = false;  // the last render wrote the caller's device buffers
    // specrel_extrude_tubes: inputs, per-block counts and offsets, outputs.
    DevBuf ex_tubes, ex_tris, ex_samples, ex_decided, ex_counts, ex_offsets, ex_first, ex_geoms, ex_attrs, ex_boxes;
    DevBuf adj_target, adj_mask_in, adj_mask_out, adj_obj, adj_pix, adj_sig, adj_objid, adj_inside, adj_partials,
        adj_counts, adj_result;
    std::vector<specrel_frame> last_frames;
    uint32_t last_scene = 0;
    uint64_t generation = 0;  // bumped whenever a buffer a captured graph points to moves
    int cus = 0;
    hipEvent_t ev[11] = {};  // 0-5 stages on the main stream, 7-8 adjoint, 9-10 star pass on the side stream
    // Second stream (default priority) for the star pass, which the primary rays do not need;
    // fork / join order it against the main stream (also inside a graph capture).
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    // The next frame's tile order runs on the side stream beside the shading and bloom passes.
    hipEvent_t order_fork = nullptr, order_join = nullptr;
    bool timed_stages = true, stars_overlapped = false, shade_fused = false;
    int stage_level = 2;  // SPECREL_OPT_STAGE_EVENTS of the last directly launched render
    specrel_timings timings{};
    specrel_stats stats_host{};
    int opt_wave_clock = 0;
    int opt_stats = 0, opt_cull = 1, opt_stack = STACK_DEFAULT, opt_persistent = 0, opt_shadow_queue = 0;
    int opt_graph = 1;
    int opt_merged = 0;
    int opt_shadow_trace = 1;
    int opt_trace_waves = 1;
    int opt_node_width = 2;  // SPECREL_OPT_NODE_WIDTH
    int opt_lds_top = 0;     // SPECREL_OPT_LDS_TOP: top levels copied into LDS (0: none)
    int opt_fuse_shade = 1;  // SPECREL_OPT_FUSED_SHADE
    int opt_stage_events = 2;  // SPECREL_OPT_STAGE_EVENTS
    int opt_wave_shape = 0;
    int opt_tile_order = 1;
    int opt_postpone = 1;
    int opt_star_overlap = 0;
    int opt_adj_strict = 0;
    bool have_hits = false;
    // The captured launch sequence of the last batch shape (SPECREL_OPT_GRAPH).
    hipGraphExec_t graph = nullptr;
    std::vector<uint64_t> graph_key;
};

namespace {

std::string g_create_err;

void check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw Fail{SPECREL_EDEVICE, std::string(what) + ": " + hipGetErrorString(e)};
}

// The scratch area's address on `device`, which must be current; looked up on the first call for
// that device (specrel_create).
void *scratch_address(int device) {
    if (device < 0 || device >= SCRATCH_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_lock);
    if (!g_scratch[device]) {
        void *ptr = nullptr;
        check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch[device] = ptr;
    }
    return g_scratch[device];
}

void dfree(DevBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

void dalloc(DevBuf &b, size_t bytes) {
    if (b.bytes >= bytes && b.p) return;
    dfree(b);
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        (void)hipGetLastError();
        throw Fail{SPECREL_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed"};
    }
    b.bytes = bytes;
}

// dalloc for buffers that captured graphs reference: a move invalidates the graph.
void galloc(specrel_ctx *ctx, DevBuf &b, size_t bytes) {
    void *before = b.p;
    dalloc(b, bytes);
    if (b.p != before) ctx->generation++;
}

void upload(DevBuf &b, const void *src, size_t bytes, hipStream_t s, size_t cap = 0) {
    dalloc(b, std::max(bytes, cap));
    if (bytes) check(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
}

template <class F>
int guarded(specrel_ctx *ctx, F &&f) {
    try {
        return f();
    } catch (const Fail &e) {
        if (ctx) ctx->err = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        if (ctx) ctx->err = e.what();
        return SPECREL_EINVAL;
    }
}

// A tree over nothing (BvhNode4D::NOTHING, bvh4d.rs:31: no primitives, an empty box, as the root
// of an empty arena): no ray enters it, so it is packed as a leaf without primitives -- entry
// LEAF, large leaf 0, which every scene's large-leaf table holds as (0, 0). (The reference's slab
// test admits the empty box, and its traversal then pushes node 0 as its own children forever.)
bool is_nothing(const specrel_node &n) { return n.count == 0 && !(n.min[0] <= n.max[0]); }

// Large leaf 0: no primitives, for NOTHING nodes and the root of a scene without nodes.
std::vector<uint2> empty_big_table() { return {make_uint2(0, 0)}; }

// The device entry of node j of `a`: a leaf (count and first slot, or an index into the large
// leaves, appended to `big`) or, for an interior node, j itself.
uint32_t encode_entry(const specrel_scene_arrays *a, uint32_t j, std::vector<uint2> &big) {
    const specrel_node &n = a->nodes[j];
    if (is_nothing(n)) return LEAF;
    if (n.count > 0) {
        if (n.count <= COUNT_MAX && n.right_or_first <= FIRST_MASK)
            return LEAF | (n.count << COUNT_SHIFT) | n.right_or_first;
        if (big.size() > FIRST_MASK) throw Fail{SPECREL_EINVAL, "too many large leaves"};
        big.push_back(make_uint2(n.right_or_first, n.count));
        return LEAF | (uint32_t)(big.size() - 1);
    }
    return j;
}

// Walks the tree of `a` from the root: every node reached once, children in range, leaf ranges
// within the primitives. parent[j] receives each reached node's parent slot (parent << 1 | right
// side); the root and the nodes the tree does not reach get ~0u. `interior` (may be NULL) receives
// the reached interior nodes. Malformed input throws, so no traversal can loop.
void walk_tree(const specrel_scene_arrays *a, std::vector<uint32_t> &parent, std::vector<uint32_t> *interior) {
    const uint32_t nn = a->num_nodes, np = a->num_prims;
    parent.assign(nn, ~0u);
    if (interior) interior->clear();
    if (!nn) return;
    std::vector<uint8_t> seen(nn, 0);
    std::vector<uint32_t> todo{0};
    seen[0] = 1;
    while (!todo.empty()) {
        const uint32_t i = todo.back();
        todo.pop_back();
        const specrel_node &n = a->nodes[i];
        if (is_nothing(n)) continue;
        if (n.count > 0) {
            if ((uint64_t)n.right_or_first + n.count > np) throw Fail{SPECREL_EINVAL, "leaf range out of bounds"};
            continue;
        }
        const uint32_t l = n.left, r = n.right_or_first;
        if (l >= nn || r >= nn) throw Fail{SPECREL_EINVAL, "child index out of range"};
        if (seen[l] || seen[r] || l == r) throw Fail{SPECREL_EINVAL, "node graph is not a tree"};
        seen[l] = seen[r] = 1;
        parent[l] = i << 1;
        parent[r] = (i << 1) | 1u;
        if (interior) interior->push_back(i);
        todo.push_back(r);
        todo.push_back(l);
    }
}

// The materials every primitive references exist (Scene4D::materials index, renderer.rs:422) and
// the stars' segments lie within their words (pack_stars layout, gpu.rs:102-119).
void check_prims_and_stars(const specrel_scene_arrays *a, uint32_t first, uint32_t count) {
    for (uint32_t i = first; i < first + count; ++i) {
        uint32_t m;
        std::memcpy(&m, &a->attrs[i].misc[2], 4);
        if (m >= a->num_materials) throw Fail{SPECREL_EINVAL, "primitive " + std::to_string(i) + " has material " +
                                                                  std::to_string(m) + " out of range"};
        if (a->order[i] >= a->num_prims) throw Fail{SPECREL_EINVAL, "order entry out of range"};
    }
}
void check_stars(const specrel_scene_arrays *a) {
    if (!a->num_stars) return;
    if (!a->star_words || a->num_star_words < a->num_stars) throw Fail{SPECREL_EINVAL, "star words too short"};
    uint64_t nseg = (a->num_star_words - a->num_stars) / 3;
    for (uint32_t i = 0; i < a->num_stars; ++i) {
        uint32_t first, count;
        std::memcpy(&first, &a->star_words[4 * i + 2], 4);
        std::memcpy(&count, &a->star_words[4 * i + 3], 4);
        if ((uint64_t)first + count > nseg) throw Fail{SPECREL_EINVAL, "star segments out of range"};
    }
}

// The interior nodes of the top TOP_LEVELS_MAX levels of the uploaded tree, breadth first, from
// the entries as uploaded (ninfo_host, the root entry): their node indices, and their child
// entries with those naming another top node rewritten as TOP_BIT | its position.
void build_top(SceneSlot &slot) {
    std::vector<uint32_t> idx;
    std::vector<uint2> info;
    const uint32_t root = slot.S.root_entry;
    int l = 0;
    if (!slot.S.empty && !(root & LEAF)) {
        std::vector<uint32_t> level{root}, next;
        for (; l < TOP_LEVELS_MAX && !level.empty(); ++l) {
            slot.top_count[l] = (uint32_t)idx.size();
            next.clear();
            for (uint32_t j : level) {
                idx.push_back(j);
                const uint2 in = slot.ninfo_host[j];
                if (!(in.x & LEAF)) next.push_back(in.x);
                if (!(in.y & LEAF)) next.push_back(in.y);
            }
            level.swap(next);
        }
        std::unordered_map<uint32_t, uint32_t> pos;
        for (uint32_t k = 0; k < idx.size(); ++k) pos[idx[k]] = k;
        auto enc = [&](uint32_t e) {
            if (e & LEAF) return e;
            auto it = pos.find(e);
            return it == pos.end() ? e : (TOP_BIT | it->second);
        };
        for (uint32_t j : idx) info.push_back(make_uint2(enc(slot.ninfo_host[j].x), enc(slot.ninfo_host[j].y)));
    }
    for (; l <= TOP_LEVELS_MAX; ++l) slot.top_count[l] = (uint32_t)idx.size();  // levels past the tree's depth
    dalloc(slot.top_idx, std::max<size_t>(1, idx.size()) * 4);
    dalloc(slot.top_info, std::max<size_t>(1, info.size()) * 8);
    if (!idx.empty()) {
        check(hipMemcpy(slot.top_idx.p, idx.data(), idx.size() * 4, hipMemcpyHostToDevice), "top levels");
        check(hipMemcpy(slot.top_info.p, info.data(), info.size() * 8, hipMemcpyHostToDevice), "top levels");
    }
    slot.S.top_idx = (const uint32_t *)slot.top_idx.p;
    slot.S.top_info = (const uint2 *)slot.top_info.p;
    slot.S.top_n = 0;
    slot.S.top_root = TOP_BIT;
}

// Node j's two child boxes (left min, left max, right min, right max) and entries.
void node_row(const specrel_scene_arrays *a, uint32_t j, float4 *row, uint2 &info, std::vector<uint2> &big) {
    const specrel_node &n = a->nodes[j];
    const specrel_node &L = a->nodes[n.left], &R = a->nodes[n.right_or_first];
    row[0] = make_float4(L.min[0], L.min[1], L.min[2], L.min[3]);
    row[1] = make_float4(L.max[0], L.max[1], L.max[2], L.max[3]);
    row[2] = make_float4(R.min[0], R.min[1], R.min[2], R.min[3]);
    row[3] = make_float4(R.max[0], R.max[1], R.max[2], R.max[3]);
    info = make_uint2(encode_entry(a, n.left, big), encode_entry(a, n.right_or_first, big));
}

// Room the device arrays of a scene edited in place get at upload: half as much again, as the
// reference's GPU buffers of such a scene (gpu.rs:639-656).
size_t with_room(size_t bytes, bool room) { return room ? bytes + bytes / 2 : bytes; }

// Re-pack the reference arrays into the device layout, validating the tree on the way. With
// `room`, the node, primitive and star arrays are allocated half again as large, so that later
// edits can grow in place.
void build_device_scene(specrel_ctx *ctx, SceneSlot &slot, const specrel_scene_arrays *a, bool room = false) {
    if (!a) throw Fail{SPECREL_EINVAL, "scene is NULL"};
    const uint32_t nn = a->num_nodes, np = a->num_prims;
    if (nn && !a->nodes) throw Fail{SPECREL_EINVAL, "nodes is NULL"};
    if (np && (!a->geoms || !a->attrs || !a->order)) throw Fail{SPECREL_EINVAL, "prim arrays are NULL"};
    if (a->num_materials == 0 && np) throw Fail{SPECREL_EINVAL, "no materials"};
    if (np >= 0x7FFFFFFFu) throw Fail{SPECREL_EINVAL, "too many primitives"};
    if (nn > 0x7FFFFFFFu / 2) throw Fail{SPECREL_EINVAL, "too many nodes"};
    check_prims_and_stars(a, 0, np);
    check_stars(a);

    std::vector<float4> nbox;
    std::vector<uint2> ninfo;
    std::vector<uint2> big = empty_big_table();
    std::vector<uint32_t> interior;
    walk_tree(a, slot.parent, &interior);
    Scene &S = slot.S;
    S.empty = nn == 0 || is_nothing(a->nodes[0]);
    S.root_entry = LEAF;
    S.root_min = make_float4(INFINITY, INFINITY, INFINITY, INFINITY);
    S.root_max = make_float4(-INFINITY, -INFINITY, -INFINITY, -INFINITY);
    if (nn) {
        nbox.assign((size_t)nn * 4, make_float4(0, 0, 0, 0));
        ninfo.assign(nn, make_uint2(0, 0));
        for (uint32_t i : interior) node_row(a, i, &nbox[4 * (size_t)i], ninfo[i], big);
        const specrel_node &root = a->nodes[0];
        S.root_entry = encode_entry(a, 0, big);
        S.root_min = make_float4(root.min[0], root.min[1], root.min[2], root.min[3]);
        S.root_max = make_float4(root.max[0], root.max[1], root.max[2], root.max[3]);
    }
    hipStream_t s = ctx->stream;
    const size_t sb = sizeof(float4), gb = sizeof(specrel_geom), ab = sizeof(specrel_attr);
    upload(slot.nbox, nbox.data(), nbox.size() * sb, s, with_room(nbox.size() * sb, room));
    upload(slot.ninfo, ninfo.data(), ninfo.size() * sizeof(uint2), s, with_room(ninfo.size() * sizeof(uint2), room));
    upload(slot.big, big.data(), big.size() * sizeof(uint2), s, with_room(big.size() * sizeof(uint2) + 64, room));
    upload(slot.geoms, a->geoms, (size_t)np * gb, s, with_room((size_t)np * gb, room));
    upload(slot.attrs, a->attrs, (size_t)np * ab, s, with_room((size_t)np * ab, room));
    upload(slot.order, a->order, (size_t)np * 4, s, with_room((size_t)np * 4, room));
    upload(slot.mats, a->materials, (size_t)a->num_materials * sizeof(specrel_material), s);
    upload(slot.stars, a->star_words, (size_t)a->num_star_words * 16, s, with_room((size_t)a->num_star_words * 16, room));
    check(hipStreamSynchronize(s), "upload");
    S.nbox = (const float4 *)slot.nbox.p;
    S.ninfo = (const uint2 *)slot.ninfo.p;
    S.big_leaves = (const uint2 *)slot.big.p;
    S.geoms = (const float4 *)slot.geoms.p;
    S.attrs = (const float4 *)slot.attrs.p;
    S.order = (const uint32_t *)slot.order.p;
    S.materials = (const float4 *)slot.mats.p;
    S.num_prims = np;
    S.light = make_float4(a->light_pos[0], a->light_pos[1], a->light_pos[2], a->light_pos[3]);
    slot.num_stars = a->num_stars;
    slot.num_star_words = a->num_star_words;
    slot.num_materials = a->num_materials;
    slot.num_nodes = nn;
    slot.num_prims = np;
    slot.ninfo_host.swap(ninfo);
    slot.big_host.swap(big);
    slot.used = true;
    build_top(slot);
}

// The encoded child entry of node j of `a` as it stands, compared with what was uploaded
// (`have`): equal when j is the same kind of node with the same children or primitive range.
bool same_entry(const SceneSlot &slot, const specrel_scene_arrays *a, uint32_t j, uint32_t have) {
    const specrel_node &n = a->nodes[j];
    if (n.count == 0) return have == j;
    if ((have & LEAF) == 0) return false;
    if (n.count <= COUNT_MAX && n.right_or_first <= FIRST_MASK)
        return have == (LEAF | (n.count << COUNT_SHIFT) | n.right_or_first);
    uint32_t c = (have >> COUNT_SHIFT) & COUNT_MAX, idx = have & FIRST_MASK;
    return c == 0 && idx < slot.big_host.size() && slot.big_host[idx].x == n.right_or_first &&
           slot.big_host[idx].y == n.count;
}

// GpuRenderer::update_scene with dirty ranges (gpu.rs:704-731, PackedSceneGpuData::update
// gpu.rs:297-344): node bounds and primitive records of the given ranges are copied into the
// uploaded layout. Returns false when the arrays no longer match the uploaded tree (counts,
// children or leaf ranges), in which case the caller uploads everything again.
bool patch_device_scene(specrel_ctx *ctx, SceneSlot &slot, const specrel_scene_arrays *a, const specrel_range *nr,
                        uint32_t n_nr, const specrel_range *pr, uint32_t n_pr) {
    if (a->num_nodes != slot.num_nodes || a->num_prims != slot.num_prims || a->num_materials != slot.num_materials ||
        a->num_stars != slot.num_stars || a->num_star_words != slot.num_star_words || slot.num_nodes == 0)
        return false;
    if ((n_nr && !nr) || (n_pr && !pr)) throw Fail{SPECREL_EINVAL, "range list is NULL"};
    const uint32_t nn = slot.num_nodes, np = slot.num_prims;
    std::vector<uint32_t> pslot;
    std::vector<float4> bounds;
    for (uint32_t r = 0; r < n_nr; ++r) {
        if ((uint64_t)nr[r].first + nr[r].count > nn) throw Fail{SPECREL_EINVAL, "node range out of bounds"};
        for (uint32_t j = nr[r].first; j < nr[r].first + nr[r].count; ++j) {
            const specrel_node &n = a->nodes[j];
            if (n.count == 0) {
                if (n.left >= nn || n.right_or_first >= nn) return false;
                if (!same_entry(slot, a, n.left, slot.ninfo_host[j].x) ||
                    !same_entry(slot, a, n.right_or_first, slot.ninfo_host[j].y))
                    return false;
            }
            if (j == 0) {
                slot.S.root_min = make_float4(n.min[0], n.min[1], n.min[2], n.min[3]);
                slot.S.root_max = make_float4(n.max[0], n.max[1], n.max[2], n.max[3]);
                continue;
            }
            if (slot.parent[j] == ~0u) return false;
            const uint32_t p = slot.parent[j] >> 1;
            const uint32_t have = (slot.parent[j] & 1u) ? slot.ninfo_host[p].y : slot.ninfo_host[p].x;
            if (!same_entry(slot, a, j, have)) return false;
            pslot.push_back(slot.parent[j]);
            bounds.push_back(make_float4(n.min[0], n.min[1], n.min[2], n.min[3]));
            bounds.push_back(make_float4(n.max[0], n.max[1], n.max[2], n.max[3]));
        }
    }
    for (uint32_t r = 0; r < n_pr; ++r) {
        if ((uint64_t)pr[r].first + pr[r].count > np) throw Fail{SPECREL_EINVAL, "primitive range out of bounds"};
        for (uint32_t i = pr[r].first; i < pr[r].first + pr[r].count; ++i) {
            uint32_t m;
            std::memcpy(&m, &a->attrs[i].misc[2], 4);
            if (m >= slot.num_materials) throw Fail{SPECREL_EINVAL, "patched primitive has material out of range"};
            if (a->order[i] >= np) throw Fail{SPECREL_EINVAL, "patched order entry out of range"};
        }
    }
    hipStream_t s = ctx->stream;
    for (uint32_t r = 0; r < n_pr; ++r) {
        const size_t f = pr[r].first, c = pr[r].count;
        if (!c) continue;
        check(hipMemcpyAsync((specrel_geom *)slot.geoms.p + f, a->geoms + f, c * sizeof(specrel_geom),
                             hipMemcpyHostToDevice, s), "patch geoms");
        check(hipMemcpyAsync((specrel_attr *)slot.attrs.p + f, a->attrs + f, c * sizeof(specrel_attr),
                             hipMemcpyHostToDevice, s), "patch attrs");
        check(hipMemcpyAsync((uint32_t *)slot.order.p + f, a->order + f, c * 4, hipMemcpyHostToDevice, s),
              "patch order");
    }
    if (!pslot.empty()) {
        DevBuf tmp;
        const size_t n = pslot.size();
        try {
            dalloc(tmp, n * (32 + 4));
            float4 *d_b = (float4 *)tmp.p;
            uint32_t *d_ps = (uint32_t *)(d_b + 2 * n);
            check(hipMemcpyAsync(d_ps, pslot.data(), n * 4, hipMemcpyHostToDevice, s), "patch parents");
            check(hipMemcpyAsync(d_b, bounds.data(), n * 32, hipMemcpyHostToDevice, s), "patch bounds");
            k_patch_nodes<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(d_ps, d_b, (uint32_t)n, (float4 *)slot.nbox.p);
            check(hipGetLastError(), "k_patch_nodes launch");
            check(hipStreamSynchronize(s), "patch");
        } catch (...) {
            dfree(tmp);
            throw;
        }
        dfree(tmp);
    }
    if (a->num_materials)
        check(hipMemcpyAsync(slot.mats.p, a->materials, (size_t)a->num_materials * sizeof(specrel_material),
                             hipMemcpyHostToDevice, s), "patch materials");
    if (a->num_star_words)
        check(hipMemcpyAsync(slot.stars.p, a->star_words, (size_t)a->num_star_words * 16, hipMemcpyHostToDevice, s),
              "patch stars");
    slot.S.light = make_float4(a->light_pos[0], a->light_pos[1], a->light_pos[2], a->light_pos[3]);
    check(hipStreamSynchronize(s), "patch");
    return true;
}

// The device entry `have` (as uploaded) still describes node j of `a`: the same kind of node,
// with the same children or primitive range.
bool entry_matches(const std::vector<uint2> &big, const specrel_scene_arrays *a, uint32_t j, uint32_t have) {
    const specrel_node &n = a->nodes[j];
    if (is_nothing(n)) return have == LEAF;
    if (n.count == 0) return have == j;
    if ((have & LEAF) == 0) return false;
    if (n.count <= COUNT_MAX && n.right_or_first <= FIRST_MASK)
        return have == (LEAF | (n.count << COUNT_SHIFT) | n.right_or_first);
    const uint32_t c = (have >> COUNT_SHIFT) & COUNT_MAX, idx = have & FIRST_MASK;
    return c == 0 && idx < big.size() && big[idx].x == n.right_or_first && big[idx].y == n.count;
}

// GpuRenderer::update_scene's patch (gpu.rs:704-731) for the edit that takes the slot's arrays
// to `a`: the nodes and primitives of the ranges that edit wrote (bvh4d.rs:158-215), re-encoded
// into the device layout, plus the stars, materials and light. The arrays may have grown. The
// device holds, per interior node, the boxes and entries of its two children, and only for the
// nodes the uploaded tree reached, so the rows written are those of the reached interior nodes
// that are listed, that hold a listed node's box (its parent), that the previous tree did not
// reach (a subtree linked in again: its rows were never made or are stale), or whose links
// differ from the device's (a node rewritten but left out of the ranges). Before any device
// write the new tree is walked: every node reached once, links and leaf ranges in bounds.
// Returns false, having written nothing, when the grown arrays do not fit the slot's buffers:
// the caller then uploads everything.
bool update_device_scene(specrel_ctx *ctx, SceneSlot &slot, const specrel_scene_arrays *a, const specrel_range *nr,
                         uint32_t n_nr, const specrel_range *pr, uint32_t n_pr) {
    const uint32_t nn = a->num_nodes, np = a->num_prims;
    if (nn && !a->nodes) throw Fail{SPECREL_EINVAL, "nodes is NULL"};
    if (np && (!a->geoms || !a->attrs || !a->order)) throw Fail{SPECREL_EINVAL, "prim arrays are NULL"};
    if ((n_nr && !nr) || (n_pr && !pr)) throw Fail{SPECREL_EINVAL, "range list is NULL"};
    if (a->num_materials == 0 && np) throw Fail{SPECREL_EINVAL, "no materials"};
    if (np >= 0x7FFFFFFFu || nn > 0x7FFFFFFFu / 2) throw Fail{SPECREL_EINVAL, "too many nodes or primitives"};
    for (uint32_t r = 0; r < n_nr; ++r)
        if ((uint64_t)nr[r].first + nr[r].count > nn) throw Fail{SPECREL_EINVAL, "node range out of bounds"};
    for (uint32_t r = 0; r < n_pr; ++r) {
        if ((uint64_t)pr[r].first + pr[r].count > np) throw Fail{SPECREL_EINVAL, "primitive range out of bounds"};
        check_prims_and_stars(a, pr[r].first, pr[r].count);
    }
    check_stars(a);
    if (nn == 0 || slot.num_nodes == 0) return false;
    const bool fits = (size_t)nn * 64 <= slot.nbox.bytes && (size_t)nn * 8 <= slot.ninfo.bytes &&
                      (size_t)np * sizeof(specrel_geom) <= slot.geoms.bytes &&
                      (size_t)np * sizeof(specrel_attr) <= slot.attrs.bytes && (size_t)np * 4 <= slot.order.bytes &&
                      (size_t)a->num_star_words * 16 <= slot.stars.bytes &&
                      (size_t)a->num_materials * sizeof(specrel_material) <= slot.mats.bytes;
    if (!fits) return false;
    std::vector<uint32_t> parent, interior;
    walk_tree(a, parent, &interior);
    std::vector<uint8_t> write(nn, 0);
    for (uint32_t r = 0; r < n_nr; ++r)
        for (uint32_t j = nr[r].first; j < nr[r].first + nr[r].count; ++j) {
            if (j != 0 && parent[j] == ~0u) continue;  // listed but not in the tree: free room
            const specrel_node &n = a->nodes[j];
            if (n.count == 0 && !is_nothing(n)) write[j] = 1;
            if (parent[j] != ~0u) write[parent[j] >> 1] = 1;
        }
    const size_t old_nn = slot.ninfo_host.size();
    for (uint32_t i : interior) {
        if (write[i]) continue;
        const bool was_reached = i < old_nn && i < slot.parent.size() && (i == 0 ? !slot.S.empty : slot.parent[i] != ~0u);
        const specrel_node &n = a->nodes[i];
        if (!was_reached || !entry_matches(slot.big_host, a, n.left, slot.ninfo_host[i].x) ||
            !entry_matches(slot.big_host, a, n.right_or_first, slot.ninfo_host[i].y))
            write[i] = 1;
    }
    std::vector<uint2> big = slot.big_host;
    const size_t big0 = big.size();
    std::vector<uint32_t> idx;
    std::vector<float4> rows;
    std::vector<uint2> info;
    for (uint32_t i : interior) {
        if (!write[i]) continue;
        idx.push_back(i);
        rows.resize(rows.size() + 4);
        info.push_back(make_uint2(0, 0));
        node_row(a, i, &rows[rows.size() - 4], info.back(), big);
    }
    Scene S = slot.S;
    {  // the root's own box and entry live in the scene record: always brought up to date
        const specrel_node &root = a->nodes[0];
        if (!entry_matches(big, a, 0, S.root_entry)) S.root_entry = encode_entry(a, 0, big);
        S.root_min = make_float4(root.min[0], root.min[1], root.min[2], root.min[3]);
        S.root_max = make_float4(root.max[0], root.max[1], root.max[2], root.max[3]);
        S.empty = is_nothing(root);
    }
    if (big.size() * sizeof(uint2) > slot.big.bytes) return false;
    // Everything checked: write.
    hipStream_t s = ctx->stream;
    DevBuf tmp;
    if (!idx.empty()) {
        const size_t n = idx.size();
        dalloc(tmp, n * (64 + 8 + 4));
        float4 *d_rows = (float4 *)tmp.p;
        uint2 *d_info = (uint2 *)(d_rows + 4 * n);
        uint32_t *d_idx = (uint32_t *)(d_info + n);
        check(hipMemcpyAsync(d_rows, rows.data(), n * 64, hipMemcpyHostToDevice, s), "update rows");
        check(hipMemcpyAsync(d_info, info.data(), n * 8, hipMemcpyHostToDevice, s), "update rows");
        check(hipMemcpyAsync(d_idx, idx.data(), n * 4, hipMemcpyHostToDevice, s), "update rows");
        k_patch_rows<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(d_idx, d_rows, d_info, (uint32_t)n, (float4 *)slot.nbox.p,
                                                                  (uint2 *)slot.ninfo.p);
        check(hipGetLastError(), "k_patch_rows launch");
    }
    if (big.size() > big0)
        check(hipMemcpyAsync((uint2 *)slot.big.p + big0, big.data() + big0, (big.size() - big0) * sizeof(uint2),
                             hipMemcpyHostToDevice, s),
              "update large leaves");
    for (uint32_t r = 0; r < n_pr; ++r) {
        const size_t f = pr[r].first, c = pr[r].count;
        if (!c) continue;
        check(hipMemcpyAsync((specrel_geom *)slot.geoms.p + f, a->geoms + f, c * sizeof(specrel_geom),
                             hipMemcpyHostToDevice, s), "update geoms");
        check(hipMemcpyAsync((specrel_attr *)slot.attrs.p + f, a->attrs + f, c * sizeof(specrel_attr),
                             hipMemcpyHostToDevice, s), "update attrs");
        check(hipMemcpyAsync((uint32_t *)slot.order.p + f, a->order + f, c * 4, hipMemcpyHostToDevice, s),
              "update order");
    }
    if (a->num_materials)
        check(hipMemcpyAsync(slot.mats.p, a->materials, (size_t)a->num_materials * sizeof(specrel_material),
                             hipMemcpyHostToDevice, s), "update materials");
    if (a->num_star_words)
        check(hipMemcpyAsync(slot.stars.p, a->star_words, (size_t)a->num_star_words * 16, hipMemcpyHostToDevice, s),
              "update stars");
    check(hipStreamSynchronize(s), "update");
    S.num_prims = np;
    S.light = make_float4(a->light_pos[0], a->light_pos[1], a->light_pos[2], a->light_pos[3]);
    slot.S = S;
    slot.ninfo_host.resize(nn, make_uint2(0, 0));
    for (size_t k = 0; k < idx.size(); ++k) slot.ninfo_host[idx[k]] = info[k];
    slot.parent.swap(parent);
    slot.big_host.swap(big);
    slot.num_nodes = nn;
    slot.num_prims = np;
    slot.num_materials = a->num_materials;
    slot.num_stars = a->num_stars;
    slot.num_star_words = a->num_star_words;
    build_top(slot);
    return true;
}

uint32_t num_super(uint32_t w, uint32_t h) { return ((w + SUPER - 1) / SUPER) * ((h + SUPER - 1) / SUPER); }

constexpr const char *NOALLOC_LAST =
    "the last render was specrel_render_noalloc: its outputs are in the caller's device buffers";
constexpr int NLEV = SPECREL_BLOOM_LEVELS + 1;
// Rows of the view table: NLEV tile levels, then super tiles, pixels, pyramid float4s.
constexpr int ROW_SUPER = NLEV, ROW_PIX = NLEV + 1, ROW_PYR = NLEV + 2, ROWS = NLEV + 3;
static_assert(ROWS == TABLE_ROWS, "the view table of k_reset_view has the rows of the context's table");
constexpr uint32_t MAX_VIEWS = 4096;

void check_len(uint64_t len, uint64_t want, const char *what, uint32_t view) {
    if (len != want)
        throw Fail{SPECREL_EINVAL, std::string(what) + " size mismatch (view " + std::to_string(view) + "): " +
                                       std::to_string(len) + " elements given, " + std::to_string(want) +
                                       " needed"};
}

void check_frame(const specrel_frame *f, uint32_t i) {
    const uint32_t w = f->size[0], h = f->size[1];
    std::string at = " (view " + std::to_string(i) + ")";
    if (w == 0 || h == 0 || w > 32768 || h > 32768) throw Fail{SPECREL_EINVAL, "bad frame size" + at};
    if (f->mode > 3 || f->shadow_mode > 2 || f->slice_mode > 3) throw Fail{SPECREL_EINVAL, "bad render mode" + at};
    if (f->num_flashes > SPECREL_MAX_FLASHES || f->num_haze > SPECREL_MAX_HAZE)
        throw Fail{SPECREL_EINVAL, "too many flashes or haze origins" + at};
}

// Sizes every per-launch buffer for these views (growing only) and uploads the view table when
// the shape changed.
void ensure_views(specrel_ctx *ctx, const specrel_frame *frames, uint32_t n, uint32_t num_stars) {
    std::vector<uint32_t> sizes(2 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
        sizes[2 * i] = frames[i].size[0];
        sizes[2 * i + 1] = frames[i].size[1];
    }
    if (sizes != ctx->sizes) {
        const size_t stride = (size_t)n + 1;
        std::vector<uint32_t> tab((size_t)ROWS * stride, 0u);
        uint64_t acc[ROWS] = {};
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t w = sizes[2 * i], h = sizes[2 * i + 1];
            for (int k = 0; k < NLEV; ++k) {
                uint32_t lw, lh;
                level_size(w, h, k, lw, lh);
                acc[k] += (uint64_t)((lw + 15) / 16) * ((lh + 15) / 16);
            }
            acc[ROW_SUPER] += num_super(w, h);
            acc[ROW_PIX] += (uint64_t)w * h;
            acc[ROW_PYR] += pyramid_len(w, h);
            for (int r = 0; r < ROWS; ++r) {
                if (acc[r] >= 0xFFFFFFFFull || acc[0] >= (1ull << 24))
                    throw Fail{SPECREL_EINVAL, "views too large for one launch"};
                tab[(size_t)r * stride + i + 1] = (uint32_t)acc[r];
            }
        }
        ctx->total_tiles = (uint32_t)acc[0];
        ctx->total_supers = (uint32_t)acc[ROW_SUPER];
        ctx->total_pix = (uint32_t)acc[ROW_PIX];
        ctx->total_pyr = (uint32_t)acc[ROW_PYR];
        const size_t np = ctx->total_pix;
        galloc(ctx, ctx->hdr, (size_t)ctx->total_pyr * sizeof(float4));
        galloc(ctx, ctx->rgb, np * 4);
        galloc(ctx, ctx->hit, np * 4);
        galloc(ctx, ctx->lam, np * 4);
        galloc(ctx, ctx->tau, np * 4);
        galloc(ctx, ctx->slot, np * 4);
        galloc(ctx, ctx->mask, np * 4);
        galloc(ctx, ctx->super_counts, (size_t)ctx->total_supers * 4);
        galloc(ctx, ctx->super_lists, (size_t)ctx->total_supers * SUPER_CAP * 4);
        galloc(ctx, ctx->frames, (size_t)n * sizeof(specrel_frame));
        galloc(ctx, ctx->tab, tab.size() * 4);
        // Tile order of the primary rays: row order until a frame has measured the tiles' costs.
        galloc(ctx, ctx->tile_cost, (size_t)ctx->total_tiles * 16);
        galloc(ctx, ctx->tile_cones, (size_t)ctx->total_tiles * 16);
        galloc(ctx, ctx->tile_order, (size_t)ctx->total_tiles * 4);
        k_tile_init<<<(ctx->total_tiles + 255) / 256, 256, 0, ctx->stream>>>((uint32_t *)ctx->tile_order.p,
                                                                            (uint32_t *)ctx->tile_cost.p,
                                                                            ctx->total_tiles);
        check(hipGetLastError(), "k_tile_init launch");
        ctx->tab_host.swap(tab);
        check(hipMemcpyAsync(ctx->tab.p, ctx->tab_host.data(), ctx->tab_host.size() * 4, hipMemcpyHostToDevice,
                             ctx->stream),
              "view table upload");
        ctx->sizes.swap(sizes);
        ctx->n_views = n;
    }
    const size_t ns = (size_t)n * (num_stars ? num_stars : 1);
    galloc(ctx, ctx->prep, ns * sizeof(StarPrep));
    galloc(ctx, ctx->dirs, ns * sizeof(StarDir));
}

Views device_views(specrel_ctx *ctx) {
    Views V;
    const uint32_t *t = (const uint32_t *)ctx->tab.p;
    const size_t stride = (size_t)ctx->n_views + 1;
    V.frames = (const specrel_frame *)ctx->frames.p;
    for (int k = 0; k < NLEV; ++k) V.level_start[k] = t + k * stride;
    V.super_start = t + ROW_SUPER * stride;
    V.pix_start = t + ROW_PIX * stride;
    V.pyr_start = t + ROW_PYR * stride;
    V.n = ctx->n_views;
    return V;
}

uint32_t host_start(const specrel_ctx *ctx, int row, uint32_t view) {
    return ctx->tab_host[(size_t)row * (ctx->n_views + 1) + view];
}

template <class Kern>
dim3 resident_grid(specrel_ctx *ctx, Kern kern) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    return dim3((unsigned)(per_cu * ctx->cus));
}

// The traversal variant the primary-ray kernel runs (specrel_timings.trace_variant): merged load
// rounds when asked for; else, for tiled rays with an LDS stack of 8 or 16 positions, the top
// levels from LDS (SPECREL_OPT_LDS_TOP) or the width-4 fetch of postponed-leaf traversal
// (SPECREL_OPT_NODE_WIDTH); plain rounds otherwise. Instrumented renders run plain or merged
// rounds (the variants visit the same sequence).
int trace_variant(const specrel_ctx *ctx, bool stats) {
    if (ctx->opt_persistent || ctx->opt_stack <= 0) return TV_PLAIN;
    if (ctx->opt_merged) return TV_MERGED;
    if (stats || (ctx->opt_stack != 8 && ctx->opt_stack != 16)) return TV_PLAIN;
    if (ctx->opt_lds_top > 0) return TV_LDSTOP;
    if (ctx->opt_node_width == 4 && ctx->opt_postpone) return TV_WIDE4;
    return TV_PLAIN;
}

// Primary rays with an LDS stack in blocks of 1, 2 or 4 waves (SPECREL_OPT_TRACE_WAVES), each wave
// one quarter of a 16x16 tile; TV: traversal variant; SH: shadow queries traced here.
template <int K, int TV, int SH>
void launch_lds_w(specrel_ctx *ctx, bool stats, dim3 tiles, hipStream_t s, const TraceArgs &ta) {
    const int w = ctx->opt_trace_waves;
    const dim3 g(tiles.x * 4 / w);
    if constexpr ((TV == TV_PLAIN || TV == TV_MERGED) && SH != 2) {
        if (stats) {
            if (w == 1) k_trace<true, K, TV, SH, 1><<<g, 64, 0, s>>>(ta);
            else if (w == 2) k_trace<true, K, TV, SH, 2><<<g, 128, 0, s>>>(ta);
            else k_trace<true, K, TV, SH, 4><<<g, 256, 0, s>>>(ta);
            return;
        }
    }
    if (w == 1) k_trace<false, K, TV, SH, 1><<<g, 64, 0, s>>>(ta);
    else if (w == 2) k_trace<false, K, TV, SH, 2><<<g, 128, 0, s>>>(ta);
    else k_trace<false, K, TV, SH, 4><<<g, 256, 0, s>>>(ta);
}
template <int K, int TV>
void launch_lds(specrel_ctx *ctx, bool stats, dim3 tiles, hipStream_t s, const TraceArgs &ta) {
    if (ta.shade_here) {
        if constexpr ((K == 8 || K == 16) && TV != TV_MERGED) launch_lds_w<K, TV, 2>(ctx, stats, tiles, s, ta);
        else throw Fail{SPECREL_EINVAL, "no fused shading for this traversal"};
    } else if (ta.mask) {
        launch_lds_w<K, TV, 1>(ctx, stats, tiles, s, ta);
    } else {
        launch_lds_w<K, TV, 0>(ctx, stats, tiles, s, ta);
    }
}

template <int K>
void launch_k(specrel_ctx *ctx, bool stats, dim3 tiles, hipStream_t s, const TraceArgs &ta) {
    if (ctx->opt_persistent) {
        if (stats) k_trace_p<true, K><<<resident_grid(ctx, k_trace_p<true, K>), 256, 0, s>>>(ta);
        else k_trace_p<false, K><<<resident_grid(ctx, k_trace_p<false, K>), 256, 0, s>>>(ta);
    } else if constexpr (K > 0) {
        const int tv = trace_variant(ctx, stats);
        if (tv == TV_MERGED) launch_lds<K, TV_MERGED>(ctx, stats, tiles, s, ta);
        else if constexpr (K == 8 || K == 16) {
            if (tv == TV_LDSTOP) launch_lds<K, TV_LDSTOP>(ctx, stats, tiles, s, ta);
            else if (tv == TV_WIDE4) launch_lds<K, TV_WIDE4>(ctx, stats, tiles, s, ta);
            else launch_lds<K, TV_PLAIN>(ctx, stats, tiles, s, ta);
        } else {
            launch_lds<K, TV_PLAIN>(ctx, stats, tiles, s, ta);
        }
    } else {
        if (stats) k_trace<true, K, TV_PLAIN, false, 4><<<tiles, 256, 0, s>>>(ta);
        else k_trace<false, K, TV_PLAIN, false, 4><<<tiles, 256, 0, s>>>(ta);
    }
}
template <int K>
void launch_k(specrel_ctx *ctx, bool stats, dim3, hipStream_t s, const OccludeArgs &oa) {
    if (stats) k_occlude_p<true, K><<<resident_grid(ctx, k_occlude_p<true, K>), 256, 0, s>>>(oa);
    else k_occlude_p<false, K><<<resident_grid(ctx, k_occlude_p<false, K>), 256, 0, s>>>(oa);
}
template <int PASS>
struct ShadePass {
    const ShadeArgs &ra;
};
template <int K, int PASS>
void launch_k(specrel_ctx *, bool stats, dim3 tiles, hipStream_t s, const ShadePass<PASS> &p) {
    if (stats) k_shade<true, K, PASS><<<tiles, 256, 0, s>>>(p.ra);
    else k_shade<false, K, PASS><<<tiles, 256, 0, s>>>(p.ra);
}
template <int K, class A>
void launch_kk(specrel_ctx *ctx, bool stats, dim3 tiles, hipStream_t s, const A &args) {
    launch_k<K>(ctx, stats, tiles, s, args);
}
template <int K, int PASS>
void launch_kk(specrel_ctx *ctx, bool stats, dim3 tiles, hipStream_t s, const ShadePass<PASS> &args) {
    launch_k<K, PASS>(ctx, stats, tiles, s, args);
}
template <class A>
void launch_any(specrel_ctx *ctx, bool stats, dim3 tiles, hipStream_t s, const A &args) {
    switch (ctx->opt_stack) {
        case 8: launch_kk<8>(ctx, stats, tiles, s, args); break;
        case 10: launch_kk<10>(ctx, stats, tiles, s, args); break;
        case 12: launch_kk<12>(ctx, stats, tiles, s, args); break;
        case -4: launch_kk<-4>(ctx, stats, tiles, s, args); break;
        case -8: launch_kk<-8>(ctx, stats, tiles, s, args); break;
        case -12: launch_kk<-12>(ctx, stats, tiles, s, args); break;
        case -16: launch_kk<-16>(ctx, stats, tiles, s, args); break;
        default: launch_kk<16>(ctx, stats, tiles, s, args); break;
    }
}

float ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0f;
    return ms;
}

// The device buffers and view-table facts of one launch sequence: the context's own buffers
// (sized by ensure_views), or for specrel_render_noalloc the fixed scratch area.
struct Bufs {
    Views V;
    uint32_t n_views = 0, total_tiles = 0, total_supers = 0, total_pix = 0;
    const uint32_t *tab_host = nullptr;  // host copy of the view table (ROWS x (n_views + 1))
    const uint32_t *sizes = nullptr;     // w, h per view
    uint32_t *frames = nullptr;          // where k_reset_frame / k_reset_view put the frame
    const TableWords *table = nullptr;   // k_reset_view writes this view table ...
    uint32_t *table_dst = nullptr;       // ... here
    uint32_t *counters = nullptr, *slot = nullptr, *mask = nullptr, *rgb = nullptr, *hit = nullptr;
    float *lam = nullptr, *tau = nullptr;
    float4 *hdr = nullptr, *tile_cones = nullptr;
    uint32_t *super_lists = nullptr, *super_counts = nullptr;
    uint32_t *tile_cost = nullptr, *tile_order = nullptr;  // NULL: tiles in row order
    StarPrep *prep = nullptr;
    StarDir *dirs = nullptr;
    ShadowQuery *squeue = nullptr;
    unsigned long long *stats = nullptr;
    specrel_wave_clock *wave_clock = nullptr;  // SPECREL_OPT_WAVE_CLOCK: one record per wave
    uint32_t start(int row, uint32_t view) const { return tab_host[(size_t)row * (n_views + 1) + view]; }
};

Bufs ctx_bufs(specrel_ctx *ctx) {
    Bufs B;
    B.V = device_views(ctx);
    B.n_views = ctx->n_views;
    B.total_tiles = ctx->total_tiles;
    B.total_supers = ctx->total_supers;
    B.total_pix = ctx->total_pix;
    B.tab_host = ctx->tab_host.data();
    B.sizes = ctx->sizes.data();
    B.frames = (uint32_t *)ctx->frames.p;
    B.counters = (uint32_t *)ctx->counters.p;
    B.slot = (uint32_t *)ctx->slot.p;
    B.mask = (uint32_t *)ctx->mask.p;
    B.rgb = (uint32_t *)ctx->rgb.p;
    B.hit = (uint32_t *)ctx->hit.p;
    B.lam = (float *)ctx->lam.p;
    B.tau = (float *)ctx->tau.p;
    B.hdr = (float4 *)ctx->hdr.p;
    B.tile_cones = (float4 *)ctx->tile_cones.p;
    B.super_lists = (uint32_t *)ctx->super_lists.p;
    B.super_counts = (uint32_t *)ctx->super_counts.p;
    B.tile_cost = (uint32_t *)ctx->tile_cost.p;
    B.tile_order = (uint32_t *)ctx->tile_order.p;
    B.prep = (StarPrep *)ctx->prep.p;
    B.dirs = (StarDir *)ctx->dirs.p;
    B.squeue = (ShadowQuery *)ctx->squeue.p;
    B.stats = (unsigned long long *)ctx->stats.p;
    B.wave_clock = ctx->opt_wave_clock ? (specrel_wave_clock *)ctx->wave_clock.p : nullptr;
    return B;
}

// What a launch over the current views needs, decided on the host from the frames.
struct Plan {
    bool stars = false;    // some view is shaded and the scene has stars (renderer.rs:357)
    bool queries = false;  // some view is shaded with shadows on (renderer.rs:631-641, 680-682)
    uint32_t max_flashes = 0;
    bool want_hits = false;
    bool frame_in_reset = false;  // one view launched directly: its frame goes to k_reset_frame
    FrameWords frame;
};

// Waves per block of the primary-ray kernel as launched: SPECREL_OPT_TRACE_WAVES for tiled rays
// with LDS stacks, else 4 (persistent rays, register stacks).
uint32_t trace_block_waves(const specrel_ctx *ctx) {
    return (ctx->opt_stack > 0 && !ctx->opt_persistent) ? (uint32_t)ctx->opt_trace_waves : 4u;
}

// Where a render's shadow queries are traced (specrel_timings.shadow_path): queued for a dense
// pass, in the primary-ray kernel (LDS stacks, tiled rays), or by the shading pass itself.
uint32_t shadow_path(const specrel_ctx *ctx, const Plan &pl) {
    if (!pl.queries) return SPECREL_SHADOWS_NONE;
    if (ctx->opt_shadow_queue) return SPECREL_SHADOWS_QUEUED;
    if (ctx->opt_shadow_trace && ctx->opt_stack > 0 && !ctx->opt_persistent) return SPECREL_SHADOWS_IN_TRACE;
    return SPECREL_SHADOWS_IN_SHADE;
}

// Validates the frames, sizes the buffers, uploads the frames: everything before the launches.
Plan prepare(specrel_ctx *ctx, uint32_t scene_idx, const specrel_frame *frames, uint32_t n, bool want_hits,
             bool frame_in_reset) {
    if (!frames) throw Fail{SPECREL_EINVAL, "frame is NULL"};
    if (n == 0 || n > MAX_VIEWS) throw Fail{SPECREL_EINVAL, "view count must be 1.." + std::to_string(MAX_VIEWS)};
    if (scene_idx >= ctx->scenes.size() || !ctx->scenes[scene_idx].used)
        throw Fail{SPECREL_EINVAL, "no scene uploaded at index " + std::to_string(scene_idx)};
    SceneSlot &sc = ctx->scenes[scene_idx];
    Plan pl;
    for (uint32_t i = 0; i < n; ++i) {
        const specrel_frame &f = frames[i];
        check_frame(&f, i);
        if (f.mode == SPECREL_MODE_SHADED) {
            pl.stars = pl.stars || sc.num_stars > 0;
            if (f.shadow_mode != SPECREL_SHADOW_OFF) {
                pl.queries = true;
                pl.max_flashes = std::max(pl.max_flashes, f.num_flashes);
            }
        }
    }
    pl.want_hits = want_hits || ctx->opt_stats != 0;
    ensure_views(ctx, frames, n, sc.num_stars);
    ctx->last_frames.assign(frames, frames + n);
    ctx->last_scene = scene_idx;
    if (pl.queries && ctx->opt_shadow_queue) {
        uint64_t cap = (uint64_t)ctx->total_pix * (1u + pl.max_flashes);
        if (cap > 0xFFFFFFFFull) throw Fail{SPECREL_EINVAL, "views too large for the shadow queue"};
        galloc(ctx, ctx->squeue, (size_t)cap * sizeof(ShadowQuery));
    }
    if (frame_in_reset) {  // k_reset_frame writes the one frame
        std::memcpy(pl.frame.w, frames, sizeof(specrel_frame));
        pl.frame_in_reset = true;
        return pl;
    }
    specrel_ctx::Stage &st = ctx->stage[ctx->stage_next++ % specrel_ctx::STAGES];
    const size_t bytes = (size_t)n * sizeof(specrel_frame);
    if (st.used) check(hipEventSynchronize(st.done), "frame staging");
    if (st.cap < bytes) {
        if (st.host) check(hipHostFree(st.host), "hipHostFree");
        st.host = nullptr;
        st.cap = 0;
        check(hipHostMalloc(&st.host, bytes, hipHostMallocDefault), "hipHostMalloc");
        st.cap = bytes;
    }
    std::memcpy(st.host, frames, bytes);
    check(hipMemcpyAsync(ctx->frames.p, st.host, bytes, hipMemcpyHostToDevice, ctx->stream), "frame upload");
    check(hipEventRecord(st.done, ctx->stream), "frame staging");
    st.used = true;
    return pl;
}

// The launch sequence of one render of the uploaded views; stage events only when `events`
// (never inside a graph capture).
void record_launches(specrel_ctx *ctx, SceneSlot &sc, const Plan &pl, bool events, const Bufs &B) {
    hipStream_t s = ctx->stream;
    const bool stats = ctx->opt_stats != 0;
    // Stage events (SPECREL_OPT_STAGE_EVENTS): 2 every stage boundary, 1 only the pair around the
    // primary-ray kernel, 0 none. Each record leaves the device idle for a few microseconds
    // between the kernels it separates.
    const int level = events ? ctx->opt_stage_events : 0;
    ctx->stage_level = level;
    const Views V = B.V;
    const uint32_t nv = B.n_views;
    uint32_t *counters = B.counters;
    Scene S = sc.S;
    if (stats) {  // distinct records touched (cleared in enqueue_views)
        uint32_t *t = (uint32_t *)ctx->touch.p;
        S.touch_nodes = t;
        S.touch_prims = t + ctx->touch_words[0];
        S.touch_hits = t + ctx->touch_words[0] + ctx->touch_words[1];
    }
    // The primary rays come first (they need only the frame); the star pass follows them, or runs
    // beside them on the side stream when overlapped. The shading pass waits for both.
    const bool overlap = pl.stars && ctx->opt_star_overlap;
    ctx->stars_overlapped = overlap;
    const dim3 tiles(B.total_tiles);
    const uint32_t path = shadow_path(ctx, pl);
    const bool queued = path == SPECREL_SHADOWS_QUEUED, in_trace = path == SPECREL_SHADOWS_IN_TRACE;
    const bool tiled = ctx->opt_stack > 0 && !ctx->opt_persistent;
    // The shading pass fused into the primary-ray kernel (for single views by default: their frame
    // time is their heaviest tiles' traversal, under which the other tiles' shading hides; a batch
    // of many light views is bound by throughput instead): tiled rays with an LDS stack of 8 or 16,
    // shadow queries (if any) traced there, no instrumentation, the star pass not overlapped (it
    // runs first instead, since the fused kernel shades with its lists).
    const bool fuse_asked = ctx->opt_fuse_shade == 2 || (ctx->opt_fuse_shade == 1 && nv == 1);
    const bool fuse = fuse_asked && !stats && tiled && (ctx->opt_stack == 8 || ctx->opt_stack == 16) &&
                      trace_variant(ctx, stats) != TV_MERGED && !queued && !overlap &&
                      (in_trace || path == SPECREL_SHADOWS_NONE);
    ctx->shade_fused = fuse;
    // One view launched directly whose star pass comes first (the fused kernel shades with its
    // lists): the star pass's first kernel also clears the counters and writes the frame.
    const bool reset_in_prep = pl.frame_in_reset && !B.table && fuse && pl.stars;
    // ev 0 and 1 bracket the primary rays when they come first (ev 10 and 1 when the fused kernel
    // follows the star pass); the others mark the later stages.
    auto mark = [&](int i) {
        if (level == 2 || (level == 1 && (i == 1 || (i == 0 && !fuse))))
            check(hipEventRecord(ctx->ev[i], s), "event");
    };
    if (B.table) {  // one view without staging: frame and view table from the kernel arguments
        k_reset_view<<<1, 256, 0, s>>>(counters, COUNTERS, pl.frame, B.frames, *B.table, B.table_dst);
    } else if (!reset_in_prep) {
        if (pl.frame_in_reset) k_reset_frame<<<1, 256, 0, s>>>(counters, COUNTERS, pl.frame, B.frames);
        else k_reset<<<1, 256, 0, s>>>(counters, COUNTERS);
    }
    check(hipGetLastError(), "k_reset launch");
    mark(0);
    if (overlap) check(hipEventRecord(ctx->fork, s), "fork");
    // order_tiles: k_star_bin also sorts the next render's tile order (one block's share of tiles).
    auto star_pass = [&](hipStream_t ss, uint32_t order_tiles) {
        if (reset_in_prep)
            k_star_prep_first<<<(sc.num_stars + 255) / 256, 256, 0, ss>>>((const float4 *)sc.stars.p, sc.num_stars,
                                                                           B.prep, B.dirs, counters, COUNTERS,
                                                                           pl.frame, B.frames);
        else
            k_star_prep<<<dim3((sc.num_stars + 255) / 256, nv), 256, 0, ss>>>(
                (const float4 *)sc.stars.p, sc.num_stars, V, B.prep, B.dirs);
        check(hipGetLastError(), "k_star_prep launch");
        if (ctx->opt_cull) {
            k_star_bin<<<B.total_supers + (order_tiles ? 1u : 0u), 256, 0, ss>>>(
                V, B.dirs, sc.num_stars, B.super_lists, B.super_counts, B.tile_cones, B.tile_cost, order_tiles,
                B.tile_order);
            check(hipGetLastError(), "k_star_bin launch");
        }
    };
    ShadeArgs ra;
    ra.S = S;
    ra.T.lut = (const uint8_t *)ctx->tables.p;
    ra.T.glyphs = (const uint8_t *)ctx->tables.p + SPECREL_LUT_SIZE;
    ra.V = V;
    ra.stars = B.prep;
    ra.dirs = B.dirs;
    ra.super_lists = B.super_lists;
    ra.super_counts = B.super_counts;
    ra.tile_cones = B.tile_cones;
    ra.num_stars = sc.num_stars;
    ra.cull = ctx->opt_cull;
    ra.want_hits = pl.want_hits ? 1 : 0;
    ra.slot = B.slot;
    ra.hdr = B.hdr;
    ra.hit = B.hit;
    ra.lambda = B.lam;
    ra.tau = B.tau;
    ra.mask = B.mask;
    ra.queue = nullptr;
    ra.qcount = counters + 1;
    ra.qcap = 0;
    ra.stats = B.stats;
    TraceArgs ta;
    ta.S = S;
    if (trace_variant(ctx, stats) == TV_LDSTOP) ta.S.top_n = sc.top_count[std::min(ctx->opt_lds_top, TOP_LEVELS_MAX)];
    ta.V = V;
    ta.num_tiles = B.total_tiles;
    ta.slot = B.slot;
    ta.mask = in_trace || fuse ? B.mask : nullptr;
    ta.queue = counters + 0;
    ta.stats = B.stats;
    ta.wave_shape = (uint32_t)ctx->opt_wave_shape;
    // Slowest tiles first only when they cannot all start at once (more than four per CU): then
    // the order decides when the slowest ones start; otherwise it only costs.
    const bool order = B.tile_order && ctx->opt_tile_order && !ctx->opt_persistent &&
                       B.total_tiles > 4u * (uint32_t)ctx->cus;
    ta.order = order ? B.tile_order : nullptr;
    ta.cost = order ? B.tile_cost : nullptr;
    ta.prio_slots = order ? (uint32_t)ctx->cus : 0u;
    ta.postpone = ctx->opt_postpone == 1 ? (uint32_t)(STACK - 2) : (uint32_t)ctx->opt_postpone;
    // Per-wave clock records: the tiled kernel only (one record per scheduling slot).
    ta.clock = tiled ? B.wave_clock : nullptr;
    ctx->wave_clock_n = ta.clock ? 4ull * B.total_tiles : 0ull;
    ta.shade = ra;
    ta.shade_here = fuse;
    // The tile order: sorted by a leading block of the star binning pass when that pass runs on
    // this stream with one block's share of tiles (before the fused kernel: the previous render's
    // costs, for this one; after the primary rays: this render's costs, for the next one), else
    // on the side stream beside the later passes (this render's costs, for the next one).
    const bool order_in_bin =
        order && pl.stars && !overlap && ctx->opt_cull && B.total_tiles <= TILE_ORDER_ONE_BLOCK;
    const bool order_side = order && !order_in_bin;
    if (fuse) {
        if (pl.stars) star_pass(s, order_in_bin ? B.total_tiles : 0u);
        if (level >= 1) check(hipEventRecord(ctx->ev[10], s), "event");
    }
    launch_any(ctx, stats, tiles, s, ta);
    check(hipGetLastError(), "trace launch");
    mark(1);
    if (overlap) {
        check(hipStreamWaitEvent(ctx->side, ctx->fork, 0), "fork wait");
        if (level == 2) check(hipEventRecord(ctx->ev[9], ctx->side), "event");
        star_pass(ctx->side, 0u);
        if (level == 2) check(hipEventRecord(ctx->ev[10], ctx->side), "event");
        check(hipEventRecord(ctx->join, ctx->side), "join");
    }
    if (order_side) {
        check(hipEventRecord(ctx->order_fork, s), "fork");
        check(hipStreamWaitEvent(ctx->side, ctx->order_fork, 0), "fork wait");
        const uint32_t *cost = B.tile_cost, n = B.total_tiles;
        uint32_t *torder = B.tile_order;
        if (n <= TILE_ORDER_ONE_BLOCK) {
            k_tile_order<<<1, 1024, 0, ctx->side>>>(cost, n, torder);
        } else {
            const uint32_t nb = (n + TILE_ORDER_CHUNK - 1) / TILE_ORDER_CHUNK;
            k_tile_hist<<<nb, 256, 0, ctx->side>>>(cost, n, counters + COUNTER_TILE_TOTAL);
            k_tile_scatter<<<nb, 256, 0, ctx->side>>>(cost, n, counters + COUNTER_TILE_TOTAL,
                                                      counters + COUNTER_TILE_CURSOR, torder);
        }
        check(hipGetLastError(), "k_tile_order launch");
        check(hipEventRecord(ctx->order_join, ctx->side), "join");
    }
    if (pl.stars && !overlap && !fuse) star_pass(s, order_in_bin ? B.total_tiles : 0u);
    mark(2);
    if (overlap) check(hipStreamWaitEvent(s, ctx->join, 0), "join wait");
    if (fuse) {
        // shaded by the primary-ray kernel
    } else if (queued) {
        ra.queue = B.squeue;
        ra.qcap = (uint32_t)((uint64_t)B.total_pix * (1u + pl.max_flashes));
        launch_any(ctx, stats, tiles, s, ShadePass<1>{ra});
        check(hipGetLastError(), "k_shade setup launch");
        OccludeArgs oa;
        oa.S = S;
        oa.queue = ra.queue;
        oa.count = counters + 1;
        oa.next = counters + 2;
        oa.cap = ra.qcap;
        oa.mask = ra.mask;
        oa.stats = ra.stats;
        launch_any(ctx, stats, tiles, s, oa);
        check(hipGetLastError(), "k_occlude launch");
        launch_any(ctx, stats, tiles, s, ShadePass<2>{ra});
    } else if (in_trace) {
        launch_any(ctx, stats, tiles, s, ShadePass<2>{ra});
    } else {
        launch_any(ctx, stats, tiles, s, ShadePass<0>{ra});
    }
    check(hipGetLastError(), "k_shade launch");
    mark(3);
    float4 *pyr = B.hdr;
    // Levels whose every view holds at most BLOOM_TAIL_PIXELS pixels are done by k_bloom_tail.
    int k0 = SPECREL_BLOOM_LEVELS;
    while (k0 > 1) {
        bool small = true;
        for (uint32_t v = 0; v < nv && small; ++v) {
            uint32_t lw, lh;
            level_size(B.sizes[2 * v], B.sizes[2 * v + 1], k0 - 1, lw, lh);
            small = (uint64_t)lw * lh <= BLOOM_TAIL_PIXELS;
        }
        if (!small) break;
        --k0;
    }
    // Down pass: levels below k0 two at a time (one if a single one is left), then the tail.
    for (int k = 1; k < k0;) {
        if (k + 1 < k0) {
            const uint32_t tiles = B.start(k, nv);
            if (tiles > BLOOM_STAGE_TILES) k_bloom_down2<true><<<tiles, 256, 0, s>>>(V, k, pyr);
            else k_bloom_down2<false><<<tiles, 256, 0, s>>>(V, k, pyr);
            k += 2;
        } else {
            k_bloom_down<<<B.start(k, nv), 256, 0, s>>>(V, k, pyr);
            k += 1;
        }
    }
    k_bloom_tail<<<nv, 1024, 0, s>>>(V, k0, pyr);
    // Up pass from level k0 - 1 down: an odd one out alone, then pairs, the last two levels (or
    // the one) with the composite.
    int top = k0 - 1;
    const int nup = std::min(top, 2);
    if ((top - nup) % 2 == 1) {
        k_bloom_up<<<B.start(top, nv), 256, 0, s>>>(V, top, pyr);
        --top;
    }
    for (; top > nup; top -= 2) k_bloom_up2<<<B.start(top - 1, nv), 256, 0, s>>>(V, top - 1, pyr);
    check(hipGetLastError(), "bloom launch");
    mark(4);
    const uint8_t *lut = (const uint8_t *)ctx->tables.p;
    if (nup == 2) k_bloom_final<2><<<tiles, 256, 0, s>>>(V, pyr, lut, B.rgb);
    else if (nup == 1) k_bloom_final<1><<<tiles, 256, 0, s>>>(V, pyr, lut, B.rgb);
    else k_bloom_final<0><<<tiles, 256, 0, s>>>(V, pyr, lut, B.rgb);
    check(hipGetLastError(), "k_bloom_final launch");
    if (order_side) check(hipStreamWaitEvent(s, ctx->order_join, 0), "join wait");
    mark(5);
}

void drop_graph(specrel_ctx *ctx) {
    if (ctx->graph) (void)hipGraphExecDestroy(ctx->graph);
    ctx->graph = nullptr;
    ctx->graph_key.clear();
}

// Enqueues one render of n views. Without stats, batches (or every render, with
// SPECREL_OPT_GRAPH = 2) replay a captured graph of the launch sequence, captured again only when
// the view sizes, options or any buffer it references changed.
int enqueue_views(specrel_ctx *ctx, uint32_t scene_idx, const specrel_frame *frames, uint32_t n, bool want_hits) {
    auto t0 = std::chrono::steady_clock::now();
    const bool stats = ctx->opt_stats != 0;
    const bool use_graph =
        !stats && !ctx->opt_wave_clock && (ctx->opt_graph == 2 || (ctx->opt_graph == 1 && n > 1));
    const Plan pl = prepare(ctx, scene_idx, frames, n, want_hits, n == 1 && !use_graph);
    SceneSlot &sc = ctx->scenes[scene_idx];
    hipStream_t s = ctx->stream;
    if (ctx->opt_wave_clock) dalloc(ctx->wave_clock, (size_t)ctx->total_tiles * 4 * sizeof(specrel_wave_clock));
    if (stats) {
        check(hipMemsetAsync(ctx->stats.p, 0, 16 * 8, s), "stats clear");
        ctx->touch_words[0] = (sc.num_nodes + 31) / 32;
        ctx->touch_words[1] = ctx->touch_words[2] = (sc.num_prims + 31) / 32;
        const size_t words = (size_t)ctx->touch_words[0] + ctx->touch_words[1] + ctx->touch_words[2];
        dalloc(ctx->touch, words * 4);
        check(hipMemsetAsync(ctx->touch.p, 0, words * 4, s), "touch clear");
    }
    if (use_graph) {
        std::vector<uint64_t> key = {ctx->generation,   scene_idx,        (uint64_t)ctx->opt_cull,
                                     (uint64_t)(int64_t)ctx->opt_stack, (uint64_t)ctx->opt_persistent,
                                     (uint64_t)ctx->opt_shadow_queue,   (uint64_t)ctx->opt_shadow_trace,
                                     (uint64_t)ctx->opt_merged,         (uint64_t)ctx->opt_trace_waves,
                                     (uint64_t)ctx->opt_star_overlap,   (uint64_t)ctx->opt_wave_shape,
                                     (uint64_t)ctx->opt_tile_order,   (uint64_t)ctx->opt_postpone,
                                     (uint64_t)ctx->opt_node_width,   (uint64_t)ctx->opt_lds_top,
                                     (uint64_t)ctx->opt_fuse_shade,
                                     pl.stars,         pl.queries,
                                     pl.max_flashes,    pl.want_hits};
        key.insert(key.end(), ctx->sizes.begin(), ctx->sizes.end());
        if (!ctx->graph || key != ctx->graph_key) {
            drop_graph(ctx);
            check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            try {
                record_launches(ctx, sc, pl, false, ctx_bufs(ctx));
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            hipGraph_t g = nullptr;
            check(hipStreamEndCapture(s, &g), "hipStreamEndCapture");
            hipError_t e = hipGraphInstantiate(&ctx->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                ctx->graph = nullptr;
                check(e, "hipGraphInstantiate");
            }
            ctx->graph_key = key;
        }
        check(hipEventRecord(ctx->ev[0], s), "event");
        check(hipGraphLaunch(ctx->graph, s), "hipGraphLaunch");
        check(hipEventRecord(ctx->ev[5], s), "event");
        ctx->timed_stages = false;
    } else {
        record_launches(ctx, sc, pl, true, ctx_bufs(ctx));
        ctx->timed_stages = true;
    }
    auto t1 = std::chrono::steady_clock::now();
    ctx->timings = specrel_timings{};
    ctx->timings.submit_ms = std::chrono::duration<float, std::milli>(t1 - t0).count();
    ctx->timings.shadow_path = shadow_path(ctx, pl);
    ctx->timings.trace_block_waves = trace_block_waves(ctx);
    ctx->timings.trace_variant = (uint32_t)trace_variant(ctx, ctx->opt_stats != 0);
    ctx->have_hits = pl.want_hits;
    ctx->last_noalloc = false;
    return SPECREL_OK;
}

// Byte offsets of one view's intermediate buffers in the scratch area (256-B aligned pieces).
struct ScratchLayout {
    size_t frame, table, counters, slot, mask, hdr, super_counts, super_lists, tile_cones, prep, dirs, total;
};

ScratchLayout scratch_layout(uint64_t pixels, uint64_t pyr, uint64_t tiles, uint64_t supers, uint64_t stars) {
    ScratchLayout L;
    size_t at = 0;
    auto take = [&](uint64_t bytes) {
        const size_t here = at;
        at += (size_t)((bytes + 255) & ~(uint64_t)255);
        return here;
    };
    L.frame = take(sizeof(specrel_frame));
    L.table = take(sizeof(TableWords));
    L.counters = take(COUNTERS * 4);
    L.slot = take(pixels * 4);
    L.mask = take(pixels * 4);
    L.hdr = take(pyr * sizeof(float4));
    L.super_counts = take(supers * 4);
    L.super_lists = take(supers * SUPER_CAP * 4);
    L.tile_cones = take(tiles * 16);
    L.prep = take(std::max<uint64_t>(stars, 1) * sizeof(StarPrep));
    L.dirs = take(std::max<uint64_t>(stars, 1) * sizeof(StarDir));
    L.total = at;
    return L;
}

// specrel_render_noalloc: one view, enqueued, every intermediate buffer in this device's scratch
// area and the outputs in the caller's device buffers. No allocation, no host-to-device copy (the
// frame and the view table travel as kernel arguments), tiles in row order (a tile order would
// have to survive between calls in memory every call shares).
void enqueue_noalloc(specrel_ctx *ctx, uint32_t scene_idx, const specrel_frame *frame, uint32_t *d_rgb,
                     uint32_t *d_hit, float *d_lambda, float *d_tau, uint64_t len) {
    auto t0 = std::chrono::steady_clock::now();
    if (!frame) throw Fail{SPECREL_EINVAL, "frame is NULL"};
    if (!d_rgb) throw Fail{SPECREL_EINVAL, "d_rgb is NULL"};
    const bool hits = d_hit || d_lambda || d_tau;
    if (hits && !(d_hit && d_lambda && d_tau)) throw Fail{SPECREL_EINVAL, "d_hit, d_lambda and d_tau: all or none"};
    check_frame(frame, 0);
    const uint32_t w = frame->size[0], h = frame->size[1];
    check_len(len, (uint64_t)w * h, "framebuffer", 0);
    if (scene_idx >= ctx->scenes.size() || !ctx->scenes[scene_idx].used)
        throw Fail{SPECREL_EINVAL, "no scene uploaded at index " + std::to_string(scene_idx)};
    if (ctx->opt_stats || ctx->opt_shadow_queue)
        throw Fail{SPECREL_EINVAL, "SPECREL_OPT_STATS and SPECREL_OPT_SHADOW_QUEUE need buffers the "
                                   "allocation-free render does not have"};
    if (!ctx->scratch) throw Fail{SPECREL_EDEVICE, "no scratch area on this device"};
    SceneSlot &sc = ctx->scenes[scene_idx];
    // The view table of this one view: rows of (0, total).
    TableWords tw{};
    uint64_t tiles = 0, pyr = pyramid_len(w, h);
    for (int k = 0; k < NLEV; ++k) {
        uint32_t lw, lh;
        level_size(w, h, k, lw, lh);
        const uint64_t t = (uint64_t)((lw + 15) / 16) * ((lh + 15) / 16);
        if (k == 0) tiles = t;
        tw.w[2 * k + 1] = (uint32_t)t;
    }
    const uint64_t supers = num_super(w, h), pixels = (uint64_t)w * h;
    tw.w[2 * ROW_SUPER + 1] = (uint32_t)supers;
    tw.w[2 * ROW_PIX + 1] = (uint32_t)pixels;
    tw.w[2 * ROW_PYR + 1] = (uint32_t)pyr;
    const ScratchLayout L = scratch_layout(pixels, pyr, tiles, supers, sc.num_stars);
    if (L.total > sizeof(scratch))
        throw Fail{SPECREL_EINVAL, "a " + std::to_string(w) + "x" + std::to_string(h) + " view with " +
                                       std::to_string(sc.num_stars) + " stars needs " + std::to_string(L.total) +
                                       " bytes of scratch; the allocation-free render has " +
                                       std::to_string(sizeof(scratch))};
    Plan pl;
    if (frame->mode == SPECREL_MODE_SHADED) {
        pl.stars = sc.num_stars > 0;
        if (frame->shadow_mode != SPECREL_SHADOW_OFF) {
            pl.queries = true;
            pl.max_flashes = frame->num_flashes;
        }
    }
    pl.want_hits = hits;
    std::memcpy(pl.frame.w, frame, sizeof(specrel_frame));
    pl.frame_in_reset = true;
    char *base = (char *)ctx->scratch;
    uint32_t *table = (uint32_t *)(base + L.table);
    const uint32_t sizes[2] = {w, h};
    Bufs B;
    B.V.frames = (const specrel_frame *)(base + L.frame);
    for (int k = 0; k < NLEV; ++k) B.V.level_start[k] = table + 2 * k;
    B.V.super_start = table + 2 * ROW_SUPER;
    B.V.pix_start = table + 2 * ROW_PIX;
    B.V.pyr_start = table + 2 * ROW_PYR;
    B.V.n = 1;
    B.n_views = 1;
    B.total_tiles = (uint32_t)tiles;
    B.total_supers = (uint32_t)supers;
    B.total_pix = (uint32_t)pixels;
    B.tab_host = tw.w;
    B.sizes = sizes;
    B.frames = (uint32_t *)(base + L.frame);
    B.table = &tw;
    B.table_dst = table;
    B.counters = (uint32_t *)(base + L.counters);
    B.slot = (uint32_t *)(base + L.slot);
    B.mask = (uint32_t *)(base + L.mask);
    B.hdr = (float4 *)(base + L.hdr);
    B.super_counts = (uint32_t *)(base + L.super_counts);
    B.super_lists = (uint32_t *)(base + L.super_lists);
    B.tile_cones = (float4 *)(base + L.tile_cones);
    B.prep = (StarPrep *)(base + L.prep);
    B.dirs = (StarDir *)(base + L.dirs);
    B.rgb = d_rgb;
    B.hit = d_hit;
    B.lam = d_lambda;
    B.tau = d_tau;
    record_launches(ctx, sc, pl, true, B);
    ctx->timed_stages = true;
    auto t1 = std::chrono::steady_clock::now();
    ctx->timings = specrel_timings{};
    ctx->timings.submit_ms = std::chrono::duration<float, std::milli>(t1 - t0).count();
    ctx->timings.shadow_path = shadow_path(ctx, pl);
    ctx->timings.trace_block_waves = trace_block_waves(ctx);
    ctx->timings.trace_variant = (uint32_t)trace_variant(ctx, ctx->opt_stats != 0);
    ctx->last_noalloc = true;
}

// The pinned readback buffer, grown to at least `bytes` (the stream is idle when it moves).
char *readback_buf(specrel_ctx *ctx, size_t bytes) {
    if (ctx->readback_cap < bytes) {
        check(hipStreamSynchronize(ctx->stream), "readback");
        if (ctx->readback) check(hipHostFree(ctx->readback), "hipHostFree");
        ctx->readback = nullptr;
        ctx->readback_cap = 0;
        const size_t cap = std::max<size_t>(bytes, (size_t)1 << 20);
        check(hipHostMalloc(&ctx->readback, cap, hipHostMallocDefault), "hipHostMalloc");
        ctx->readback_cap = cap;
    }
    return (char *)ctx->readback;
}

// One device-to-host copy of a readback: `bytes` from device `src` into host `dst`.
struct Fetch {
    void *dst;
    const void *src;
    size_t bytes;
};

// Enqueues the copies into the pinned readback buffer behind whatever the stream holds, waits
// once, then fills the callers' buffers. Returns the wait and the host copy in milliseconds.
void fetch(specrel_ctx *ctx, const std::vector<Fetch> &f, float *wait_ms, float *copy_ms) {
    size_t total = 0;
    for (const Fetch &x : f) total += (x.bytes + 255) & ~(size_t)255;
    char *stage = readback_buf(ctx, total);
    size_t off = 0;
    for (const Fetch &x : f) {
        if (x.bytes) check(hipMemcpyAsync(stage + off, x.src, x.bytes, hipMemcpyDeviceToHost, ctx->stream), "readback");
        off += (x.bytes + 255) & ~(size_t)255;
    }
    auto t0 = std::chrono::steady_clock::now();
    check(hipStreamSynchronize(ctx->stream), "readback");
    auto t1 = std::chrono::steady_clock::now();
    off = 0;
    for (const Fetch &x : f) {
        if (x.bytes) std::memcpy(x.dst, stage + off, x.bytes);
        off += (x.bytes + 255) & ~(size_t)255;
    }
    auto t2 = std::chrono::steady_clock::now();
    if (wait_ms) *wait_ms = std::chrono::duration<float, std::milli>(t1 - t0).count();
    if (copy_ms) *copy_ms = std::chrono::duration<float, std::milli>(t2 - t1).count();
}

void collect(specrel_ctx *ctx) {
    specrel_timings &t = ctx->timings;
    const int level = ctx->timed_stages ? ctx->stage_level : 2;  // a graph replay: ev 0 and 5 around it
    if (ctx->timed_stages && level == 1) {
        t.trace_ms = ctx->shade_fused ? ev_ms(ctx->ev[10], ctx->ev[1]) : ev_ms(ctx->ev[0], ctx->ev[1]);
    } else if (ctx->timed_stages && level == 2) {
        if (ctx->shade_fused) {  // star pass, then primary rays with the shading (ev 0, 10, 1)
            t.star_prep_ms = ev_ms(ctx->ev[0], ctx->ev[10]);
            t.trace_ms = ev_ms(ctx->ev[10], ctx->ev[1]);
            t.shade_ms = 0.0f;
        } else {
            t.trace_ms = ev_ms(ctx->ev[0], ctx->ev[1]);
            t.star_prep_ms = ctx->stars_overlapped ? ev_ms(ctx->ev[9], ctx->ev[10]) : ev_ms(ctx->ev[1], ctx->ev[2]);
            t.shade_ms = ev_ms(ctx->ev[2], ctx->ev[3]);
        }
        t.bloom_ms = ev_ms(ctx->ev[3], ctx->ev[4]);
        t.composite_ms = ev_ms(ctx->ev[4], ctx->ev[5]);
    }
    if (level == 2) t.kernel_ms = ev_ms(ctx->ev[0], ctx->ev[5]);
    if (ctx->opt_stats) {
        unsigned long long *dst = (unsigned long long *)ctx->stats.p;
        const uint32_t *bits = (const uint32_t *)ctx->touch.p;
        for (int k = 0; k < 3; ++k) {
            const uint32_t n = ctx->touch_words[k];
            if (n) k_popcount<<<std::min<uint32_t>((n + 255) / 256, 1024), 256, 0, ctx->stream>>>(bits, n, dst + 10 + k);
            check(hipGetLastError(), "k_popcount launch");
            bits += n;
        }
        unsigned long long c[16];
        fetch(ctx, {{c, ctx->stats.p, sizeof c}}, nullptr, nullptr);
        specrel_stats &st = ctx->stats_host;
        std::memset(&st, 0, sizeof st);
        st.rays = c[0];
        st.node_pops = c[1];
        st.box_tests = c[2];
        st.prim_tests = c[3];
        st.hits = c[4];
        st.shadow_rays = c[5];
        st.shadow_node_pops = c[6];
        st.shadow_box_tests = c[7];
        st.shadow_prim_tests = c[8];
        st.star_tests = c[9];
        st.distinct_nodes = c[10];
        st.distinct_prims = c[11];
        st.distinct_hits = c[12];
    }
}


// The readback copies of view `view` of the last render (NULL pointers are skipped); every
// non-NULL buffer holds `len` elements, which must be the view's width * height.
void view_fetches(specrel_ctx *ctx, uint32_t view, uint32_t *rgb, uint32_t *hit, float *lam, float *tau, uint64_t len,
                  std::vector<Fetch> &f) {
    if (ctx->n_views == 0) throw Fail{SPECREL_EINVAL, "nothing rendered yet"};
    if (ctx->last_noalloc) throw Fail{SPECREL_EINVAL, NOALLOC_LAST};
    if (view >= ctx->n_views) throw Fail{SPECREL_EINVAL, "view index out of range"};
    if ((hit || lam || tau) && !ctx->have_hits) throw Fail{SPECREL_EINVAL, "the last render did not write hit buffers"};
    const size_t off = host_start(ctx, ROW_PIX, view);
    const size_t n = (size_t)ctx->sizes[2 * view] * ctx->sizes[2 * view + 1];
    if (rgb || hit || lam || tau) check_len(len, n, "framebuffer", view);
    if (rgb) f.push_back({rgb, (const uint32_t *)ctx->rgb.p + off, n * 4});
    if (hit) f.push_back({hit, (const uint32_t *)ctx->hit.p + off, n * 4});
    if (lam) f.push_back({lam, (const float *)ctx->lam.p + off, n * 4});
    if (tau) f.push_back({tau, (const float *)ctx->tau.p + off, n * 4});
}

// Copies view `view` of the last render to the host.
void download_view(specrel_ctx *ctx, uint32_t view, uint32_t *rgb, uint32_t *hit, float *lam, float *tau,
                   uint64_t len) {
    std::vector<Fetch> f;
    view_fetches(ctx, view, rgb, hit, lam, tau, len, f);
    fetch(ctx, f, nullptr, nullptr);
}

// specrel_adjoint / specrel_adjoint_batch on the views of the last render: per-view gradients,
// one launch per pass over every view's tiles.
int adjoint_views(specrel_ctx *ctx, const float *target, const uint8_t *mask_in, uint8_t *mask_out, uint64_t n_pixels,
                  const specrel_adj_object *objects, uint32_t n_objects, specrel_adjoint_out *out, double *d_objects) {
    check(hipSetDevice(ctx->device), "hipSetDevice");
    if (!target) throw Fail{SPECREL_EINVAL, "target is NULL"};
    if (ctx->n_views == 0) throw Fail{SPECREL_EINVAL, "nothing rendered yet"};
    if (n_pixels != ctx->total_pix)
        throw Fail{SPECREL_EINVAL, "target size mismatch: " + std::to_string(n_pixels) + " pixels given, the last "
                                   "render has " + std::to_string(ctx->total_pix)};
    if (n_objects && (!objects || !d_objects)) throw Fail{SPECREL_EINVAL, "objects or d_objects is NULL"};
    if (n_objects > 1024) throw Fail{SPECREL_EINVAL, "at most 1024 objects"};
    const uint32_t nv = ctx->n_views;
    if (nv == 0 || ctx->last_frames.size() != nv) throw Fail{SPECREL_EINVAL, "nothing rendered yet"};
    if (ctx->last_noalloc) throw Fail{SPECREL_EINVAL, NOALLOC_LAST};
    for (uint32_t v = 0; v < nv; ++v) {
        const specrel_frame &f = ctx->last_frames[v];
        if (f.mode != SPECREL_MODE_SHADED || f.slice_mode != SPECREL_VIEW_PAST_CONE ||
            f.shadow_mode == SPECREL_SHADOW_CAMERA_SLICE || f.num_flashes != 0 || f.num_haze != 0)
            throw Fail{SPECREL_EINVAL, "view " + std::to_string(v) + ": the adjoint covers shaded past-light-cone "
                                       "views with null or no shadows, without flashes or haze"};
    }
    if (ctx->last_scene >= ctx->scenes.size() || !ctx->scenes[ctx->last_scene].used)
        throw Fail{SPECREL_EINVAL, "the last render's scene is gone"};
    SceneSlot &sc = ctx->scenes[ctx->last_scene];
    std::vector<uint32_t> ofirst(n_objects), ocount(n_objects);
    std::vector<float4> ocenter(n_objects);
    for (uint32_t k = 0; k < n_objects; ++k) {
        const specrel_adj_object &o = objects[k];
        if (o.prim_count == 0 || (uint64_t)o.prim_first + o.prim_count > sc.num_prims)
            throw Fail{SPECREL_EINVAL, "object " + std::to_string(k) + " has a bad primitive range"};
        if (k && o.prim_first < (uint64_t)objects[k - 1].prim_first + objects[k - 1].prim_count)
            throw Fail{SPECREL_EINVAL, "objects must be sorted by prim_first and disjoint"};
        ofirst[k] = o.prim_first;
        ocount[k] = o.prim_count;
        ocenter[k] = make_float4(o.center[0], o.center[1], o.center[2], o.center[3]);
    }
    const size_t np = ctx->total_pix;
    const uint32_t tiles = ctx->total_tiles, stride = (uint32_t)(ADJ_CAM + 3 * n_objects);
    hipStream_t s = ctx->stream;
    dalloc(ctx->adj_target, np * 3 * sizeof(float));
    dalloc(ctx->adj_pix, np * ADJ_PIX * sizeof(double));
    dalloc(ctx->adj_sig, np * 4);
    dalloc(ctx->adj_objid, np * 4);
    dalloc(ctx->adj_inside, np);
    dalloc(ctx->adj_partials, (size_t)tiles * stride * sizeof(double));
    dalloc(ctx->adj_counts, (size_t)tiles * ADJ_CLASSES * 4);
    dalloc(ctx->adj_result, (size_t)nv * (stride * sizeof(double) + ADJ_CLASSES * 8));
    dalloc(ctx->adj_obj, (size_t)n_objects * 24 + 16);
    if (mask_in) dalloc(ctx->adj_mask_in, np);
    if (mask_out) dalloc(ctx->adj_mask_out, np);
    check(hipMemcpyAsync(ctx->adj_target.p, target, np * 3 * sizeof(float), hipMemcpyHostToDevice, s), "target");
    if (mask_in) check(hipMemcpyAsync(ctx->adj_mask_in.p, mask_in, np, hipMemcpyHostToDevice, s), "mask");
    float4 *d_center = (float4 *)ctx->adj_obj.p;
    uint32_t *d_first = (uint32_t *)(d_center + n_objects), *d_count = d_first + n_objects;
    if (n_objects) {
        check(hipMemcpyAsync(d_center, ocenter.data(), n_objects * 16, hipMemcpyHostToDevice, s), "objects");
        check(hipMemcpyAsync(d_first, ofirst.data(), n_objects * 4, hipMemcpyHostToDevice, s), "objects");
        check(hipMemcpyAsync(d_count, ocount.data(), n_objects * 4, hipMemcpyHostToDevice, s), "objects");
    }
    AdjArgs a;
    a.S = sc.S;
    a.T.lut = (const uint8_t *)ctx->tables.p;
    a.T.glyphs = (const uint8_t *)ctx->tables.p + SPECREL_LUT_SIZE;
    a.V = device_views(ctx);
    a.slot = (const uint32_t *)ctx->slot.p;
    a.hdr = (const float4 *)ctx->hdr.p;
    a.target = (const float *)ctx->adj_target.p;
    a.mask_in = mask_in ? (const uint8_t *)ctx->adj_mask_in.p : nullptr;
    a.mask_out = mask_out ? (uint8_t *)ctx->adj_mask_out.p : nullptr;
    a.obj_first = d_first;
    a.obj_count = d_count;
    a.obj_center = d_center;
    a.n_obj = n_objects;
    a.pix = (double *)ctx->adj_pix.p;
    a.sig = (uint32_t *)ctx->adj_sig.p;
    a.obj = (uint32_t *)ctx->adj_objid.p;
    a.inside = (uint8_t *)ctx->adj_inside.p;
    a.strict = ctx->opt_adj_strict ? 1u : 0u;
    a.partials = (double *)ctx->adj_partials.p;
    a.block_counts = (uint32_t *)ctx->adj_counts.p;
    a.stride = stride;
    a.result = (double *)ctx->adj_result.p;
    a.result_counts = (unsigned long long *)(a.result + (size_t)nv * stride);
    check(hipEventRecord(ctx->ev[7], s), "event");
    k_adj_pixel<16><<<tiles, 256, 0, s>>>(a);
    check(hipGetLastError(), "k_adj_pixel launch");
    k_adj_reduce<<<tiles, 256, 0, s>>>(a);
    check(hipGetLastError(), "k_adj_reduce launch");
    k_adj_final<<<dim3(nv, (stride + ADJ_CLASSES + 15) / 16), 256, 0, s>>>(a);
    check(hipGetLastError(), "k_adj_final launch");
    check(hipEventRecord(ctx->ev[8], s), "event");
    std::vector<double> res((size_t)nv * stride);
    std::vector<uint64_t> cnt((size_t)nv * ADJ_CLASSES);
    std::vector<Fetch> f = {{res.data(), a.result, res.size() * sizeof(double)},
                            {cnt.data(), a.result_counts, cnt.size() * 8}};
    if (mask_out) f.push_back({mask_out, ctx->adj_mask_out.p, np});
    fetch(ctx, f, nullptr, nullptr);
    ctx->timings.adjoint_ms = ev_ms(ctx->ev[7], ctx->ev[8]);
    for (uint32_t v = 0; v < nv; ++v) {
        const double *r = res.data() + (size_t)v * stride;
        const uint64_t *c = cnt.data() + (size_t)v * ADJ_CLASSES;
        if (out) {
            specrel_adjoint_out &o = out[v];
            std::memset(&o, 0, sizeof o);
            o.loss = r[0];
            for (int k = 0; k < 4; ++k) o.d_event[k] = r[1 + k];
            for (int e = 0; e < 4; ++e)
                for (int k = 0; k < 4; ++k) o.d_tetrad[e][k] = r[5 + 4 * e + k];
            o.valid = c[0];
            o.miss = c[1];
            o.starlight = c[2];
            o.border = c[3];
            o.silhouette = c[4];
            o.edge = c[5];
        }
        for (uint32_t k = 0; k < 3 * n_objects; ++k) d_objects[(size_t)v * 3 * n_objects + k] = r[ADJ_CAM + k];
    }
    return SPECREL_OK;
}

}  // namespace

extern "C" {

int specrel_create(int device_ordinal, specrel_ctx **out) {
    if (!out) return SPECREL_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        g_create_err = "no HIP device";
        return SPECREL_ENODEV;
    }
    if (device_ordinal < 0 || device_ordinal >= n) {
        g_create_err = "device ordinal out of range";
        return SPECREL_ENODEV;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device_ordinal) != hipSuccess) {
        g_create_err = "hipGetDeviceProperties failed";
        return SPECREL_ENODEV;
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        g_create_err = std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950";
        return SPECREL_ENODEV;
    }
    specrel_ctx *ctx = new specrel_ctx();
    int rc = guarded(ctx, [&] {
        ctx->device = device_ordinal;
        ctx->name = std::string(prop.name) + " (" + prop.gcnArchName + ", HIP)";
        check(hipSetDevice(device_ordinal), "hipSetDevice");
        ctx->scratch = scratch_address(device_ordinal);
        check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
        // The stage events only time (the host waits on the stream, not on them): no system-scope
        // cache writeback and invalidation at each record.
        for (auto &e : ctx->ev) check(hipEventCreateWithFlags(&e, hipEventDisableSystemFence), "hipEventCreate");
        check(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking), "hipStreamCreate");
        check(hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming), "hipEventCreate");
        check(hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming), "hipEventCreate");
        check(hipEventCreateWithFlags(&ctx->order_fork, hipEventDisableTiming), "hipEventCreate");
        check(hipEventCreateWithFlags(&ctx->order_join, hipEventDisableTiming), "hipEventCreate");
        for (auto &st : ctx->stage) check(hipEventCreateWithFlags(&st.done, hipEventDisableTiming), "hipEventCreate");
        std::vector<uint32_t> words(SPECREL_LUT_SIZE + 14 * 7);
        sr::build_tables(words.data());
        std::vector<uint8_t> bytes(words.size());
        for (size_t i = 0; i < words.size(); ++i) bytes[i] = (uint8_t)words[i];
        upload(ctx->tables, bytes.data(), bytes.size(), ctx->stream);
        dalloc(ctx->stats, 16 * 8);
        dalloc(ctx->counters, COUNTERS * 4);
        ctx->cus = prop.multiProcessorCount;
        if (hipDeviceGetAttribute(&ctx->clock_khz, hipDeviceAttributeWallClockRate, ctx->device) != hipSuccess)
            ctx->clock_khz = 0;
        check(hipStreamSynchronize(ctx->stream), "create");
        return SPECREL_OK;
    });
    if (rc != SPECREL_OK) {
        g_create_err = ctx->err;
        specrel_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return SPECREL_OK;
}

void specrel_destroy(specrel_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &sl : ctx->scenes) {
        dfree(sl.nbox), dfree(sl.ninfo), dfree(sl.big), dfree(sl.geoms), dfree(sl.attrs), dfree(sl.order);
        dfree(sl.mats), dfree(sl.stars);
    }
    dfree(ctx->hdr), dfree(ctx->rgb), dfree(ctx->hit), dfree(ctx->lam), dfree(ctx->tau), dfree(ctx->prep);
    dfree(ctx->slot), dfree(ctx->dirs), dfree(ctx->super_lists), dfree(ctx->super_counts);
    dfree(ctx->counters), dfree(ctx->mask), dfree(ctx->squeue), dfree(ctx->touch), dfree(ctx->wave_clock);
    dfree(ctx->tile_cost), dfree(ctx->tile_order), dfree(ctx->tile_cones);
    dfree(ctx->frames), dfree(ctx->tab), dfree(ctx->tables), dfree(ctx->stats);
    dfree(ctx->ex_tubes), dfree(ctx->ex_tris), dfree(ctx->ex_samples), dfree(ctx->ex_decided);
    dfree(ctx->ex_counts), dfree(ctx->ex_offsets), dfree(ctx->ex_first), dfree(ctx->ex_geoms);
    dfree(ctx->ex_attrs), dfree(ctx->ex_boxes);
    dfree(ctx->adj_target), dfree(ctx->adj_mask_in), dfree(ctx->adj_mask_out), dfree(ctx->adj_obj);
    dfree(ctx->adj_pix), dfree(ctx->adj_sig), dfree(ctx->adj_objid), dfree(ctx->adj_inside), dfree(ctx->adj_partials);
    dfree(ctx->adj_counts), dfree(ctx->adj_result);
    drop_graph(ctx);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side), (void)hipStreamDestroy(ctx->side);
    if (ctx->fork) (void)hipEventDestroy(ctx->fork);
    if (ctx->join) (void)hipEventDestroy(ctx->join);
    if (ctx->order_fork) (void)hipEventDestroy(ctx->order_fork);
    if (ctx->order_join) (void)hipEventDestroy(ctx->order_join);
    for (auto &st : ctx->stage) {
        if (st.done) (void)hipEventDestroy(st.done);
        if (st.host) (void)hipHostFree(st.host);
    }
    if (ctx->readback) (void)hipHostFree(ctx->readback);
    delete ctx;
}

const char *specrel_last_error(const specrel_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }
const char *specrel_device_name(const specrel_ctx *ctx) { return ctx ? ctx->name.c_str() : ""; }

int specrel_set_option(specrel_ctx *ctx, int option, int value) {
    if (!ctx) return SPECREL_EINVAL;
    if (option == SPECREL_OPT_STATS) ctx->opt_stats = value != 0;
    else if (option == SPECREL_OPT_WAVE_CLOCK) ctx->opt_wave_clock = value != 0;
    else if (option == SPECREL_OPT_STAR_CULL) ctx->opt_cull = value != 0;
    else if (option == SPECREL_OPT_STACK) {
        if (value != 8 && value != 10 && value != 12 && value != 16 && value != -4 && value != -8 && value != -12 &&
            value != -16) {
            ctx->err = "stack must be 8, 10, 12 or 16 (LDS positions) or -4, -8, -12, -16 (register entries)";
            return SPECREL_EINVAL;
        }
        ctx->opt_stack = value;
    } else if (option == SPECREL_OPT_PERSISTENT) ctx->opt_persistent = value != 0;
    else if (option == SPECREL_OPT_SHADOW_QUEUE) ctx->opt_shadow_queue = value != 0;
    else if (option == SPECREL_OPT_MERGED_LOADS) ctx->opt_merged = value != 0;
    else if (option == SPECREL_OPT_SHADOW_IN_TRACE) ctx->opt_shadow_trace = value != 0;
    else if (option == SPECREL_OPT_STAR_OVERLAP) ctx->opt_star_overlap = value != 0;
    else if (option == SPECREL_OPT_TRACE_WAVES) {
        if (value != 1 && value != 2 && value != 4) {
            ctx->err = "trace waves must be 1, 2 or 4";
            return SPECREL_EINVAL;
        }
        ctx->opt_trace_waves = value;
    } else if (option == SPECREL_OPT_TILE_ORDER) ctx->opt_tile_order = value != 0;
    else if (option == SPECREL_OPT_ADJOINT_MASK) {
        if (value != 0 && value != 1) {
            ctx->err = "adjoint mask must be 0 (silhouettes) or 1 (every shading edge)";
            return SPECREL_EINVAL;
        }
        ctx->opt_adj_strict = value;
    }
    else if (option == SPECREL_OPT_LEAF_POSTPONE) {
        if (value < 0 || value > STACK - 2) {
            ctx->err = "leaf postponement must be 0, 1 or a stack limit of 2..62 entries";
            return SPECREL_EINVAL;
        }
        ctx->opt_postpone = value;
    }
    else if (option == SPECREL_OPT_WAVE_SHAPE) {
        if (value < 0 || value > 2) {
            ctx->err = "wave shape must be 0 (8x8), 1 (16x4) or 2 (4x16)";
            return SPECREL_EINVAL;
        }
        ctx->opt_wave_shape = value;
    }
    else if (option == SPECREL_OPT_STAGE_EVENTS) {
        if (value < 0 || value > 2) {
            ctx->err = "stage events must be 0 (none), 1 (the primary-ray kernel) or 2 (every stage)";
            return SPECREL_EINVAL;
        }
        ctx->opt_stage_events = value;
    }
    else if (option == SPECREL_OPT_FUSED_SHADE) {
        if (value < 0 || value > 2) {
            ctx->err = "fused shading must be 0 (never), 1 (single views) or 2 (every render)";
            return SPECREL_EINVAL;
        }
        ctx->opt_fuse_shade = value;
    }
    else if (option == SPECREL_OPT_NODE_WIDTH) {
        if (value != 2 && value != 4) {
            ctx->err = "node width must be 2 (one box pair per step) or 4 (the children's pairs fetched too)";
            return SPECREL_EINVAL;
        }
        ctx->opt_node_width = value;
    }
    else if (option == SPECREL_OPT_LDS_TOP) {
        if (value < 0 || value > TOP_LEVELS_MAX) {
            ctx->err = "LDS top levels must be 0..8";
            return SPECREL_EINVAL;
        }
        ctx->opt_lds_top = value;
    }
    else if (option == SPECREL_OPT_GRAPH) {
        if (value < 0 || value > 2) {
            ctx->err = "graph must be 0 (never), 1 (batches) or 2 (every render)";
            return SPECREL_EINVAL;
        }
        ctx->opt_graph = value;
    } else {
        ctx->err = "unknown option";
        return SPECREL_EINVAL;
    }
    return SPECREL_OK;
}

int specrel_upload_scene(specrel_ctx *ctx, uint32_t scene_idx, const specrel_scene_arrays *scene) {
    if (!ctx) return SPECREL_EINVAL;
    return guarded(ctx, [&] {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        if (scene_idx > 4096) throw Fail{SPECREL_EINVAL, "scene index too large"};
        if (ctx->scenes.size() <= scene_idx) ctx->scenes.resize(scene_idx + 1);
        // Keep the old device buffers for reuse; rebuild the slot's contents.
        SceneSlot &slot = ctx->scenes[scene_idx];
        slot.used = false;
        slot.versioned = false;
        ctx->generation++;  // a captured graph holds the old scene pointers
        build_device_scene(ctx, slot, scene);
        return SPECREL_OK;
    });
}

int specrel_upload_scene_versioned(specrel_ctx *ctx, uint32_t scene_idx, const specrel_scene_arrays *scene,
                                   const specrel_bvh_version *version) {
    if (!ctx) return SPECREL_EINVAL;
    return guarded(ctx, [&] {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        if (!version) throw Fail{SPECREL_EINVAL, "version is NULL"};
        if (scene_idx > 4096) throw Fail{SPECREL_EINVAL, "scene index too large"};
        if (ctx->scenes.size() <= scene_idx) ctx->scenes.resize(scene_idx + 1);
        SceneSlot &slot = ctx->scenes[scene_idx];
        slot.used = false;
        slot.versioned = false;
        ctx->generation++;
        build_device_scene(ctx, slot, scene, version->edit > 0);
        slot.versioned = true;
        slot.layout = version->layout;
        slot.edit = version->edit;
        return SPECREL_OK;
    });
}

int specrel_update_scene(specrel_ctx *ctx, uint32_t scene_idx, const specrel_scene_arrays *scene,
                         const specrel_bvh_version *version, const specrel_range *node_ranges, uint32_t n_node_ranges,
                         const specrel_range *prim_ranges, uint32_t n_prim_ranges, int *out_patched) {
    if (!ctx) return SPECREL_EINVAL;
    return guarded(ctx, [&] {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        if (!scene || !version) throw Fail{SPECREL_EINVAL, "scene or version is NULL"};
        if (scene_idx >= ctx->scenes.size() || !ctx->scenes[scene_idx].used)
            throw Fail{SPECREL_EINVAL, "no scene uploaded at index " + std::to_string(scene_idx)};
        SceneSlot &slot = ctx->scenes[scene_idx];
        ctx->generation++;  // a captured graph holds the scene by value
        // follows (gpu.rs:341-343): the device holds the version just before this edit
        const bool follows = slot.versioned && slot.layout == version->layout && slot.edit + 1 == version->edit;
        bool patched = follows && update_device_scene(ctx, slot, scene, node_ranges, n_node_ranges, prim_ranges,
                                                       n_prim_ranges);
        if (!patched) {
            slot.used = false;
            slot.versioned = false;
            build_device_scene(ctx, slot, scene, version->edit > 0);
        }
        slot.versioned = true;
        slot.layout = version->layout;
        slot.edit = version->edit;
        if (out_patched) *out_patched = patched ? 1 : 0;
        return SPECREL_OK;
    });
}

int specrel_upload_scene_file(specrel_ctx *ctx, uint32_t scene_idx, const char *path, int verify) {
    if (!ctx) return SPECREL_EINVAL;
    return guarded(ctx, [&] {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        if (scene_idx > 4096) throw Fail{SPECREL_EINVAL, "scene index too large"};
        sr::file::Mapped m;
        specrel_scene_arrays a;
        try {
            sr::file::open(path, verify != 0, m, a);
        } catch (const sr::file::Error &e) {
            throw Fail{SPECREL_EINVAL, e.msg};
        }
        if (ctx->scenes.size() <= scene_idx) ctx->scenes.resize(scene_idx + 1);
        SceneSlot &slot = ctx->scenes[scene_idx];
        slot.used = false;
        slot.versioned = false;
        ctx->generation++;
        // copies out of the mapping before it is unmapped; a file of an edited BVH gets room to grow
        build_device_scene(ctx, slot, &a, m.h.bvh_layout != 0 && m.h.bvh_edit > 0);
        // the file's BVH version (layout 0: none; the reference numbers layouts from 1, bvh4d.rs:57-61)
        slot.versioned = m.h.bvh_layout != 0;
        slot.layout = m.h.bvh_layout;
        slot.edit = m.h.bvh_edit;
        return SPECREL_OK;
    });
}

int specrel_extrude_tubes(specrel_ctx *ctx, const specrel_tube *tubes, uint32_t n_tubes, specrel_geom *geoms,
                          specrel_attr *attrs, float *boxes, uint64_t cap, uint64_t *n_out, uint64_t *counts) {
    if (!ctx) return SPECREL_EINVAL;
    return guarded(ctx, [&] {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        if ((n_tubes && !tubes) || !n_out) throw Fail{SPECREL_EINVAL, "tubes or n_out is NULL"};
        ctx->ex_records = 0;  // until this extrusion has completed
        // Host tables: the tubes' prepared triangles (TubeTriangle::new), samples, decided_at.
        std::vector<dev::ExTube> desc(n_tubes);
        std::vector<sr::TubeTri> tris;
        std::vector<sr::Sample> samples;
        std::vector<float> decided;
        uint64_t n_prisms = 0;
        for (uint32_t i = 0; i < n_tubes; ++i) {
            const specrel_tube &t = tubes[i];
            const std::string who = "tube " + std::to_string(i) + ": ";
            dev::ExTube &d = desc[i];
            d.prism_first = n_prisms;
            d.tri_first = (uint32_t)tris.size();
            d.n_tris = t.n_triangles;
            d.sample_first = (uint32_t)samples.size();
            d.n_samples = t.n_samples;
            d.decided_first = ~0u;
            d.material = t.material;
            if (t.n_samples < 2 || t.n_triangles == 0) {  // nothing to extrude (mesh4d.rs:728)
                d.n_tris = std::max(1u, t.n_triangles);
                continue;
            }
            if (!t.vertices || !t.tri_index || !t.tri_uv || !t.tri_normal || !t.samples)
                throw Fail{SPECREL_EINVAL, who + "arrays are NULL"};
            std::vector<sr::V3> verts(t.n_vertices);
            for (uint32_t k = 0; k < t.n_vertices; ++k)
                verts[k] = sr::v3(t.vertices[3 * k], t.vertices[3 * k + 1], t.vertices[3 * k + 2]);
            for (uint32_t k = 0; k < t.n_samples; ++k) {
                const specrel_wl_sample &w = t.samples[k];
                sr::Sample sm = sr::sample_from(w.center, w.proper_time, w.phi, w.proper_accel_local, w.e);
                for (const sr::V3 &p : verts)  // extend_worldtube_simplices' horizon check (mesh4d.rs:738-747)
                    if (!(1.0f + sr::dot(sm.accel_local, p) > 0.0f))
                        throw Fail{SPECREL_EINVAL, who + "Born-rigid worldtube crosses its Rindler horizon at sample " +
                                                       std::to_string(k)};
                samples.push_back(sm);
            }
            for (uint32_t k = 0; k < t.n_triangles; ++k) {
                uint32_t idx[3];
                sr::V3 pos[3];
                float uv[3][2];
                for (int c = 0; c < 3; ++c) {
                    idx[c] = t.tri_index[3 * k + c];
                    if (idx[c] >= t.n_vertices) throw Fail{SPECREL_EINVAL, who + "triangle vertex index out of range"};
                    pos[c] = verts[idx[c]];
                    uv[c][0] = t.tri_uv[6 * k + 2 * c];
                    uv[c][1] = t.tri_uv[6 * k + 2 * c + 1];
                }
                tris.push_back(sr::tube_tri(
                    idx, pos, uv, sr::v3(t.tri_normal[3 * k], t.tri_normal[3 * k + 1], t.tri_normal[3 * k + 2])));
            }
            if (t.decided_at) {
                d.decided_first = (uint32_t)decided.size();
                decided.insert(decided.end(), t.decided_at, t.decided_at + t.n_samples);
            }
            n_prisms += (uint64_t)(t.n_samples - 1) * t.n_triangles;
        }
        const uint64_t n_blocks = (n_prisms + 255) / 256;
        if (n_blocks > 0x7FFFFFFFull) throw Fail{SPECREL_EINVAL, "too many prisms for one launch"};
        hipStream_t s = ctx->stream;
        upload(ctx->ex_tubes, desc.data(), desc.size() * sizeof(dev::ExTube), s);
        upload(ctx->ex_tris, tris.data(), tris.size() * sizeof(sr::TubeTri), s);
        upload(ctx->ex_samples, samples.data(), samples.size() * sizeof(sr::Sample), s);
        upload(ctx->ex_decided, decided.data(), decided.size() * sizeof(float), s);
        dalloc(ctx->ex_counts, std::max<uint64_t>(1, n_blocks) * 4);
        dalloc(ctx->ex_offsets, (n_blocks + 1) * 8);
        dalloc(ctx->ex_first, std::max<uint32_t>(1, n_tubes) * 8);
        dev::ExArgs a;
        a.tubes = (const dev::ExTube *)ctx->ex_tubes.p;
        a.n_tubes = n_tubes;
        a.n_prisms = n_prisms;
        a.tris = (const sr::TubeTri *)ctx->ex_tris.p;
        a.samples = (const sr::Sample *)ctx->ex_samples.p;
        a.decided = (const float *)ctx->ex_decided.p;
        uint64_t total = 0;
        float ms = 0.0f;
        bool direct = false;
        if (n_prisms) {
            // Optimistic pass: records at slot 3p + k; when all are proper these are final.
            const uint64_t slots = 3 * n_prisms, slot_blocks = (slots + 255) / 256;
            if (slot_blocks > 0x7FFFFFFFull) throw Fail{SPECREL_EINVAL, "too many prisms for one launch"};
            dalloc(ctx->ex_geoms, slots * sizeof(specrel_geom));
            dalloc(ctx->ex_attrs, slots * sizeof(specrel_attr));
            dalloc(ctx->ex_boxes, slots * 32);
            dalloc(ctx->ex_counts, slot_blocks * 4);
            dalloc(ctx->ex_offsets, (slot_blocks + 1) * 8);
            check(hipEventRecord(ctx->ev[7], s), "event");
            dev::k_extrude_direct<<<dim3((unsigned)slot_blocks), 256, 0, s>>>(
                a, (uint32_t *)ctx->ex_counts.p, (specrel_geom *)ctx->ex_geoms.p, (specrel_attr *)ctx->ex_attrs.p,
                (float4 *)ctx->ex_boxes.p);
            check(hipGetLastError(), "k_extrude_direct launch");
            dev::k_extrude_scan<<<1, 1024, 0, s>>>((uint32_t *)ctx->ex_counts.p, (uint32_t)slot_blocks,
                                                   (uint64_t *)ctx->ex_offsets.p);
            check(hipGetLastError(), "k_extrude_scan launch");
            check(hipEventRecord(ctx->ev[8], s), "event");
            check(hipMemcpyAsync(&total, (const uint64_t *)ctx->ex_offsets.p + slot_blocks, 8, hipMemcpyDeviceToHost,
                                 s),
                  "total readback");
            check(hipStreamSynchronize(s), "extrude");
            ms = ev_ms(ctx->ev[7], ctx->ev[8]);
            direct = total == slots;
            if (!direct) {
                // Some tetrahedra were degenerate: count per block of prisms, scan, and write each
                // record again at its index after the drops.
                check(hipEventRecord(ctx->ev[7], s), "event");
                dev::k_extrude_count<<<dim3((unsigned)n_blocks), 256, 0, s>>>(a, (uint32_t *)ctx->ex_counts.p);
                check(hipGetLastError(), "k_extrude_count launch");
                dev::k_extrude_scan<<<1, 1024, 0, s>>>((uint32_t *)ctx->ex_counts.p, (uint32_t)n_blocks,
                                                       (uint64_t *)ctx->ex_offsets.p);
                check(hipGetLastError(), "k_extrude_scan launch");
                dev::k_extrude_emit<<<dim3((unsigned)n_blocks), 256, 0, s>>>(
                    a, (const uint64_t *)ctx->ex_offsets.p, (specrel_geom *)ctx->ex_geoms.p,
                    (specrel_attr *)ctx->ex_attrs.p, (float4 *)ctx->ex_boxes.p, (uint64_t *)ctx->ex_first.p);
                check(hipGetLastError(), "k_extrude_emit launch");
                check(hipEventRecord(ctx->ev[8], s), "event");
                check(hipStreamSynchronize(s), "extrude emit");
                ms += ev_ms(ctx->ev[7], ctx->ev[8]);
            }
        }
        ctx->timings.extrude_ms = ms;
        ctx->ex_records = total;
        *n_out = total;
        std::vector<uint64_t> first(n_tubes, 0);
        if (direct)
            for (uint32_t i = 0; i < n_tubes; ++i) first[i] = 3 * desc[i].prism_first;
        else if (n_tubes && n_prisms)
            check(hipMemcpy(first.data(), ctx->ex_first.p, n_tubes * 8, hipMemcpyDeviceToHost), "tube offsets");
        if (counts) {
            // Records of tube i: from its first record to the next tube with prisms (or the end).
            uint64_t next = total;
            for (uint32_t i = n_tubes; i-- > 0;) {
                const bool has = tubes[i].n_samples >= 2 && tubes[i].n_triangles > 0;
                counts[i] = has ? next - first[i] : 0;
                if (has) next = first[i];
            }
        }
        if (geoms || attrs || boxes) {
            if (total > cap)
                throw Fail{SPECREL_EINVAL, "output buffers hold " + std::to_string(cap) + " records, " +
                                               std::to_string(total) + " needed"};
            if (geoms && total)
                check(hipMemcpy(geoms, ctx->ex_geoms.p, total * sizeof(specrel_geom), hipMemcpyDeviceToHost), "geoms");
            if (attrs && total)
                check(hipMemcpy(attrs, ctx->ex_attrs.p, total * sizeof(specrel_attr), hipMemcpyDeviceToHost), "attrs");
            if (boxes && total) check(hipMemcpy(boxes, ctx->ex_boxes.p, total * 32, hipMemcpyDeviceToHost), "boxes");
        }
        return SPECREL_OK;
    });
}

int specrel_build_bvh(specrel_ctx *ctx, const float *boxes, uint64_t n, specrel_node *nodes, uint32_t *order,
                      uint32_t *n_nodes, specrel_bvh_stats *stats) {
    if (!ctx) return SPECREL_EINVAL;
    return guarded(ctx, [&] {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        if (!nodes || !order || !n_nodes) throw Fail{SPECREL_EINVAL, "nodes, order or n_nodes is NULL"};
        if (n > 0x7FFFFFFFull) throw Fail{SPECREL_EINVAL, "too many records"};
        hipStream_t s = ctx->stream;
        const float4 *d_boxes;
        DevBuf in;
        if (boxes) {
            dalloc(in, std::max<uint64_t>(1, n) * 32);
            check(hipMemcpyAsync(in.p, boxes, n * 32, hipMemcpyHostToDevice, s), "boxes upload");
            d_boxes = (const float4 *)in.p;
        } else {
            if (n != ctx->ex_records)
                throw Fail{SPECREL_EINVAL, "n is " + std::to_string(n) + " but the last extrusion left " +
                                               std::to_string(ctx->ex_records) + " records on the device"};
            d_boxes = (const float4 *)ctx->ex_boxes.p;
        }
        DevBuf d_nodes, d_order;
        dalloc(d_nodes, std::max<uint64_t>(1, 2 * n) * sizeof(specrel_node));
        dalloc(d_order, std::max<uint64_t>(1, n) * 4);
        dev::BvhBuildStats st{};
        uint32_t total = 0;
        std::string err = dev::build_bvh(d_boxes, (uint32_t)n, (specrel_node *)d_nodes.p, (uint32_t *)d_order.p,
                                         &total, s, &st);
        if (!err.empty()) throw Fail{SPECREL_EDEVICE, err};
        if (total) check(hipMemcpy(nodes, d_nodes.p, (size_t)total * sizeof(specrel_node), hipMemcpyDeviceToHost),
                         "nodes readback");
        if (n) check(hipMemcpy(order, d_order.p, n * 4, hipMemcpyDeviceToHost), "order readback");
        dfree(in), dfree(d_nodes), dfree(d_order);
        *n_nodes = total;
        if (stats) {
            std::memset(stats, 0, sizeof *stats);
            stats->ms = st.ms;
            stats->levels = st.levels;
            stats->large_segments = st.large_segments;
            stats->small_subtrees = st.small_subtrees;
            stats->fallback_segments = st.fallback_segments;
            stats->jump_passes = st.jump_passes;
        }
        return SPECREL_OK;
    });
}

int specrel_patch_scene(specrel_ctx *ctx, uint32_t scene_idx, const specrel_scene_arrays *scene,
                        const specrel_range *node_ranges, uint32_t n_node_ranges, const specrel_range *prim_ranges,
                        uint32_t n_prim_ranges, int *out_full) {
    if (!ctx) return SPECREL_EINVAL;
    return guarded(ctx, [&] {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        if (!scene) throw Fail{SPECREL_EINVAL, "scene is NULL"};
        if (scene_idx >= ctx->scenes.size() || !ctx->scenes[scene_idx].used)
            throw Fail{SPECREL_EINVAL, "no scene uploaded at index " + std::to_string(scene_idx)};
        SceneSlot &slot = ctx->scenes[scene_idx];
        ctx->generation++;  // a captured graph holds the scene by value
        bool ok = patch_device_scene(ctx, slot, scene, node_ranges, n_node_ranges, prim_ranges, n_prim_ranges);
        slot.versioned = false;  // an unversioned edit: the next versioned update uploads afresh
        if (!ok) {
            slot.used = false;
            build_device_scene(ctx, slot, scene);
        }
        if (out_full) *out_full = ok ? 0 : 1;
        return SPECREL_OK;
    });
}

uint32_t specrel_num_scenes(const specrel_ctx *ctx) { return ctx ? (uint32_t)ctx->scenes.size() : 0; }

int specrel_render(specrel_ctx *ctx, uint32_t scene_idx, const specrel_frame *frame, uint32_t *out_rgb,
                   uint32_t *out_hit, float *out_lambda, float *out_tau, uint64_t len) {
    if (!ctx) return SPECREL_EINVAL;
    return guarded(ctx, [&] {
        if (!out_rgb) throw Fail{SPECREL_EINVAL, "out_rgb is NULL"};
        if (!frame) throw Fail{SPECREL_EINVAL, "frame is NULL"};
        // The framebuffer must hold width * height pixels (gpu.rs:808), checked before any work.
        check_len(len, (uint64_t)frame->size[0] * frame->size[1], "framebuffer", 0);
        check(hipSetDevice(ctx->device), "hipSetDevice");
        bool hits = out_hit || out_lambda || out_tau;
        enqueue_views(ctx, scene_idx, frame, 1, hits);
        std::vector<Fetch> f;
        view_fetches(ctx, 0, out_rgb, out_hit, out_lambda, out_tau, len, f);
        float wait_ms = 0.0f, copy_ms = 0.0f;
        fetch(ctx, f, &wait_ms, &copy_ms);
        collect(ctx);
        ctx->timings.wait_ms = wait_ms;
        ctx->timings.readback_ms = copy_ms;
        return SPECREL_OK;
    });
}

int specrel_render_uniforms(specrel_ctx *ctx, uint32_t scene_idx, const specrel_gpu_uniforms *u, uint32_t *out_rgb,
                            uint64_t len) {
    if (!ctx) return SPECREL_EINVAL;
    specrel_frame f;
    const int rc = guarded(ctx, [&] {
        if (!u) throw Fail{SPECREL_EINVAL, "uniforms is NULL"};
        if (scene_idx >= ctx->scenes.size() || !ctx->scenes[scene_idx].used)
            throw Fail{SPECREL_EINVAL, "no scene uploaded at index " + std::to_string(scene_idx)};
        const SceneSlot &sc = ctx->scenes[scene_idx];
        const float4 l = sc.S.light;
        const float light[4] = {l.x, l.y, l.z, l.w};
        if (u->num_nodes != sc.num_nodes || u->num_stars != sc.num_stars ||
            std::memcmp(light, u->light_pos, sizeof light) != 0)
            throw Fail{SPECREL_EINVAL, "uniforms were built for another scene (num_nodes, num_stars or light_pos "
                                       "differ from scene " + std::to_string(scene_idx) + "'s)"};
        if (u->num_flashes > SPECREL_MAX_FLASHES) throw Fail{SPECREL_EINVAL, "num_flashes out of range"};
        std::memset(&f, 0, sizeof f);
        std::memcpy(f.cam_pos, u->cam_pos, sizeof f.cam_pos);
        std::memcpy(f.e, u->e, sizeof f.e);
        std::memcpy(f.half_wh, u->half_wh, sizeof f.half_wh);
        std::memcpy(f.size, u->size, sizeof f.size);
        f.mode = u->mode;
        f.shadow_mode = u->shadow_mode;
        f.slice_mode = u->slice_mode;
        f.num_flashes = u->num_flashes;
        std::memcpy(f.frontier_origin, u->frontier_origin, sizeof f.frontier_origin);
        f.frontier_tau = u->frontier_tau;
        f.frontier_view_offset = u->frontier_view_offset;
        f.prophecy_delta = u->prophecy_delta;
        f.num_haze = u->num_haze;
        std::memcpy(f.haze_origin, u->haze_origin, sizeof f.haze_origin);
        std::memcpy(f.flash_at, u->flash_at, sizeof f.flash_at);
        std::memcpy(f.flash_color_width, u->flash_col_width, sizeof f.flash_color_width);
        for (uint32_t i = 0; i < SPECREL_MAX_FLASHES; ++i) f.flash_intensity[i] = i < u->num_flashes ? 1.0f : 0.0f;
        return SPECREL_OK;
    });
    if (rc != SPECREL_OK) return rc;
    return specrel_render(ctx, scene_idx, &f, out_rgb, nullptr, nullptr, nullptr, len);
}

int specrel_render_resident(specrel_ctx *ctx, uint32_t scene_idx, const specrel_frame *frame, int want_hits) {
    if (!ctx) return SPECREL_EINVAL;
    return guarded(ctx, [&] {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        return enqueue_views(ctx, scene_idx, frame, 1, want_hits != 0);
    });
}

int specrel_render_noalloc(specrel_ctx *ctx, uint32_t scene_idx, const specrel_frame *frame, uint32_t *d_rgb,
                           uint32_t *d_hit, float *d_lambda, float *d_tau, uint64_t len) {
    if (!ctx) return SPECREL_EINVAL;
    return guarded(ctx, [&] {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        enqueue_noalloc(ctx, scene_idx, frame, d_rgb, d_hit, d_lambda, d_tau, len);
        return SPECREL_OK;
    });
}

int specrel_render_batch(specrel_ctx *ctx, uint32_t scene_idx, const specrel_frame *frames, uint32_t n_views,
                         const specrel_view_out *outs) {
    if (!ctx) return SPECREL_EINVAL;
    return guarded(ctx, [&] {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        bool hits = false;
        if (outs && frames)
            for (uint32_t i = 0; i < n_views; ++i) {
                const specrel_view_out &o = outs[i];
                hits = hits || o.hit || o.lambda || o.tau;
                if (o.rgb || o.hit || o.lambda || o.tau)
                    check_len(o.len, (uint64_t)frames[i].size[0] * frames[i].size[1], "view output", i);
            }
        enqueue_views(ctx, scene_idx, frames, n_views, hits);
        std::vector<Fetch> f;
        if (outs)
            for (uint32_t i = 0; i < n_views; ++i)
                view_fetches(ctx, i, outs[i].rgb, outs[i].hit, outs[i].lambda, outs[i].tau, outs[i].len, f);
        float wait_ms = 0.0f, copy_ms = 0.0f;
        fetch(ctx, f, &wait_ms, &copy_ms);
        collect(ctx);
        ctx->timings.wait_ms = wait_ms;
        ctx->timings.readback_ms = copy_ms;
        return SPECREL_OK;
    });
}

int specrel_render_batch_resident(specrel_ctx *ctx, uint32_t scene_idx, const specrel_frame *frames,
                                  uint32_t n_views, int want_hits) {
    if (!ctx) return SPECREL_EINVAL;
    return guarded(ctx, [&] {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        return enqueue_views(ctx, scene_idx, frames, n_views, want_hits != 0);
    });
}

int specrel_sync(specrel_ctx *ctx) {
    if (!ctx) return SPECREL_EINVAL;
    return guarded(ctx, [&] {
        auto t0 = std::chrono::steady_clock::now();
        check(hipStreamSynchronize(ctx->stream), "sync");
        auto t1 = std::chrono::steady_clock::now();
        collect(ctx);
        ctx->timings.wait_ms = std::chrono::duration<float, std::milli>(t1 - t0).count();
        return SPECREL_OK;
    });
}

int specrel_download(specrel_ctx *ctx, uint32_t *out_rgb, uint32_t *out_hit, float *out_lambda, float *out_tau,
                     uint64_t len) {
    return specrel_download_view(ctx, 0, out_rgb, out_hit, out_lambda, out_tau, len);
}

int specrel_download_view(specrel_ctx *ctx, uint32_t view, uint32_t *out_rgb, uint32_t *out_hit, float *out_lambda,
                          float *out_tau, uint64_t len) {
    if (!ctx) return SPECREL_EINVAL;
    return guarded(ctx, [&] {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        download_view(ctx, view, out_rgb, out_hit, out_lambda, out_tau, len);
        return SPECREL_OK;
    });
}

int specrel_download_linear(specrel_ctx *ctx, uint32_t view, float *out_rgb3, uint64_t len) {
    if (!ctx) return SPECREL_EINVAL;
    return guarded(ctx, [&] {
        if (!out_rgb3) throw Fail{SPECREL_EINVAL, "output is NULL"};
        if (ctx->n_views == 0) throw Fail{SPECREL_EINVAL, "nothing rendered yet"};
        if (ctx->last_noalloc) throw Fail{SPECREL_EINVAL, NOALLOC_LAST};
        if (view >= ctx->n_views) throw Fail{SPECREL_EINVAL, "view index out of range"};
        const size_t n = (size_t)ctx->sizes[2 * view] * ctx->sizes[2 * view + 1];
        check_len(len, 3 * (uint64_t)n, "linear image", view);
        check(hipSetDevice(ctx->device), "hipSetDevice");
        std::vector<float4> tmp(n);
        fetch(ctx, {{tmp.data(), (const float4 *)ctx->hdr.p + host_start(ctx, ROW_PYR, view), n * sizeof(float4)}},
              nullptr, nullptr);
        for (size_t i = 0; i < n; ++i) {
            out_rgb3[3 * i] = tmp[i].x;
            out_rgb3[3 * i + 1] = tmp[i].y;
            out_rgb3[3 * i + 2] = tmp[i].z;
        }
        return SPECREL_OK;
    });
}

int specrel_autotune(specrel_ctx *ctx, uint32_t scene_idx, const specrel_frame *frames, uint32_t n_views,
                     uint32_t iters, specrel_tuning *out) {
    if (!ctx) return SPECREL_EINVAL;
    return guarded(ctx, [&] {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        if (ctx->opt_stats) throw Fail{SPECREL_EINVAL, "autotune runs with SPECREL_OPT_STATS off"};
        if (iters == 0 || iters > 1000) throw Fail{SPECREL_EINVAL, "iters must be 1..1000"};
        struct Cand {
            int stack, persistent, queue, graph, in_trace, shape, postpone, waves, width, top, fuse;
        };
        const Cand saved{ctx->opt_stack,        ctx->opt_persistent, ctx->opt_shadow_queue, ctx->opt_graph,
                         ctx->opt_shadow_trace, ctx->opt_wave_shape, ctx->opt_postpone,   ctx->opt_trace_waves,
                         ctx->opt_node_width,   ctx->opt_lds_top,    ctx->opt_fuse_shade};
        std::vector<Cand> cands;
        for (int st : {16, 8})
            for (int pe : {0, 1})
                for (int sh : {0, 1, 2})  // shadows in the shading pass, queued, in the trace kernel
                    for (int g : {0, 2})
                        for (int ws : {0, 1, 2})  // a wave's pixels: 8x8, 16x4, 4x16 (tiled primary rays only)
                            if (ws == 0 || !pe)
                                cands.push_back({st, pe, sh == 1, g, sh == 2, ws, saved.postpone, saved.waves,
                                                 saved.width, saved.top, saved.fuse});
        auto apply = [&](const Cand &c) {
            ctx->opt_stack = c.stack;
            ctx->opt_persistent = c.persistent;
            ctx->opt_shadow_queue = c.queue;
            ctx->opt_graph = c.graph;
            ctx->opt_shadow_trace = c.in_trace;
            ctx->opt_wave_shape = c.shape;
            ctx->opt_postpone = c.postpone;
            ctx->opt_trace_waves = c.waves;
            ctx->opt_node_width = c.width;
            ctx->opt_lds_top = c.top;
            ctx->opt_fuse_shade = c.fuse;
        };
        // Median wall time per render of the batch, enqueue to completion.
        auto time_cand = [&](const Cand &c) {
            apply(c);
            enqueue_views(ctx, scene_idx, frames, n_views, false);
            check(hipStreamSynchronize(ctx->stream), "autotune warmup");
            std::vector<float> ms;
            for (uint32_t i = 0; i < iters; ++i) {
                auto t0 = std::chrono::steady_clock::now();
                enqueue_views(ctx, scene_idx, frames, n_views, false);
                check(hipStreamSynchronize(ctx->stream), "autotune");
                ms.push_back(std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count());
            }
            std::sort(ms.begin(), ms.end());
            return ms[ms.size() / 2];
        };
        float base_ms = 0.0f, best_ms = 0.0f;
        Cand best = saved;
        try {
            base_ms = time_cand(saved);
            best_ms = base_ms;
            for (const Cand &c : cands) {
                float t = time_cand(c);
                if (t < best_ms) {
                    best_ms = t;
                    best = c;
                }
            }
            // Then the best one with leaf postponement switched.
            Cand flip = best;
            flip.postpone = best.postpone ? 0 : 1;
            cands.push_back(flip);
            float t = time_cand(flip);
            if (t < best_ms) {
                best_ms = t;
                best = flip;
            }
            // Then the best one with the other block sizes of tiled LDS-stack primary rays.
            if (best.stack > 0 && !best.persistent) {
                const Cand base = best;
                for (int w : {1, 2, 4}) {
                    if (w == base.waves) continue;
                    Cand c = base;
                    c.waves = w;
                    cands.push_back(c);
                    float tw = time_cand(c);
                    if (tw < best_ms) {
                        best_ms = tw;
                        best = c;
                    }
                }
            }
            // Then the best one with the other BVH node fetch width, and with 4, 6 or 8 top levels
            // of the tree in LDS (tiled primary rays with an LDS stack of 8 or 16).
            if ((best.stack == 8 || best.stack == 16) && !best.persistent) {
                const Cand base = best;
                std::vector<Cand> more;
                Cand c = base;
                c.width = base.width == 4 ? 2 : 4;
                c.top = 0;
                more.push_back(c);
                for (int lv : {0, 4, 6, 8}) {
                    if (lv == base.top) continue;
                    c = base;
                    c.top = lv;
                    more.push_back(c);
                }
                for (const Cand &m : more) {
                    cands.push_back(m);
                    float tm = time_cand(m);
                    if (tm < best_ms) {
                        best_ms = tm;
                        best = m;
                    }
                }
                // Then the best one with the shading fused into the primary rays for every render,
                // or never.
                Cand c2 = best;
                c2.fuse = n_views > 1 ? 2 : 0;
                if (c2.fuse != best.fuse) {
                    cands.push_back(c2);
                    float tf = time_cand(c2);
                    if (tf < best_ms) {
                        best_ms = tf;
                        best = c2;
                    }
                }
            }
        } catch (...) {
            apply(saved);
            throw;
        }
        apply(best);
        if (out) {
            std::memset(out, 0, sizeof *out);
            out->stack = best.stack;
            out->persistent = best.persistent;
            out->shadow_queue = best.queue;
            out->graph = best.graph;
            out->shadow_in_trace = best.in_trace;
            out->wave_shape = best.shape;
            out->leaf_postpone = best.postpone != 0;
            out->trace_waves = (best.stack > 0 && !best.persistent) ? best.waves : 4;
            out->node_width = best.width;
            out->lds_top = best.top;
            out->fused_shade = best.fuse;
            out->candidates = (uint32_t)cands.size() + 1;
            out->best_ms = best_ms;
            out->default_ms = base_ms;
        }
        return SPECREL_OK;
    });
}

int specrel_adjoint(specrel_ctx *ctx, const float *target, const uint8_t *mask_in, uint8_t *mask_out,
                    uint64_t n_pixels, const specrel_adj_object *objects, uint32_t n_objects,
                    specrel_adjoint_out *out, double *d_objects) {
    if (!ctx) return SPECREL_EINVAL;
    return guarded(ctx, [&] {
        if (ctx->n_views != 1) throw Fail{SPECREL_EINVAL, "the last render is not a single view"};
        return adjoint_views(ctx, target, mask_in, mask_out, n_pixels, objects, n_objects, out, d_objects);
    });
}

int specrel_adjoint_batch(specrel_ctx *ctx, const float *targets, const uint8_t *masks_in, uint8_t *masks_out,
                          uint64_t n_pixels, const specrel_adj_object *objects, uint32_t n_objects,
                          specrel_adjoint_out *outs, double *d_objects) {
    if (!ctx) return SPECREL_EINVAL;
    return guarded(ctx, [&] {
        return adjoint_views(ctx, targets, masks_in, masks_out, n_pixels, objects, n_objects, outs, d_objects);
    });
}

int specrel_last_timings(const specrel_ctx *ctx, specrel_timings *out) {
    if (!ctx || !out) return SPECREL_EINVAL;
    *out = ctx->timings;
    return SPECREL_OK;
}

int specrel_last_stats(const specrel_ctx *ctx, specrel_stats *out) {
    if (!ctx || !out) return SPECREL_EINVAL;
    *out = ctx->stats_host;
    return SPECREL_OK;
}

int specrel_last_wave_clock(specrel_ctx *ctx, specrel_wave_clock *out, uint64_t cap, uint64_t *n_waves,
                            uint32_t *clock_khz) {
    if (!ctx || !n_waves || !clock_khz || (cap && !out)) return SPECREL_EINVAL;
    return guarded(ctx, [&] {
        check(hipSetDevice(ctx->device), "hipSetDevice");
        check(hipStreamSynchronize(ctx->stream), "wave clock");
        const uint64_t n = ctx->last_noalloc ? 0ull : ctx->wave_clock_n;
        *n_waves = n;
        *clock_khz = (uint32_t)ctx->clock_khz;
        const uint64_t k = std::min(cap, n);
        if (k) check(hipMemcpy(out, ctx->wave_clock.p, k * sizeof(specrel_wave_clock), hipMemcpyDeviceToHost),
                     "wave clock copy");
        return SPECREL_OK;
    });
}

}  // extern "C"
