// The split decision of Bvh4D::build (bvh4d.rs:254-427), shared by the host builder (scene.cpp)
// and the device builder (bvh_build.hip), so both choose the same split bit for bit: 8 centroid
// bins per live axis, the spatial + 0.22 x temporal surface measure, the first lowest SAH cost
// over axes t, x, y, z and bins right to left, and the fallbacks (midpoint of the widest centroid
// extent; a leaf when every extent is below 1e-5). Bounds and bins are unions (min / max) and
// counts, so the order references are folded in does not change them.
#pragma once
#include <stdint.h>

#include "extrude.h"
#include "minkowski.h"

namespace sr {

const int BVH_BINS = 8;

SR_HD Box4 box_empty() {
    const float inf = __builtin_inff();
    return Box4{splat4(inf), splat4(-inf)};
}
SR_HD Box4 box_union(Box4 a, Box4 b) { return Box4{vmin(a.mn, b.mn), vmax(a.mx, b.mx)}; }
SR_HD void box_add_point(Box4 &a, V4 p) {
    a.mn = vmin(a.mn, p);
    a.mx = vmax(a.mx, p);
}

SR_HD float box_measure(const Box4 &b) {
    V4 e = b.mx - b.mn;
    float spatial_sa = e.x * e.y + e.y * e.z + e.z * e.x;
    float temporal_sa = 0.22f * e.t * (e.x + e.y + e.z);
    return spatial_sa + temporal_sa;
}

// Per-axis binning frame of a node: which axes have extent, their minimum and bins / extent.
struct BinFrame {
    bool live[4];
    float axis_min[4], inv_ext[4], ext[4];
};

SR_HD BinFrame bin_frame(const Box4 &cb) {
    BinFrame f;
    V4 e = cb.mx - cb.mn;
    for (int a = 0; a < 4; ++a) {
        f.ext[a] = axis(e, a);
        f.live[a] = f.ext[a] >= 1e-4f;
        f.axis_min[a] = axis(cb.mn, a);
        f.inv_ext[a] = (float)BVH_BINS / f.ext[a];
    }
    return f;
}

SR_HD uint32_t bin_of(const BinFrame &f, int a, float c) {
    uint64_t k = f2u_sat((c - f.axis_min[a]) * f.inv_ext[a]);
    return k > (uint64_t)(BVH_BINS - 1) ? (uint32_t)(BVH_BINS - 1) : (uint32_t)k;
}

struct Bins {
    Box4 b[4][BVH_BINS];
    uint32_t n[4][BVH_BINS];
};

SR_HD void bins_clear(Bins &B) {
    for (int a = 0; a < 4; ++a)
        for (int i = 0; i < BVH_BINS; ++i) {
            B.b[a][i] = box_empty();
            B.n[a][i] = 0;
        }
}

enum { SPLIT_SAH = 0, SPLIT_MIDPOINT = 1, SPLIT_LEAF = 2 };

// The SAH cost of a split into boxes l and r holding nl and nr references, and the position of the
// split after bin i of an axis starting at axis_min with extent ext (bvh4d.rs:307-417).
SR_HD float sah_cost(const Box4 &l, uint32_t nl, const Box4 &r, uint32_t nr) {
    return (float)nl * box_measure(l) + (float)nr * box_measure(r);
}
SR_HD float sah_split_pos(float axis_min, float ext, int i) {
    return axis_min + ((float)(i + 1) / (float)BVH_BINS) * ext;
}

// One live axis a of the SAH sweep (bvh4d.rs:307-417) over that axis's bins b / n: every split
// between bins, the cheapest so far kept (strictly cheaper replaces, axes in order 0..3).
SR_HD void sah_axis(const Box4 (&b)[BVH_BINS], const uint32_t (&n)[BVH_BINS], const BinFrame &f, int a,
                    float &best_cost, int &best_axis, float &best_pos) {
    Box4 left_box[BVH_BINS - 1];
    uint32_t left_cnt[BVH_BINS - 1];
    Box4 cur_b = box_empty();
    uint32_t cur_c = 0;
    for (int i = 0; i < BVH_BINS - 1; ++i) {
        cur_b = box_union(cur_b, b[i]);
        cur_c += n[i];
        left_box[i] = cur_b;
        left_cnt[i] = cur_c;
    }
    Box4 right_b = box_empty();
    uint32_t right_c = 0;
    for (int i = BVH_BINS - 2; i >= 0; --i) {
        right_b = box_union(right_b, b[i + 1]);
        right_c += n[i + 1];
        if (left_cnt[i] > 0 && right_c > 0) {
            float cost = sah_cost(left_box[i], left_cnt[i], right_b, right_c);
            if (cost < best_cost) {
                best_cost = cost;
                best_axis = a;
                best_pos = sah_split_pos(f.axis_min[a], f.ext[a], i);
            }
        }
    }
}

// After the sweeps (best_cost from them, +inf if none): SPLIT_SAH, else the midpoint of the
// widest centroid axis (SPLIT_MIDPOINT), or SPLIT_LEAF when every axis is flat.
SR_HD int split_after_sweeps(const Box4 &cb, const BinFrame &f, float best_cost, int &best_axis, float &best_pos) {
    if (best_cost != __builtin_inff()) return SPLIT_SAH;
    float max_e = -1.0f;
    for (int a = 0; a < 4; ++a) {
        if (f.ext[a] > max_e) {
            max_e = f.ext[a];
            best_axis = a;
        }
    }
    if (max_e < 1e-5f) return SPLIT_LEAF;
    best_pos = axis((cb.mn + cb.mx) * 0.5f, best_axis);
    return SPLIT_MIDPOINT;
}

// The split of a node with centroid bounds cb and filled bins: SPLIT_SAH or SPLIT_MIDPOINT with
// (axis, pos), or SPLIT_LEAF.
SR_HD int choose_split(const Box4 &cb, const BinFrame &f, const Bins &bins, int &best_axis, float &best_pos) {
    best_axis = 0;
    best_pos = 0.0f;
    float best_cost = __builtin_inff();
    for (int a = 0; a < 4; ++a)
        if (f.live[a]) sah_axis(bins.b[a], bins.n[a], f, a, best_cost, best_axis, best_pos);
    return split_after_sweeps(cb, f, best_cost, best_axis, best_pos);
}

}  // namespace sr
