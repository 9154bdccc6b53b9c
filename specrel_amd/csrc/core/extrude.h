// Worldtube extrusion, shared by the scene producer (host C++) and k_extrude (HIP), so both
// emit the same records bit for bit: TubeTriangle::new (mesh4d.rs:662-718), the prism step of
// extend_worldtube_simplices (mesh4d.rs:721-854) and Simplex4D::new (mesh4d.rs:312-362) with the
// GpuGeom / GpuAttr packing (gpu.rs:68-92, 358-381).
#pragma once
#include <stdint.h>

#include "minkowski.h"
#include "specrel_types.h"

namespace sr {

// WorldlineSample (mesh4d.rs:615-630).
struct Sample {
    V4 center;
    float proper_time;
    V3 phi;
    V3 accel_local;
    Tetrad tetrad;
};

// TubeTriangle: the corners sorted by vertex index (stable), their rest positions, UVs and
// standard barycentric corners, and the UV->position gradients tu, tv.
struct TubeTri {
    uint32_t idx[3];
    V3 p[3];
    float uv[3][2];
    V3 bary[3];
    V3 tu, tv, normal;
};

struct Box4 {
    V4 mn, mx;
};

const float DECIDED_AT_STATIC = -1.0e9f;

// TubeTriangle::new from a triangle's vertex indices, rest positions, corner UVs and normal.
SR_HD TubeTri tube_tri(const uint32_t idx[3], const V3 pos[3], const float uv[3][2], V3 normal) {
    const V3 std_bary[3] = {v3(1, 0, 0), v3(0, 1, 0), v3(0, 0, 1)};
    // slice::sort_by_key on the vertex index: a stable sort of three corners.
    int perm[3] = {0, 1, 2};
    for (int i = 1; i < 3; ++i)
        for (int j = i; j > 0 && idx[perm[j]] < idx[perm[j - 1]]; --j) {
            int t = perm[j];
            perm[j] = perm[j - 1];
            perm[j - 1] = t;
        }
    TubeTri t;
    for (int k = 0; k < 3; ++k) {
        t.idx[k] = idx[perm[k]];
        t.p[k] = pos[perm[k]];
        t.uv[k][0] = uv[perm[k]][0];
        t.uv[k][1] = uv[perm[k]][1];
        t.bary[k] = std_bary[perm[k]];
    }
    V3 dp1 = t.p[1] - t.p[0], dp2 = t.p[2] - t.p[0];
    float du1 = t.uv[1][0] - t.uv[0][0], dv1 = t.uv[1][1] - t.uv[0][1];
    float du2 = t.uv[2][0] - t.uv[0][0], dv2 = t.uv[2][1] - t.uv[0][1];
    float uv_det = du1 * dv2 - dv1 * du2;
    if (fabsf(uv_det) > 1e-10f) {
        float inv_det = 1.0f / uv_det;
        t.tu = (dp1 * dv2 - dp2 * dv1) * inv_det;
        t.tv = (dp2 * du1 - dp1 * du2) * inv_det;
    } else {
        t.tu = v3(0, 0, 0);
        t.tv = v3(0, 0, 0);
    }
    t.normal = normal;
    return t;
}

SR_HD void put4(float *dst, V4 v) {
    dst[0] = v.t;
    dst[1] = v.x;
    dst[2] = v.y;
    dst[3] = v.z;
}

// Simplex4D::new + packing. Returns false for a degenerate 3-simplex (no record).
SR_HD bool make_simplex(const V4 verts[4], const float uv[4][2], const V3 tri_bary[4], const float tau[4],
                        const float grad[4][2], float uvws0, float uvws1, V4 normal, V4 four_vel,
                        const float decided[4], uint32_t material, specrel_geom &g, specrel_attr &a, Box4 &box) {
    V4 v0 = verts[0];
    V4 e1 = verts[1] - v0, e2 = verts[2] - v0, e3 = verts[3] - v0;
    V4 n = cross4(e1, e2, e3);
    float n_len_sq = edot(n, n);
    if (n_len_sq < 1e-16f) return false;
    float inv_n_len = 1.0f / sqrtf(n_len_sq);
    V4 unit_n = n * inv_n_len;
    V4 raw_b1 = cross4(e2, e3, unit_n);
    float d1 = edot(raw_b1, e1);
    V4 raw_b2 = cross4(e3, e1, unit_n);
    float d2 = edot(raw_b2, e2);
    V4 raw_b3 = cross4(e1, e2, unit_n);
    float d3 = edot(raw_b3, e3);
    if (fabsf(d1) < 1e-12f || fabsf(d2) < 1e-12f || fabsf(d3) < 1e-12f) return false;
    const float inf = __builtin_inff();
    box.mn = splat4(inf);
    box.mx = splat4(-inf);
    for (int i = 0; i < 4; ++i) {
        box.mn = vmin(box.mn, verts[i]);
        box.mx = vmax(box.mx, verts[i]);
    }
    V4 eps = splat4(1e-4f);
    box.mn -= eps;
    box.mx += eps;
    put4(g.v0, v0);
    put4(g.n, unit_n);
    put4(g.b1, raw_b1 * (1.0f / d1));
    put4(g.b2, raw_b2 * (1.0f / d2));
    put4(g.b3, raw_b3 * (1.0f / d3));
    for (int i = 0; i < 4; ++i) {
        a.uv_u[i] = uv[i][0];
        a.uv_v[i] = uv[i][1];
        a.tb_x[i] = tri_bary[i].x;
        a.tb_y[i] = tri_bary[i].y;
        a.tb_z[i] = tri_bary[i].z;
        a.tau[i] = tau[i];
        a.grad_u[i] = grad[i][0];
        a.grad_v[i] = grad[i][1];
        a.decided_at[i] = decided[i];
    }
    put4(a.normal, normal);
    put4(a.four_vel, four_vel);
    a.misc[0] = uvws0;
    a.misc[1] = uvws1;
    __builtin_memcpy(&a.misc[2], &material, 4);
    a.misc[3] = 0.0f;
    return true;
}

// The event of a rest-frame body point at a worldline sample (mesh4d.rs:749).
SR_HD V4 embed(const Sample &s, V3 p) { return s.center + local_spatial_to_world(s.tetrad, p); }

// What the three tetrahedra of one prism share (triangle `tri` swept from sample sc to sample
// sn, mesh4d.rs:755-799): the six corner events, proper times, UV gradients of tau, the step's
// surface normal and average four-velocity.
struct PrismShared {
    V4 u_avg, normal;
    V4 ev[2][3];       // [time][corner a, b, c]
    float tau[2][3];   // [time][corner]
    float grad[2][2];  // [time][u, v]
    float uvws0, uvws1;
};

SR_HD PrismShared prism_shared(const TubeTri &tri, const Sample &sc, const Sample &sn) {
    PrismShared q;
    V4 u_raw = sc.tetrad.e[0] + sn.tetrad.e[0];
    float u_norm = sqrtf(fmax_(-mdot(u_raw, u_raw), 1e-12f));
    q.u_avg = u_raw * (1.0f / u_norm);
    for (int c = 0; c < 3; ++c) {
        q.ev[0][c] = embed(sc, tri.p[c]);
        q.ev[1][c] = embed(sn, tri.p[c]);
    }
    for (int c = 0; c < 3; ++c) {
        q.tau[0][c] = sc.proper_time + dot(tri.p[c], sc.phi);
        q.tau[1][c] = sn.proper_time + dot(tri.p[c], sn.phi);
    }
    q.uvws0 = length(tri.tu);
    q.uvws1 = length(tri.tv);
    q.grad[0][0] = dot(tri.tu, sc.phi);
    q.grad[0][1] = dot(tri.tv, sc.phi);
    q.grad[1][0] = dot(tri.tu, sn.phi);
    q.grad[1][1] = dot(tri.tv, sn.phi);
    V4 n4c = local_spatial_to_world(sc.tetrad, tri.normal);
    V4 n4n = local_spatial_to_world(sn.tetrad, tri.normal);
    V4 n4_raw = n4c + n4n;
    float n4_len = sqrtf(fmax_(mdot(n4_raw, n4_raw), 1e-12f));
    q.normal = n4_raw * (1.0f / n4_len);
    return q;
}

// Tetrahedron k of the prism split (A0,B0,C0)->(A1,B1,C1) (mesh4d.rs:801-830):
//   0: (A0, B0, C0, C1)   1: (A0, B0, B1, C1)   2: (A0, A1, B1, C1)
// Vertex 0 is always A0 and vertex 3 C1; vertex 1 is B0 for k < 2, else A1; vertex 2 is C0 for
// k = 0, else B1. Every attribute follows its vertex's (time, corner).
SR_HD bool prism_tet(const PrismShared &q, const TubeTri &tri, float dec0, float dec1, uint32_t material, int k,
                     specrel_geom &g, specrel_attr &a, Box4 &box) {
    // Selects, not indexed arrays, so a thread's k costs no private memory on the device.
    const bool b1 = k < 2, b2 = k == 0;  // vertex 1 is B0 (else A1); vertex 2 is C0 (else B1)
    const V4 vv[4] = {q.ev[0][0], b1 ? q.ev[0][1] : q.ev[1][0], b2 ? q.ev[0][2] : q.ev[1][1], q.ev[1][2]};
    const float uv[4][2] = {{tri.uv[0][0], tri.uv[0][1]},
                            {b1 ? tri.uv[1][0] : tri.uv[0][0], b1 ? tri.uv[1][1] : tri.uv[0][1]},
                            {b2 ? tri.uv[2][0] : tri.uv[1][0], b2 ? tri.uv[2][1] : tri.uv[1][1]},
                            {tri.uv[2][0], tri.uv[2][1]}};
    const V3 tb[4] = {tri.bary[0], b1 ? tri.bary[1] : tri.bary[0], b2 ? tri.bary[2] : tri.bary[1], tri.bary[2]};
    const float ta[4] = {q.tau[0][0], b1 ? q.tau[0][1] : q.tau[1][0], b2 ? q.tau[0][2] : q.tau[1][1], q.tau[1][2]};
    const float gg[4][2] = {{q.grad[0][0], q.grad[0][1]},
                            {b1 ? q.grad[0][0] : q.grad[1][0], b1 ? q.grad[0][1] : q.grad[1][1]},
                            {b2 ? q.grad[0][0] : q.grad[1][0], b2 ? q.grad[0][1] : q.grad[1][1]},
                            {q.grad[1][0], q.grad[1][1]}};
    const float dd[4] = {dec0, b1 ? dec0 : dec1, b2 ? dec0 : dec1, dec1};
    return make_simplex(vv, uv, tb, ta, gg, q.uvws0, q.uvws1, q.normal, q.u_avg, dd, material, g, a, box);
}

// The three tetrahedra of one prism in order: sink(ok, geom, attr, box) once per tetrahedron, ok
// false for a degenerate one (geom, attr and box then unset).
template <class Sink>
SR_HD void prism(const TubeTri &tri, const Sample &sc, const Sample &sn, float dec0, float dec1, uint32_t material,
                 Sink &&sink) {
    const PrismShared q = prism_shared(tri, sc, sn);
    specrel_geom g;
    specrel_attr a;
    Box4 box;
    for (int k = 0; k < 3; ++k) {
        bool ok = prism_tet(q, tri, dec0, dec1, material, k, g, a, box);
        sink(ok, g, a, box);
    }
}

// Sample from the C ABI layout (specrel_wl_sample).
SR_HD Sample sample_from(const float center[4], float proper_time, const float phi[3], const float accel[3],
                         const float e[4][4]) {
    Sample s;
    s.center = v4(center[0], center[1], center[2], center[3]);
    s.proper_time = proper_time;
    s.phi = v3(phi[0], phi[1], phi[2]);
    s.accel_local = v3(accel[0], accel[1], accel[2]);
    for (int k = 0; k < 4; ++k) s.tetrad.e[k] = v4(e[k][0], e[k][1], e[k][2], e[k][3]);
    return s;
}

}  // namespace sr
