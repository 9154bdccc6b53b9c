// Host-side scene producer for the SpecRel HIP backend.
//
// Restates the reference's scene pipeline (the caller side of the renderer boundary):
//   Mesh3D::{cube, uv_sphere, ground_grid}         mesh4d.rs:452-610
//   sample_inertial_worldline                      mesh4d.rs:912-937
//   sample_accelerated_worldline (f64, Simpson, Fermi-Walker)  mesh4d.rs:947-1052
//   sync_proper_time_at                            mesh4d.rs:1056-1087
//   TubeTriangle + extend_worldtube_simplices      mesh4d.rs:662-854
//   Simplex4D::new                                 mesh4d.rs:312-371
//   Bvh4D::build (8-bin SAH over t,x,y,z)          bvh4d.rs:122-141, 254-427
//   PackedSceneGpuData::from_scene / pack_stars    gpu.rs:102-119, 247-289, 346-381
//   Scene4D::common_materials                      scene.rs:181-280
//   starfield                                      star.rs:290-318
// and adds seeded generators for the synthetic workloads (icosphere field with a star sphere,
// accelerated Born-rigid bodies with light clocks).
#include <algorithm>
#include <array>
#include <cmath>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../core/bvh_split.h"
#include "../core/extrude.h"
#include "../core/minkowski.h"
#include "../core/scene_file.hpp"
#include "../core/tables.h"
#include "specrel_file.h"
#include "specrel_scene.h"

using namespace sr;

namespace {

thread_local std::string g_err;

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// ----------------------------------------------------------------------------------------
// Meshes (mesh4d.rs:437-611)
// ----------------------------------------------------------------------------------------
struct Tri {
    size_t idx[3];
    float uv[3][2];
    V3 normal;
};
struct Mesh {
    std::vector<V3> v;
    std::vector<Tri> t;
};

Mesh mesh_cube(V3 he) {
    Mesh m;
    const float zs[2] = {-he.z, he.z}, ys[2] = {-he.y, he.y}, xs[2] = {-he.x, he.x};
    for (float z : zs)
        for (float y : ys)
            for (float x : xs) m.v.push_back(v3(x, y, z));
    float ux_y = fmax_(round_(he.x / fmax_(he.y, 1e-4f)), 1.0f);
    float uz_y = fmax_(round_(he.z / fmax_(he.y, 1e-4f)), 1.0f);
    float ux_z = fmax_(round_(he.x / fmax_(he.z, 1e-4f)), 1.0f);
    struct Face {
        size_t q[4];
        V3 n;
        float us, vs;
    };
    const Face faces[6] = {
        {{0, 1, 3, 2}, v3(0, 0, -1), ux_y, 1.0f}, {{5, 4, 6, 7}, v3(0, 0, 1), ux_y, 1.0f},
        {{4, 0, 2, 6}, v3(-1, 0, 0), uz_y, 1.0f}, {{1, 5, 7, 3}, v3(1, 0, 0), uz_y, 1.0f},
        {{4, 5, 1, 0}, v3(0, -1, 0), ux_z, 1.0f}, {{2, 3, 7, 6}, v3(0, 1, 0), ux_z, 1.0f}};
    for (const Face &f : faces) {
        float uv0[2] = {0.0f, 0.0f}, uv1[2] = {f.us, 0.0f}, uv2[2] = {f.us, f.vs},
              uv3[2] = {0.0f, f.vs};
        Tri a{{f.q[0], f.q[1], f.q[2]}, {{uv0[0], uv0[1]}, {uv1[0], uv1[1]}, {uv2[0], uv2[1]}}, f.n};
        Tri b{{f.q[0], f.q[2], f.q[3]}, {{uv0[0], uv0[1]}, {uv2[0], uv2[1]}, {uv3[0], uv3[1]}}, f.n};
        m.t.push_back(a);
        m.t.push_back(b);
    }
    return m;
}

Mesh mesh_uv_sphere(float radius, size_t lat, size_t lon) {
    const float PI = 3.14159265358979323846f, TAU = 6.28318530717958647692f;
    Mesh m;
    for (size_t i = 0; i <= lat; ++i) {
        float v = (float)i / (float)lat;
        float theta = v * PI;
        float sin_t = sinf_(theta), cos_t = cosf_(theta);
        for (size_t j = 0; j <= lon; ++j) {
            float u = (float)j / (float)lon;
            float phi = u * TAU;
            float sin_p = sinf_(phi), cos_p = cosf_(phi);
            m.v.push_back(v3(radius * sin_t * cos_p, radius * cos_t, radius * sin_t * sin_p));
        }
    }
    size_t stride = lon + 1;
    for (size_t i = 0; i < lat; ++i) {
        for (size_t j = 0; j < lon; ++j) {
            size_t i00 = i * stride + j, i10 = i00 + 1, i01 = (i + 1) * stride + j, i11 = i01 + 1;
            float u0 = (float)j / (float)lon * 4.0f;
            float u1 = (float)(j + 1) / (float)lon * 4.0f;
            float v0 = (float)i / (float)lat * 2.0f;
            float v1 = (float)(i + 1) / (float)lat * 2.0f;
            if (i > 0) {
                V3 n = normalized(m.v[i00] + m.v[i10] + m.v[i11]);
                m.t.push_back(Tri{{i00, i11, i10}, {{u0, v0}, {u1, v1}, {u1, v0}}, n});
            }
            if (i + 1 < lat) {
                V3 n = normalized(m.v[i00] + m.v[i01] + m.v[i11]);
                m.t.push_back(Tri{{i00, i01, i11}, {{u0, v0}, {u0, v1}, {u1, v1}}, n});
            }
        }
    }
    return m;
}

Mesh mesh_ground(float half_x, float half_z, float y, size_t dx, size_t dz) {
    Mesh m;
    for (size_t iz = 0; iz <= dz; ++iz) {
        float fz = (float)iz / (float)dz;
        float z = -half_z + fz * (2.0f * half_z);
        for (size_t ix = 0; ix <= dx; ++ix) {
            float fx = (float)ix / (float)dx;
            float x = -half_x + fx * (2.0f * half_x);
            m.v.push_back(v3(x, y, z));
        }
    }
    size_t stride = dx + 1;
    V3 n = v3(0, 1, 0);
    for (size_t iz = 0; iz < dz; ++iz) {
        for (size_t ix = 0; ix < dx; ++ix) {
            size_t i00 = iz * stride + ix, i10 = i00 + 1, i01 = (iz + 1) * stride + ix, i11 = i01 + 1;
            V3 p00 = m.v[i00], p10 = m.v[i10], p01 = m.v[i01], p11 = m.v[i11];
            m.t.push_back(Tri{{i00, i01, i11}, {{p00.x, p00.z}, {p01.x, p01.z}, {p11.x, p11.z}}, n});
            m.t.push_back(Tri{{i00, i11, i10}, {{p00.x, p00.z}, {p11.x, p11.z}, {p10.x, p10.z}}, n});
        }
    }
    return m;
}

// Icosphere: an icosahedron whose faces are split 4-to-1 `subdiv` times with shared edge
// midpoints (so the extruded tube stays watertight), projected onto the sphere. UVs are
// longitude/latitude (u in [0,4), v in [0,2] like uv_sphere), unwrapped per triangle across
// the seam; normals are the normalised vertex sums, as uv_sphere's.
Mesh mesh_icosphere(float radius, int subdiv) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<std::array<double, 3>> P = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0},
                                            {0, -1, t}, {0, 1, t},  {0, -1, -t}, {0, 1, -t},
                                            {t, 0, -1}, {t, 0, 1},  {-t, 0, -1}, {-t, 0, 1}};
    std::vector<std::array<size_t, 3>> F = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                                            {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                            {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                                            {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
    for (auto &p : P) {
        double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        p = {p[0] / l, p[1] / l, p[2] / l};
    }
    for (int s = 0; s < subdiv; ++s) {
        std::vector<std::array<size_t, 3>> G;
        std::vector<std::pair<uint64_t, size_t>> cache;
        auto mid = [&](size_t a, size_t b) {
            uint64_t key = ((uint64_t)std::min(a, b) << 32) | (uint64_t)std::max(a, b);
            for (auto &kv : cache)
                if (kv.first == key) return kv.second;
            std::array<double, 3> m = {(P[a][0] + P[b][0]) * 0.5, (P[a][1] + P[b][1]) * 0.5,
                                       (P[a][2] + P[b][2]) * 0.5};
            double l = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
            P.push_back({m[0] / l, m[1] / l, m[2] / l});
            cache.push_back({key, P.size() - 1});
            return P.size() - 1;
        };
        for (auto &f : F) {
            size_t a = mid(f[0], f[1]), b = mid(f[1], f[2]), c = mid(f[2], f[0]);
            G.push_back({f[0], a, c});
            G.push_back({f[1], b, a});
            G.push_back({f[2], c, b});
            G.push_back({a, b, c});
        }
        F.swap(G);
    }
    Mesh m;
    for (auto &p : P) m.v.push_back(v3((float)(p[0] * radius), (float)(p[1] * radius), (float)(p[2] * radius)));
    const double PI = 3.14159265358979323846;
    for (auto &f : F) {
        Tri tr;
        double us[3], vs[3];
        for (int k = 0; k < 3; ++k) {
            const auto &p = P[f[k]];
            us[k] = (std::atan2(p[2], p[0]) / (2.0 * PI) + 0.5) * 4.0;
            vs[k] = std::acos(std::max(-1.0, std::min(1.0, p[1]))) / PI * 2.0;
        }
        double umax = std::max(us[0], std::max(us[1], us[2]));
        for (int k = 0; k < 3; ++k)
            if (umax - us[k] > 2.0) us[k] += 4.0;
        for (int k = 0; k < 3; ++k) {
            tr.idx[k] = f[k];
            tr.uv[k][0] = (float)us[k];
            tr.uv[k][1] = (float)vs[k];
        }
        tr.normal = normalized(m.v[f[0]] + m.v[f[1]] + m.v[f[2]]);
        m.t.push_back(tr);
    }
    return m;
}

Mesh make_mesh(uint32_t kind, const float *p) {
    switch (kind) {
        case SPECREL_MESH_CUBE: return mesh_cube(v3(p[0], p[1], p[2]));
        case SPECREL_MESH_UV_SPHERE: return mesh_uv_sphere(p[0], (size_t)p[1], (size_t)p[2]);
        case SPECREL_MESH_GROUND: return mesh_ground(p[0], p[1], p[2], (size_t)p[3], (size_t)p[4]);
        case SPECREL_MESH_ICOSPHERE: return mesh_icosphere(p[0], (int)p[1]);
    }
    throw Error("unknown mesh kind");
}

// ----------------------------------------------------------------------------------------
// Worldlines (mesh4d.rs:613-1087)
// ----------------------------------------------------------------------------------------
std::vector<Sample> sample_inertial(V3 pos, V3 beta, float t_min, float t_max, size_t steps) {
    Tetrad T = tetrad_from_beta(beta);
    double inv_gamma = 1.0 / (double)T.e[0].t;
    std::vector<Sample> out;
    out.reserve(steps + 1);
    for (size_t i = 0; i <= steps; ++i) {
        double frac = (double)i / (double)steps;
        double t = (double)t_min + frac * (double)(t_max - t_min);
        V3 p = pos + beta * (float)t;
        float tau = (float)(t * inv_gamma);
        out.push_back(Sample{from_ts((float)t, p), tau, v3(0, 0, 0), v3(0, 0, 0), T});
    }
    return out;
}

using PosFn = std::function<void(double, double *)>;

std::vector<Sample> sample_accelerated(const PosFn &pos_fn, float t_min, float t_max, size_t steps) {
    const double dt_deriv = 1e-5;
    const int SUBSTEPS = 8;
    auto velocity = [&](double t, double v[3]) {
        double pm[3], pp[3];
        pos_fn(t - dt_deriv, pm);
        pos_fn(t + dt_deriv, pp);
        double inv_2h = 0.5 / dt_deriv;
        v[0] = (pp[0] - pm[0]) * inv_2h;
        v[1] = (pp[1] - pm[1]) * inv_2h;
        v[2] = (pp[2] - pm[2]) * inv_2h;
        double speed_sq = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        if (!(speed_sq < 1.0)) throw Error("worldline becomes superluminal");
    };
    auto four_vel = [&](double t, double *inv_gamma_out) {
        double v[3];
        velocity(t, v);
        double b2 = std::min(v[0] * v[0] + v[1] * v[1] + v[2] * v[2], 0.999999999);
        double inv_gamma = std::sqrt(1.0 - b2);
        double gamma = 1.0 / inv_gamma;
        *inv_gamma_out = inv_gamma;
        return v4((float)gamma, (float)(gamma * v[0]), (float)(gamma * v[1]), (float)(gamma * v[2]));
    };
    auto proper_accel_world = [&](double t) {
        double ig, igm, igp;
        four_vel(t, &ig);
        V4 um = four_vel(t - dt_deriv, &igm);
        V4 up = four_vel(t + dt_deriv, &igp);
        V4 du_dt = (up - um) * (0.5f / (float)dt_deriv);
        return du_dt * (1.0f / (float)ig);
    };
    double t_min_d = (double)t_min;
    double t_span_d = (double)(t_max - t_min);
    double v0[3];
    velocity(t_min_d, v0);
    Tetrad T = tetrad_from_beta(v3((float)v0[0], (float)v0[1], (float)v0[2]));
    double tau = 0.0;
    V3 phi = v3(0, 0, 0);
    double prev_t = t_min_d;
    std::vector<Sample> out;
    out.reserve(steps + 1);
    for (size_t i = 0; i <= steps; ++i) {
        double frac = (double)i / (double)steps;
        double t = t_min_d + frac * t_span_d;
        if (i > 0) {
            double sub_dt = (t - prev_t) / (double)SUBSTEPS;
            for (int s = 0; s < SUBSTEPS; ++s) {
                double ta = prev_t + (double)s * sub_dt;
                double tb = ta + sub_dt;
                double tm = 0.5 * (ta + tb);
                double inv_ga, inv_gm, inv_gb;
                V4 ua = four_vel(ta, &inv_ga);
                four_vel(tm, &inv_gm);
                V4 ub = four_vel(tb, &inv_gb);
                tau += (sub_dt / 6.0) * (inv_ga + 4.0 * inv_gm + inv_gb);
                V4 du = ub - ua;
                V3 before = world_to_local_spatial(T, du);
                boost_to(T, ub);
                V3 after = world_to_local_spatial(T, du);
                phi += (before + after) * 0.5f;
            }
        }
        prev_t = t;
        double p[3];
        pos_fn(t, p);
        V4 a_world = proper_accel_world(t);
        V3 a_local = world_to_local_spatial(T, a_world);
        out.push_back(Sample{v4((float)t, (float)p[0], (float)p[1], (float)p[2]), (float)tau, phi,
                             a_local, T});
    }
    return out;
}

void sync_proper_time_at(std::vector<Sample> &s, float t_sync) {
    if (s.empty()) return;
    float tau_at = s[0].proper_time;
    V3 phi_at = s[0].phi;
    if (t_sync <= s[0].center.t) {
        tau_at = s[0].proper_time;
        phi_at = s[0].phi;
    } else if (t_sync >= s.back().center.t) {
        tau_at = s.back().proper_time;
        phi_at = s.back().phi;
    } else {
        for (size_t i = 0; i + 1 < s.size(); ++i) {
            const Sample &s0 = s[i], &s1 = s[i + 1];
            if (t_sync >= s0.center.t && t_sync <= s1.center.t) {
                float dt = fmax_(s1.center.t - s0.center.t, 1e-6f);
                float frac = (t_sync - s0.center.t) / dt;
                tau_at = s0.proper_time + frac * (s1.proper_time - s0.proper_time);
                phi_at = s0.phi + (s1.phi - s0.phi) * frac;
                break;
            }
        }
    }
    for (Sample &x : s) {
        x.proper_time -= tau_at;
        x.phi -= phi_at;
    }
}

// Scene 1's traveling twin (scene.rs:124-138).
void twin_x_vx(double t_rel, double *x, double *vx) {
    const double x_start = 1.4, t_dep = -14.0, t_ret = 14.0;
    if (t_rel <= t_dep || t_rel >= t_ret) {
        *x = x_start;
        *vx = 0.0;
        return;
    }
    double period = t_ret - t_dep;
    double omega = 6.283185307179586 / period;
    double amp = 4.2;
    double phase = omega * (t_rel - t_dep);
    *x = x_start + amp * (1.0 - std::cos(phase));
    *vx = amp * omega * std::sin(phase);
}

std::vector<Sample> sample_worldline(uint32_t kind, const double *d) {
    switch (kind) {
        case SPECREL_WL_INERTIAL:
            return sample_inertial(v3((float)d[0], (float)d[1], (float)d[2]),
                                   v3((float)d[3], (float)d[4], (float)d[5]), (float)d[6], (float)d[7],
                                   (size_t)d[8]);
        case SPECREL_WL_HYPERBOLIC: {
            // Constant proper acceleration a along unit dir, at rest at `centre` at t_turn, within
            // the window |a (t - t_turn)| <= w (speed up to w / sqrt(1 + w^2)), inertial beyond:
            // x = centre + dir f(|t - t_turn|), f(s) = sqrt(1/a^2 + s^2) - 1/a.
            double c[3] = {d[0], d[1], d[2]}, u[3] = {d[3], d[4], d[5]};
            double a = d[6], t_turn = d[7], w = d[8];
            double ia = 1.0 / a, tw = w / a;
            double f_edge = std::sqrt(ia * ia + tw * tw) - ia;
            double v_edge = tw / std::sqrt(ia * ia + tw * tw);
            PosFn f = [=](double t, double *p) {
                double s = std::fabs(t - t_turn);
                double x = s <= tw ? std::sqrt(ia * ia + s * s) - ia : f_edge + v_edge * (s - tw);
                p[0] = c[0] + u[0] * x;
                p[1] = c[1] + u[1] * x;
                p[2] = c[2] + u[2] * x;
            };
            return sample_accelerated(f, (float)d[9], (float)d[10], (size_t)d[11]);
        }
        case SPECREL_WL_TWIN: {
            double t_off = d[0], y = d[1], z = d[2];
            PosFn f = [=](double t, double *p) {
                double x, vx;
                twin_x_vx(t - t_off, &x, &vx);
                p[0] = x;
                p[1] = y;
                p[2] = z;
            };
            return sample_accelerated(f, (float)d[3], (float)d[4], (size_t)d[5]);
        }
        case SPECREL_WL_CIRCLE: {
            double radius = d[0], beta = d[1], omega = beta / radius;
            PosFn f = [=](double t, double *p) {
                p[0] = radius * std::cos(omega * t);
                p[1] = 0.0;
                p[2] = radius * std::sin(omega * t);
            };
            return sample_accelerated(f, (float)d[2], (float)d[3], (size_t)d[4]);
        }
        case SPECREL_WL_OSCILLATE: {
            // x = cx + amp sin(omega t) at fixed y, z (scene.rs:553-563)
            double cx = d[0], amp = d[1], omega = d[2], y = d[3], z = d[4];
            PosFn f = [=](double t, double *p) {
                p[0] = cx + amp * std::sin(omega * t);
                p[1] = y;
                p[2] = z;
            };
            return sample_accelerated(f, (float)d[5], (float)d[6], (size_t)d[7]);
        }
        case SPECREL_WL_ORBIT: {
            // angle = omega (t - t_ref) + phase; (cx + r cos, cy + wobble sin(2 angle), cz + r sin)
            // (scene.rs:567-578, 683-700)
            double cx = d[0], cy = d[1], cz = d[2], r = d[3], omega = d[4], t_ref = d[5], phase = d[6], wob = d[7];
            PosFn f = [=](double t, double *p) {
                double angle = omega * (t - t_ref) + phase;
                double sn = std::sin(angle), cs = std::cos(angle);
                p[0] = cx + r * cs;
                p[1] = cy + wob * std::sin(2.0 * angle);
                p[2] = cz + r * sn;
            };
            return sample_accelerated(f, (float)d[8], (float)d[9], (size_t)d[10]);
        }
        case SPECREL_WL_BORN_RAMP: {
            double beta_max = d[0];
            double k = 3.141592653589793 / 20.0;
            PosFn f = [=](double t, double *p) {
                double tc = std::min(std::max(t, 0.0), 20.0);
                p[0] = beta_max * (0.5 * tc - std::sin(2.0 * k * tc) / (4.0 * k));
                p[1] = 0.0;
                p[2] = 0.0;
            };
            return sample_accelerated(f, (float)d[1], (float)d[2], (size_t)d[3]);
        }
    }
    throw Error("unknown worldline kind");
}

// ----------------------------------------------------------------------------------------
// Simplices (mesh4d.rs:214-371) and worldtube extrusion (mesh4d.rs:662-854)
// ----------------------------------------------------------------------------------------
using Aabb = Box4;

struct Prims {
    std::vector<specrel_geom> geoms;
    std::vector<specrel_attr> attrs;
    std::vector<Aabb> boxes;
    size_t size() const { return geoms.size(); }
    void reserve(size_t n) {
        geoms.reserve(n);
        attrs.reserve(n);
        boxes.reserve(n);
    }
};

TubeTri tube_tri_of(const Mesh &m, const Tri &tri) {
    uint32_t idx[3];
    V3 pos[3];
    for (int k = 0; k < 3; ++k) {
        idx[k] = (uint32_t)tri.idx[k];
        pos[k] = m.v[tri.idx[k]];
    }
    return tube_tri(idx, pos, tri.uv, tri.normal);
}

// The Rindler-horizon check of extend_worldtube_simplices (mesh4d.rs:738-747).
void check_rindler(const std::vector<V3> &verts, const Sample *samples, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        const Sample &s = samples[i];
        for (const V3 &p : verts) {
            float rindler = 1.0f + dot(s.accel_local, p);
            if (!(rindler > 0.0f))
                throw Error("Born-rigid worldtube crosses its Rindler horizon (sample t " +
                            std::to_string(s.center.t) + ", proper acceleration " +
                            std::to_string(length(s.accel_local)) + ", body extent " + std::to_string(length(p)) + ")");
        }
    }
}

// extend_worldtube_simplices (mesh4d.rs:721-854) over prepared triangles: steps in order,
// triangles in order, the three tetrahedra of each prism in order; degenerate ones dropped.
void extrude_tris(const std::vector<TubeTri> &tris, const Sample *samples, size_t n_samples, const float *decided,
                  uint32_t material, Prims &out) {
    if (n_samples < 2) return;
    size_t n_steps = n_samples - 1;
    size_t need = out.size() + n_steps * 3 * tris.size();
    if (need > out.geoms.capacity()) out.reserve(std::max(need, 2 * out.geoms.capacity()));
    for (size_t step = 0; step < n_steps; ++step) {
        const Sample &sc = samples[step], &sn = samples[step + 1];
        float dec0 = decided ? decided[step] : DECIDED_AT_STATIC;
        float dec1 = decided ? decided[step + 1] : DECIDED_AT_STATIC;
        for (const TubeTri &tri : tris)
            prism(tri, sc, sn, dec0, dec1, material,
                  [&](bool ok, const specrel_geom &g, const specrel_attr &a, const Box4 &box) {
                      if (!ok) return;
                      out.geoms.push_back(g);
                      out.attrs.push_back(a);
                      out.boxes.push_back(box);
                  });
    }
}

void extrude(const Mesh &mesh, const std::vector<Sample> &samples, const float *decided, uint32_t material,
             Prims &out) {
    if (samples.size() < 2) return;
    check_rindler(mesh.v, samples.data(), samples.size());
    std::vector<TubeTri> tris;
    tris.reserve(mesh.t.size());
    for (const Tri &t : mesh.t) tris.push_back(tube_tri_of(mesh, t));
    extrude_tris(tris, samples.data(), samples.size(), decided, material, out);
}

// ----------------------------------------------------------------------------------------
// Bvh4D::build (bvh4d.rs:8-13, 122-141, 254-427)
// ----------------------------------------------------------------------------------------
struct BuildRef {
    Aabb box;
    V4 c;
    uint32_t index;
};

struct Node {
    Aabb box;
    uint32_t right_or_first, count, left;
};

const size_t PARALLEL_MIN = 4096;
const size_t FOLD_CHUNK = 1u << 16;

std::atomic<int> g_threads_live{0};
int g_thread_budget = 1;

template <class F>
void parallel_chunks(size_t n, const F &f) {
    // f(begin, end, chunk_id) over ceil(n / FOLD_CHUNK) chunks, run on up to g_thread_budget threads.
    size_t nchunks = (n + FOLD_CHUNK - 1) / FOLD_CHUNK;
    int workers = (int)std::min<size_t>(nchunks, (size_t)std::max(1, g_thread_budget));
    if (workers <= 1) {
        for (size_t c = 0; c < nchunks; ++c) f(c * FOLD_CHUNK, std::min(n, (c + 1) * FOLD_CHUNK), c);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> th;
    for (int w = 0; w < workers; ++w)
        th.emplace_back([&] {
            for (;;) {
                size_t c = next.fetch_add(1);
                if (c >= nchunks) break;
                f(c * FOLD_CHUNK, std::min(n, (c + 1) * FOLD_CHUNK), c);
            }
        });
    for (auto &t : th) t.join();
}

// The node over refs: bounds, and 0 for a leaf or the split point (refs reordered).
size_t split(BuildRef *refs, size_t count, Aabb &bounds_out) {
    Aabb bounds = box_empty(), cb = box_empty();
    if (count > FOLD_CHUNK) {
        size_t nchunks = (count + FOLD_CHUNK - 1) / FOLD_CHUNK;
        std::vector<Aabb> pb(nchunks, box_empty()), pc(nchunks, box_empty());
        parallel_chunks(count, [&](size_t b, size_t e, size_t c) {
            Aabb x = box_empty(), y = box_empty();
            for (size_t i = b; i < e; ++i) {
                x = box_union(x, refs[i].box);
                box_add_point(y, refs[i].c);
            }
            pb[c] = x;
            pc[c] = y;
        });
        for (size_t c = 0; c < nchunks; ++c) {
            bounds = box_union(bounds, pb[c]);
            cb = box_union(cb, pc[c]);
        }
    } else {
        for (size_t i = 0; i < count; ++i) {
            bounds = box_union(bounds, refs[i].box);
            box_add_point(cb, refs[i].c);
        }
    }
    bounds_out = bounds;
    if (count <= 3) return 0;

    const BinFrame f = bin_frame(cb);
    auto bin_range = [&](Bins &B, size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
            const BuildRef &r = refs[i];
            for (int a = 0; a < 4; ++a) {
                if (!f.live[a]) continue;
                uint32_t k = bin_of(f, a, axis(r.c, a));
                B.b[a][k] = box_union(B.b[a][k], r.box);
                B.n[a][k] += 1;
            }
        }
    };
    Bins bins;
    bins_clear(bins);
    if (count > FOLD_CHUNK) {
        size_t nchunks = (count + FOLD_CHUNK - 1) / FOLD_CHUNK;
        std::vector<Bins> part(nchunks);
        parallel_chunks(count, [&](size_t b, size_t e, size_t c) {
            bins_clear(part[c]);
            bin_range(part[c], b, e);
        });
        for (size_t c = 0; c < nchunks; ++c)
            for (int a = 0; a < 4; ++a)
                for (int i = 0; i < BVH_BINS; ++i) {
                    bins.b[a][i] = box_union(bins.b[a][i], part[c].b[a][i]);
                    bins.n[a][i] += part[c].n[a][i];
                }
    } else {
        bin_range(bins, 0, count);
    }
    int best_axis;
    float best_split_pos;
    if (choose_split(cb, f, bins, best_axis, best_split_pos) == SPLIT_LEAF) return 0;
    // Partition by swapping (bvh4d.rs:409-416): the left side keeps its order.
    size_t mid = 0;
    for (size_t i = 0; i < count; ++i) {
        if (axis(refs[i].c, best_axis) < best_split_pos) {
            std::swap(refs[i], refs[mid]);
            mid += 1;
        }
    }
    if (mid == 0 || mid == count) {
        // select_nth_unstable_by fallback (bvh4d.rs:418-425): any permutation that puts the
        // count/2 smallest centroids on the axis first satisfies it; the one used here is a
        // stable sort by that centroid (ties keep their order), which the device builder
        // reproduces exactly.
        mid = count / 2;
        int ax = best_axis;
        std::stable_sort(refs, refs + count, [ax](const BuildRef &x, const BuildRef &y) {
            return axis(x.c, ax) < axis(y.c, ax);
        });
    }
    return mid;
}

void build_recursive(std::vector<Node> &nodes, BuildRef *refs, size_t count, size_t first) {
    size_t node_idx = nodes.size();
    Aabb bounds;
    size_t mid = split(refs, count, bounds);
    if (mid == 0) {
        nodes.push_back(Node{bounds, (uint32_t)first, (uint32_t)count, 0});
        return;
    }
    nodes.push_back(Node{bounds, 0, 0, (uint32_t)(node_idx + 1)});
    build_recursive(nodes, refs, mid, first);
    uint32_t right_idx = (uint32_t)nodes.size();
    build_recursive(nodes, refs + mid, count - mid, first + mid);
    nodes[node_idx].right_or_first = right_idx;
}

inline Node moved(Node n, uint32_t dn) {
    if (n.count > 0) return n;
    n.left += dn;
    n.right_or_first += dn;
    return n;
}

// Parallel subtree build; produces the same array as the serial depth-first build.
std::vector<Node> build_subtree(BuildRef *refs, size_t count, size_t first, int depth) {
    std::vector<Node> nodes;
    if (count < PARALLEL_MIN || depth > 12) {
        nodes.reserve(count);
        build_recursive(nodes, refs, count, first);
        return nodes;
    }
    Aabb bounds;
    size_t mid = split(refs, count, bounds);
    if (mid == 0) {
        nodes.push_back(Node{bounds, (uint32_t)first, (uint32_t)count, 0});
        return nodes;
    }
    std::vector<Node> left, right;
    if (g_threads_live.fetch_add(1) < g_thread_budget) {
        std::thread th([&] { left = build_subtree(refs, mid, first, depth + 1); });
        right = build_subtree(refs + mid, count - mid, first + mid, depth + 1);
        th.join();
    } else {
        left = build_subtree(refs, mid, first, depth + 1);
        right = build_subtree(refs + mid, count - mid, first + mid, depth + 1);
    }
    g_threads_live.fetch_sub(1);
    uint32_t right_idx = 1 + (uint32_t)left.size();
    nodes.reserve(1 + left.size() + right.size());
    nodes.push_back(Node{bounds, right_idx, 0, 1});
    for (const Node &n : left) nodes.push_back(moved(n, 1));
    left.clear();
    left.shrink_to_fit();
    for (const Node &n : right) nodes.push_back(moved(n, right_idx));
    return nodes;
}

// ----------------------------------------------------------------------------------------
// Stars (star.rs:55-180, 290-318) and materials (scene.rs:181-280)
// ----------------------------------------------------------------------------------------
struct StarSeg {
    V4 event, u;
    float s_min, s_max;
};
struct Star {
    std::vector<StarSeg> segs;
    float temperature, intensity;
};
const float UNBOUNDED = 1e30f;

Star star_moving(V3 pos, V3 beta, float T, float I) {
    Star s;
    s.segs.push_back(StarSeg{from_ts(0.0f, pos), four_velocity_from_beta(beta), -UNBOUNDED, UNBOUNDED});
    s.temperature = T;
    s.intensity = I;
    return s;
}
Star star_along(const V4 *ev, size_t n, float T, float I) {
    Star s;
    for (size_t i = 0; i + 1 < n; ++i) {
        V4 d = ev[i + 1] - ev[i];
        float len = sqrtf(fmax_(-mdot(d, d), 0.0f));
        if (len > 1e-6f && d.t > 0.0f) s.segs.push_back(StarSeg{ev[i], d * (1.0f / len), 0.0f, len});
    }
    s.temperature = T;
    s.intensity = I;
    return s;
}

inline float planck(float lambda_um, float temperature) {
    float x = fmin_(14388.0f / (lambda_um * fmax_(temperature, 1.0f)), 80.0f);
    float l5 = powf_(lambda_um, 5.0f);  // powi(5) of a constant, folded in double
    return 1.0f / (l5 * (expf_(x) - 1.0f));
}
inline V3 blackbody_rgb(float T) {
    const float R = 0.610f, G = 0.550f, B = 0.465f, W = 6504.0f;
    return v3(planck(R, T) / planck(R, W), planck(G, T) / planck(G, W), planck(B, T) / planck(B, W));
}
inline float luminance(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
inline float color_norm(const Star &s) {
    float lum = luminance(blackbody_rgb(s.temperature));
    return lum > 0.0f ? s.intensity / lum : 0.0f;
}

std::vector<Star> starfield(V3 center, size_t count, float r_min, float r_max, uint32_t seed) {
    const float TAU = 6.28318530717958647692f;
    uint32_t state = seed * 747796405u + 2891336453u;
    auto next = [&]() {
        state = state * 747796405u + 2891336453u;
        uint32_t word = ((state >> ((state >> 28) + 4)) ^ state) * 277803737u;
        return (float)((word >> 22) ^ word) / 4294967296.0f;
    };
    std::vector<Star> out;
    for (size_t i = 0; i < count; ++i) {
        float z = next() * 1.7f - 0.7f;
        float phi = next() * TAU;
        float rho = sqrtf(fmax_(1.0f - z * z, 0.0f));
        V3 dir = v3(rho * cosf_(phi), z, rho * sinf_(phi));
        float r = r_min + (r_max - r_min) * next();
        float n1 = next();
        float n2 = next();
        float temperature = 2800.0f + 14000.0f * n1 * n2;
        float n3 = next();
        float intensity = r * r * 2e-4f * powf_(10.0f, 2.0f * n3);
        V3 beta = v3(0, 0, 0);
        if (i % 7 == 3) {
            float a = next() * TAU;
            float n4 = next();
            beta = v3(cosf_(a), 0.2f * (n4 - 0.5f), sinf_(a)) * 0.6f;
        }
        out.push_back(star_moving(center + dir * r, beta, temperature, intensity));
    }
    return out;
}

specrel_material mat(V3 base, V3 checker, float emission, float grid, float wire, uint32_t pattern = 0,
                     float cavity = 1.0f) {
    specrel_material m;
    m.base[0] = base.x;
    m.base[1] = base.y;
    m.base[2] = base.z;
    m.base[3] = emission;
    m.checker[0] = checker.x;
    m.checker[1] = checker.y;
    m.checker[2] = checker.z;
    m.checker[3] = grid;
    m.params[0] = wire;
    std::memcpy(&m.params[1], &pattern, 4);
    m.params[2] = cavity;
    m.params[3] = 0.0f;
    return m;
}

std::vector<specrel_material> common_materials() {
    auto hull = [](V3 c) { return mat(c, c, 0.35f, 1.0f, 0.0f); };
    return {
        mat(v3(0.22f, 0.25f, 0.30f), v3(0.13f, 0.15f, 0.19f), 0.18f, 1.0f, 0.012f),
        mat(v3(0.20f, 0.72f, 0.95f), v3(0.10f, 0.25f, 0.45f), 0.28f, 1.0f, 0.025f, 1),
        mat(v3(0.96f, 0.32f, 0.24f), v3(0.45f, 0.12f, 0.10f), 0.28f, 1.0f, 0.025f, 1),
        mat(v3(0.96f, 0.78f, 0.20f), v3(0.45f, 0.30f, 0.08f), 0.28f, 1.0f, 0.025f, 1),
        mat(v3(0.88f, 0.32f, 0.88f), v3(0.32f, 0.10f, 0.38f), 0.28f, 1.0f, 0.025f, 1),
        mat(v3(0.20f, 0.95f, 0.65f), v3(0.08f, 0.22f, 0.18f), 0.32f, 1.0f, 0.025f, 2, 1.0f),
        mat(v3(0.85f, 0.82f, 0.72f), v3(0.22f, 0.52f, 0.62f), 0.18f, 2.0f, 0.025f),
        mat(v3(0.25f, 0.92f, 1.00f), v3(0.85f, 0.94f, 0.98f), 0.28f, 2.0f, 0.030f),
        mat(v3(0.96f, 0.38f, 0.22f), v3(0.95f, 0.86f, 0.72f), 0.22f, 2.0f, 0.030f),
        mat(v3(1.00f, 0.10f, 0.025f), v3(1.00f, 0.10f, 0.025f), 1.20f, 1.0f, 0.0f),
        hull(v3(0.20f, 0.72f, 0.95f)),
        hull(v3(0.96f, 0.32f, 0.24f)),
        hull(v3(0.96f, 0.78f, 0.20f)),
        hull(v3(0.88f, 0.32f, 0.88f)),
    };
}

// SplitMix64 (Steele, Lea, Flood 2014) for the synthetic generators.
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double range(double a, double b) { return a + (b - a) * uniform(); }
    void unit(double *d) {
        double z = range(-1.0, 1.0), ph = range(0.0, 6.283185307179586);
        double r = std::sqrt(std::max(0.0, 1.0 - z * z));
        d[0] = r * std::cos(ph);
        d[1] = z;
        d[2] = r * std::sin(ph);
    }
};

}  // namespace

// ----------------------------------------------------------------------------------------
// Builder + host scene
// ----------------------------------------------------------------------------------------
// A tube recorded for deferred extrusion: its mesh and samples in the C ABI layout.
struct PendingTube {
    std::vector<float> verts, uv, normal, decided;
    std::vector<uint32_t> idx;
    std::vector<specrel_wl_sample> samples;
    uint32_t material = 0;
    size_t object = 0;
};

struct specrel_builder {
    Prims prims;
    bool deferred = false;
    std::vector<PendingTube> pending;
    std::vector<specrel_object> objects;
    std::vector<Star> stars;
    std::vector<specrel_material> materials = common_materials();
    V3 light = v3(2.0f, 6.5f, 2.0f);
};

struct specrel_host_scene {
    std::vector<specrel_node> nodes;
    std::vector<specrel_geom> geoms;
    std::vector<specrel_attr> attrs;
    std::vector<uint32_t> order;
    std::vector<specrel_material> materials;
    std::vector<float> star_words;
    std::vector<float> star_ti;
    std::vector<specrel_object> objects;
    uint32_t num_stars = 0;
    float light[4] = {0, 0, 0, 0};
};

namespace {

template <class F>
int guarded(F &&f) {
    try {
        return f();
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    } catch (...) {
        g_err = "unknown error";
        return -1;
    }
}

// Bvh4D::build over the records' padded bounds: nodes in preorder and the leaf order.
void build_tree(const Aabb *boxes, size_t n, std::vector<specrel_node> &out, std::vector<uint32_t> &order) {
    unsigned hw = std::thread::hardware_concurrency();
    g_thread_budget = (int)std::max(1u, std::min(hw ? hw : 1u, 64u));
    std::vector<Node> nodes;
    order.assign(n, 0);
    if (n > 0) {
        std::vector<BuildRef> refs(n);
        for (size_t i = 0; i < n; ++i) {
            const Aabb &bx = boxes[i];
            refs[i] = BuildRef{bx, (bx.mn + bx.mx) * 0.5f, (uint32_t)i};
        }
        g_threads_live = 0;
        nodes = build_subtree(refs.data(), n, 0, 0);
        for (size_t i = 0; i < n; ++i) order[i] = refs[i].index;
    }
    out.resize(nodes.size());
    for (size_t i = 0; i < nodes.size(); ++i) {
        specrel_node &o = out[i];
        put4(o.min, nodes[i].box.mn);
        put4(o.max, nodes[i].box.mx);
        o.right_or_first = nodes[i].right_or_first;
        o.count = nodes[i].count;
        o.left = nodes[i].left;
        o.pad = 0;
    }
}

void finish_rest(specrel_builder &b, specrel_host_scene &s);

void finish_into(specrel_builder &b, specrel_host_scene &s) {
    std::vector<uint32_t> order;
    build_tree(b.prims.boxes.data(), b.prims.size(), s.nodes, order);
    s.order = order;
    finish_rest(b, s);
}

// Everything after the tree: leaf-order records, materials, objects, packed stars, light.
void finish_rest(specrel_builder &b, specrel_host_scene &s) {
    size_t n = b.prims.size();
    const std::vector<uint32_t> &order = s.order;
    s.geoms.resize(n);
    s.attrs.resize(n);
    for (size_t i = 0; i < n; ++i) {
        s.geoms[i] = b.prims.geoms[order[i]];
        s.attrs[i] = b.prims.attrs[order[i]];
    }
    s.materials = b.materials;
    s.objects = b.objects;
    // pack_stars (gpu.rs:106-119)
    s.star_words.clear();
    s.star_ti.clear();
    uint32_t first = 0;
    for (const Star &st : b.stars) {
        uint32_t cnt = (uint32_t)st.segs.size();
        float cn = color_norm(st);
        float hdr[4] = {st.temperature, cn, 0.0f, 0.0f};
        std::memcpy(&hdr[2], &first, 4);
        std::memcpy(&hdr[3], &cnt, 4);
        s.star_words.insert(s.star_words.end(), hdr, hdr + 4);
        s.star_ti.push_back(st.temperature);
        s.star_ti.push_back(st.intensity);
        first += cnt;
    }
    for (const Star &st : b.stars) {
        for (const StarSeg &g : st.segs) {
            float w[12] = {g.event.t, g.event.x, g.event.y, g.event.z, g.u.t, g.u.x, g.u.y, g.u.z,
                           g.s_min, g.s_max, 0.0f, 0.0f};
            s.star_words.insert(s.star_words.end(), w, w + 12);
        }
    }
    s.num_stars = (uint32_t)b.stars.size();
    s.light[0] = 0.0f;
    s.light[1] = b.light.x;
    s.light[2] = b.light.y;
    s.light[3] = b.light.z;
}

specrel_wl_sample to_c(const Sample &s) {
    specrel_wl_sample o;
    std::memset(&o, 0, sizeof o);
    put4(o.center, s.center);
    o.proper_time = s.proper_time;
    o.phi[0] = s.phi.x;
    o.phi[1] = s.phi.y;
    o.phi[2] = s.phi.z;
    o.proper_accel_local[0] = s.accel_local.x;
    o.proper_accel_local[1] = s.accel_local.y;
    o.proper_accel_local[2] = s.accel_local.z;
    for (int k = 0; k < 4; ++k) put4(o.e[k], s.tetrad.e[k]);
    return o;
}

// Extrusion of one tube descriptor (the mesh and samples in the C ABI layout), appending to out.
void extrude_tube(const specrel_tube &t, Prims &out) {
    if (t.n_samples < 2) return;
    if (!t.vertices || !t.tri_index || !t.tri_uv || !t.tri_normal || !t.samples)
        throw Error("tube arrays are NULL");
    std::vector<V3> verts(t.n_vertices);
    for (uint32_t i = 0; i < t.n_vertices; ++i)
        verts[i] = v3(t.vertices[3 * i], t.vertices[3 * i + 1], t.vertices[3 * i + 2]);
    std::vector<Sample> samples(t.n_samples);
    for (uint32_t i = 0; i < t.n_samples; ++i) {
        const specrel_wl_sample &w = t.samples[i];
        samples[i] = sample_from(w.center, w.proper_time, w.phi, w.proper_accel_local, w.e);
    }
    check_rindler(verts, samples.data(), samples.size());
    std::vector<TubeTri> tris(t.n_triangles);
    for (uint32_t k = 0; k < t.n_triangles; ++k) {
        uint32_t idx[3];
        V3 pos[3];
        float uv[3][2];
        for (int c = 0; c < 3; ++c) {
            idx[c] = t.tri_index[3 * k + c];
            if (idx[c] >= t.n_vertices) throw Error("triangle vertex index out of range");
            pos[c] = verts[idx[c]];
            uv[c][0] = t.tri_uv[6 * k + 2 * c];
            uv[c][1] = t.tri_uv[6 * k + 2 * c + 1];
        }
        tris[k] = tube_tri(idx, pos, uv, v3(t.tri_normal[3 * k], t.tri_normal[3 * k + 1], t.tri_normal[3 * k + 2]));
    }
    extrude_tris(tris, samples.data(), samples.size(), t.decided_at, t.material, out);
}

specrel_tube descriptor(const PendingTube &p) {
    specrel_tube t;
    std::memset(&t, 0, sizeof t);
    t.vertices = p.verts.data();
    t.tri_index = p.idx.data();
    t.tri_uv = p.uv.data();
    t.tri_normal = p.normal.data();
    t.samples = p.samples.data();
    t.decided_at = p.decided.empty() ? nullptr : p.decided.data();
    t.n_vertices = (uint32_t)(p.verts.size() / 3);
    t.n_triangles = (uint32_t)(p.idx.size() / 3);
    t.n_samples = (uint32_t)p.samples.size();
    t.material = p.material;
    return t;
}

void add_tube(specrel_builder &b, uint32_t mesh_kind, const float *mp, uint32_t wl_kind, const double *wp,
              bool do_sync, float sync_t, uint32_t material, const float *decided, uint32_t n_decided) {
    Mesh m = make_mesh(mesh_kind, mp);
    std::vector<Sample> s = sample_worldline(wl_kind, wp);
    if (do_sync) sync_proper_time_at(s, sync_t);
    if (decided && n_decided != s.size()) throw Error("one decided_at entry per worldline sample");
    size_t first = b.prims.size();
    if (b.deferred) {
        if (s.size() >= 2) check_rindler(m.v, s.data(), s.size());
        PendingTube t;
        for (const V3 &v : m.v) t.verts.insert(t.verts.end(), {v.x, v.y, v.z});
        for (const Tri &tr : m.t) {
            for (int c = 0; c < 3; ++c) {
                t.idx.push_back((uint32_t)tr.idx[c]);
                t.uv.push_back(tr.uv[c][0]);
                t.uv.push_back(tr.uv[c][1]);
            }
            t.normal.insert(t.normal.end(), {tr.normal.x, tr.normal.y, tr.normal.z});
        }
        for (const Sample &x : s) t.samples.push_back(to_c(x));
        if (decided) t.decided.assign(decided, decided + n_decided);
        t.material = material;
        t.object = b.objects.size();
        b.pending.push_back(std::move(t));
    } else {
        extrude(m, s, decided, material, b.prims);
    }
    // The object's reference event: its worldline sample nearest to lab time 0.
    specrel_object o;
    std::memset(&o, 0, sizeof o);
    o.prim_first = (uint32_t)first;
    o.prim_count = (uint32_t)(b.prims.size() - first);
    size_t best = 0;
    for (size_t i = 1; i < s.size(); ++i)
        if (std::fabs(s[i].center.t) < std::fabs(s[best].center.t)) best = i;
    if (!s.empty()) {
        o.center[0] = s[best].center.t;
        o.center[1] = s[best].center.x;
        o.center[2] = s[best].center.y;
        o.center[3] = s[best].center.z;
    }
    o.material = material;
    b.objects.push_back(o);
}

specrel_gen_params defaults(const std::string &name) {
    specrel_gen_params p;
    std::memset(&p, 0, sizeof p);
    p.seed = 0x5EC12E1ull;
    if (name == "cube") {
        p.n_objects = 1;
        p.slices = 64;
        p.t_min = -20.0f;
        p.t_max = 20.0f;
    } else if (name == "large") {
        p.n_objects = 128;
        p.slices = 128;
        p.n_stars = 10000;
        p.icosphere_fraction = 1.0f;
        p.accel_fraction = 0.25f;
        p.accel_min = 0.01f;
        p.accel_max = 0.2f;
        p.t_min = -100.0f;
        p.t_max = 28.0f;
    } else if (name == "accel_clocks") {
        p.n_objects = 64;
        p.n_clocks = 16;
        p.slices = 256;
        p.n_stars = 0;
        p.icosphere_fraction = 0.5f;
        p.accel_fraction = 1.0f;
        p.accel_min = 0.05f;
        p.accel_max = 0.5f;
        p.t_min = -100.0f;
        p.t_max = 28.0f;
    } else {
        p.t_min = -18.0f;
        p.t_max = 34.0f;
    }
    return p;
}

// The large scene: a seeded field of bodies (icospheres of 320 triangles or 12-triangle cubes),
// radius U[0.5,3], centre U[-50,50]^3 at t = 0, speed U[0,0.9) in a random direction; a share
// with constant proper acceleration instead (at rest at the centre at a seeded t_turn in
// [-60, 0], accelerating along a random direction while |a (t - t_turn)| <= 3, so gamma <= 3.2),
// and a sphere of stars of radius 500 with temperatures log-uniform in [3000, 30000] K.
void gen_large(specrel_builder &b, const specrel_gen_params &p, bool clocks_scene) {
    Rng rng{p.seed};
    const uint32_t mats_checker[3] = {6, 7, 8};
    const uint32_t mats_clock[5] = {1, 2, 3, 4, 5};
    for (uint32_t k = 0; k < p.first_object + p.n_objects; ++k) {
        double radius = rng.range(0.5, 3.0);
        double c[3] = {rng.range(-50, 50), rng.range(-50, 50), rng.range(-50, 50)};
        double dir[3];
        rng.unit(dir);
        double speed = rng.range(0.0, 0.9);
        bool ico = rng.uniform() < p.icosphere_fraction;
        bool accel = rng.uniform() < p.accel_fraction;
        double a = rng.range(p.accel_min, p.accel_max);
        double pick = rng.uniform();
        if (k < p.first_object) continue;
        uint32_t material = pick < 0.5 ? mats_checker[(int)(pick * 6.0) % 3] : mats_clock[(int)((pick - 0.5) * 10.0) % 5];
        double window = 3.0;
        if (clocks_scene) {
            // Keep the Born-rigid body inside its Rindler horizon, a * r < 1 (mesh4d.rs:742-749),
            // with room for the f32 finite-difference proper acceleration of the sampler, whose
            // error grows like gamma^2: a * r <= 0.6 and gamma <= 2.24 (window |a dt| <= 2).
            radius = std::min(radius, 0.6 / a);
            window = 2.0;
        }
        float mp[3];
        uint32_t mk;
        if (ico) {
            mk = SPECREL_MESH_ICOSPHERE;
            mp[0] = (float)radius;
            mp[1] = 2.0f;
            mp[2] = 0.0f;
        } else {
            mk = SPECREL_MESH_CUBE;
            float he = (float)(radius / std::sqrt(3.0));
            mp[0] = he;
            mp[1] = he;
            mp[2] = he;
        }
        if (accel) {
            double t_turn = rng.range(-60.0, 0.0);
            (void)speed;
            double wp[12] = {c[0], c[1], c[2], dir[0], dir[1], dir[2], a, t_turn, window, (double)p.t_min,
                             (double)p.t_max, (double)p.slices};
            add_tube(b, mk, mp, SPECREL_WL_HYPERBOLIC, wp, true, 0.0f, material, nullptr, 0);
        } else {
            double wp[9] = {c[0], c[1], c[2], dir[0] * speed, dir[1] * speed, dir[2] * speed, (double)p.t_min,
                            (double)p.t_max, (double)p.slices};
            add_tube(b, mk, mp, SPECREL_WL_INERTIAL, wp, false, 0.0f, material, nullptr, 0);
        }
    }
    for (uint32_t k = 0; k < p.n_clocks; ++k) {
        // Einstein light-clock worldtubes (material 5, L = 1 light-second; scene.rs:227-236 and
        // the 0.55 x 0.625 x 0.55 pillar of scene.rs:298), inertial.
        double c[3] = {rng.range(-30, 30), rng.range(-10, 10), rng.range(-30, 30)};
        double dir[3];
        rng.unit(dir);
        double speed = rng.range(0.0, 0.8);
        float mp[3] = {0.55f, 0.625f, 0.55f};
        double wp[9] = {c[0], c[1], c[2], dir[0] * speed, dir[1] * speed, dir[2] * speed, (double)p.t_min,
                        (double)p.t_max, (double)p.slices};
        add_tube(b, SPECREL_MESH_CUBE, mp, SPECREL_WL_INERTIAL, wp, true, 0.0f, 5, nullptr, 0);
    }
    for (uint32_t k = 0; k < p.n_stars; ++k) {
        double d[3];
        rng.unit(d);
        double u = rng.uniform();
        float T = (float)(3000.0 * std::pow(10.0, u));
        double w = rng.uniform();
        float r = 500.0f;
        float I = r * r * 2e-4f * (float)std::pow(10.0, 2.0 * w);
        b.stars.push_back(star_moving(v3((float)(d[0] * 500.0), (float)(d[1] * 500.0), (float)(d[2] * 500.0)),
                                      v3(0, 0, 0), T, I));
    }
    b.light = v3(10.0f, 80.0f, -10.0f);
}

// Scene 1 of the reference: the twin paradox (scene.rs:283-379).
void gen_twin(specrel_builder &b) {
    const float t_min = -18.0f, t_max = 34.0f, t_loop = t_max - t_min, t_sync_base = -14.0f;
    {
        float mp[5] = {22.0f, 24.0f, -2.2f, 4, 4};
        double wp[9] = {0, 0, 0, 0, 0, 0, -160.0, 120.0, 1};
        add_tube(b, SPECREL_MESH_GROUND, mp, SPECREL_WL_INERTIAL, wp, false, 0, 0, nullptr, 0);
    }
    float twin[3] = {1.15f, 1.15f, 1.15f}, pillar[3] = {0.55f, 0.625f, 0.55f}, marker[3] = {0.6f, 0.6f, 0.6f};
    for (int k = -2; k <= 1; ++k) {
        float t_offset = (float)k * t_loop;
        float seg_t0 = t_min + t_offset, seg_t1 = t_max + t_offset, t_sync = t_sync_base + t_offset;
        double home[9] = {-1.2, -1.05, 6.0, 0, 0, 0, (double)seg_t0, (double)seg_t1, 2};
        add_tube(b, SPECREL_MESH_CUBE, twin, SPECREL_WL_INERTIAL, home, true, t_sync, 1, nullptr, 0);
        double home_lc[9] = {-3.2, -1.575, 6.0, 0, 0, 0, (double)seg_t0, (double)seg_t1, 2};
        add_tube(b, SPECREL_MESH_CUBE, pillar, SPECREL_WL_INERTIAL, home_lc, true, t_sync, 5, nullptr, 0);
        double trav[6] = {(double)t_offset, -1.05, 6.0, (double)seg_t0, (double)seg_t1, 120};
        add_tube(b, SPECREL_MESH_CUBE, twin, SPECREL_WL_TWIN, trav, true, t_sync, 2, nullptr, 0);
        double trav_lc[6] = {(double)t_offset, 0.725, 6.0, (double)seg_t0, (double)seg_t1, 120};
        add_tube(b, SPECREL_MESH_CUBE, pillar, SPECREL_WL_TWIN, trav_lc, true, t_sync, 5, nullptr, 0);
        double mk[9] = {9.2, -1.6, 7.8, 0, 0, 0, (double)seg_t0, (double)seg_t1, 2};
        add_tube(b, SPECREL_MESH_CUBE, marker, SPECREL_WL_INERTIAL, mk, true, t_sync, 3, nullptr, 0);
    }
    b.light = v3(2.0f, 6.5f, 2.0f);
    b.stars = starfield(v3(0, 0, 0), 90, 60.0f, 170.0f, 7);
}

// The ground plane of the demo scenes, static over [t0, t1] (one worldline step).
void demo_ground(specrel_builder &b, float half_x, float half_z, float divs_x, float divs_z, double t0, double t1) {
    float mp[5] = {half_x, half_z, -2.2f, divs_x, divs_z};
    double wp[9] = {0, 0, 0, 0, 0, 0, t0, t1, 1};
    add_tube(b, SPECREL_MESH_GROUND, mp, SPECREL_WL_INERTIAL, wp, false, 0, 0, nullptr, 0);
}

// A body at rest at p over [t0, t1] in `steps` worldline steps.
void demo_static(specrel_builder &b, const float *he, float px, float py, float pz, double t0, double t1,
                 double steps, uint32_t material) {
    double wp[9] = {(double)px, (double)py, (double)pz, 0, 0, 0, t0, t1, steps};
    add_tube(b, SPECREL_MESH_CUBE, he, SPECREL_WL_INERTIAL, wp, false, 0, material, nullptr, 0);
}

// Scene 2 of the reference: the barn and the pole (scene.rs:383-455). A 5 m barn at rest and an
// 8 m pole at 0.85c along x (contracted to 4.21 m), in loop copies k = -2..1 of 24 s.
void gen_barn(specrel_builder &b) {
    const float t_min = -12.0f, t_max = 12.0f, t_loop = t_max - t_min;
    demo_ground(b, 22.0f, 22.0f, 4, 4, -140.0, 140.0);
    float pillar[3] = {0.35f, 1.5f, 0.35f};
    for (float x_door : {-2.85f, 2.85f}) demo_static(b, pillar, x_door, -0.7f, 5.2f, -140.0, 140.0, 1, 1);
    float beam[3] = {3.2f, 0.8f, 0.6f};
    demo_static(b, beam, 0.0f, 1.6f, 5.8f, -140.0, 140.0, 1, 1);
    float wall[3] = {2.5f, 1.5f, 0.25f};
    demo_static(b, wall, 0.0f, -0.7f, 7.4f, -140.0, 140.0, 1, 6);
    float pole[3] = {4.0f, 0.55f, 0.45f};
    const V3 beta = v3(0.85f, 0.0f, 0.0f);
    for (int k = -2; k <= 1; ++k) {
        const float t_offset = (float)k * t_loop;
        const V3 pos = v3(0.0f, -0.85f, 6.0f) - beta * t_offset;
        double wp[9] = {pos.x, pos.y, pos.z, beta.x, beta.y, beta.z, (double)(t_offset - 55.0f),
                        (double)(t_offset + 55.0f), 12};
        add_tube(b, SPECREL_MESH_CUBE, pole, SPECREL_WL_INERTIAL, wp, true, t_offset, 2, nullptr, 0);
    }
    b.light = v3(0.0f, 6.0f, 1.5f);
    b.stars = starfield(v3(0, 0, 0), 90, 60.0f, 170.0f, 7);
}

// Scene 3 of the reference: the Terrell-Penrose street (scene.rs:460-571). A colonnade of clock
// cubes and light clocks at rest, loop copies (k = -2..1 of 28 s) of a cube flying at 0.85c along
// x and a uv sphere at 0.86c along -z, a cube oscillating along x and one orbiting in the xz plane.
void gen_penrose(specrel_builder &b) {
    const float t_min = -14.0f, t_max = 14.0f, t_loop = t_max - t_min;
    demo_ground(b, 20.0f, 28.0f, 4, 6, -140.0, 140.0);
    float pillar[3] = {0.95f, 0.95f, 0.95f}, lc[3] = {0.65f, 0.625f, 0.65f};
    for (int iz = -3; iz <= 3; ++iz) {
        const float z = (float)iz * 5.0f;
        for (float x : {-4.2f, 4.2f}) {
            if (iz == 0) demo_static(b, lc, x, -1.575f, z, -140.0, 140.0, 1, 5);
            else demo_static(b, pillar, x, -1.25f, z, -140.0, 140.0, 1, 1);
        }
    }
    float flyer[3] = {1.05f, 1.05f, 1.05f};
    float sphere[3] = {1.05f, 10, 16};
    const V3 beta_f = v3(0.85f, 0.0f, 0.0f), beta_s = v3(0.0f, 0.0f, -0.86f);
    for (int k = -2; k <= 1; ++k) {
        const float t_offset = (float)k * t_loop;
        const V3 pf = v3(0.0f, -0.75f, 4.5f) - beta_f * t_offset;
        double wf[9] = {pf.x, pf.y, pf.z, beta_f.x, beta_f.y, beta_f.z, (double)(t_offset - 55.0f),
                        (double)(t_offset + 55.0f), 10};
        add_tube(b, SPECREL_MESH_CUBE, flyer, SPECREL_WL_INERTIAL, wf, true, t_offset, 2, nullptr, 0);
        const V3 ps = v3(-1.8f, 0.6f, 5.0f) - beta_s * t_offset;
        double ws[9] = {ps.x, ps.y, ps.z, beta_s.x, beta_s.y, beta_s.z, (double)(t_offset - 55.0f),
                        (double)(t_offset + 55.0f), 10};
        add_tube(b, SPECREL_MESH_UV_SPHERE, sphere, SPECREL_WL_INERTIAL, ws, true, t_offset, 6, nullptr, 0);
    }
    const double omega0 = 6.283185307179586 / (double)t_loop;
    float accel[3] = {0.95f, 0.95f, 0.95f};
    double wa[8] = {0.0, 3.6, omega0, 1.1, 8.5, -110.0, 70.0, 260};
    add_tube(b, SPECREL_MESH_CUBE, accel, SPECREL_WL_OSCILLATE, wa, false, 0, 3, nullptr, 0);
    float orbit[3] = {0.75f, 0.75f, 0.75f};
    double wo[11] = {1.5, 0.5, 11.5, 3.2, omega0, 0.0, 0.0, 0.0, -110.0, 70.0, 260};
    add_tube(b, SPECREL_MESH_CUBE, orbit, SPECREL_WL_ORBIT, wo, false, 0, 4, nullptr, 0);
    b.light = v3(0.0f, 6.5f, 4.0f);
    b.stars = starfield(v3(0, 0, 0), 90, 60.0f, 170.0f, 7);
}

// Scene 4 of the reference: Einstein's train (scene.rs:576-643). Seven platform clocks at rest
// synchronised in the lab, and loop copies (k = -2..1 of 28 s) of a seven-wagon train at 0.8c
// synchronised in its own rest frame.
void gen_train(specrel_builder &b) {
    const float t_min = -14.0f, t_max = 14.0f, t_loop = t_max - t_min;
    demo_ground(b, 24.0f, 22.0f, 4, 4, -140.0, 140.0);
    float clock[3] = {0.95f, 0.95f, 0.95f};
    for (int i = -3; i <= 3; ++i) demo_static(b, clock, (float)i * 2.6f, -1.15f, 6.8f, -140.0, 140.0, 1, 1);
    float lc[3] = {0.55f, 0.625f, 0.55f};
    demo_static(b, lc, 0.0f, -1.575f, 5.1f, -140.0, 140.0, 1, 5);
    const V3 beta = v3(0.80f, 0.0f, 0.0f);
    const float gamma = 1.0f / sqrtf(1.0f - beta.x * beta.x);
    for (int k = -2; k <= 1; ++k) {
        const float t_offset = (float)k * t_loop;
        for (int i = -3; i <= 3; ++i) {
            const float x_rest = (float)i * 2.6f;
            const V3 pos = v3(x_rest / gamma, 1.05f, 6.8f) - beta * t_offset;
            const float t_sync = t_offset + gamma * beta.x * x_rest;
            double wp[9] = {pos.x, pos.y, pos.z, beta.x, beta.y, beta.z, (double)(t_offset - 60.0f),
                            (double)(t_offset + 60.0f), 10};
            add_tube(b, SPECREL_MESH_CUBE, clock, SPECREL_WL_INERTIAL, wp, true, t_sync, 2, nullptr, 0);
        }
    }
    b.light = v3(0.0f, 6.5f, 2.0f);
    b.stars = starfield(v3(0, 0, 0), 90, 60.0f, 170.0f, 7);
}

// Scene 5 of the reference: the relativistic carousel (scene.rs:648-719). A master clock and a
// light clock at rest, six cubes orbiting them at 0.8c, two turns per 48 s.
void gen_carousel(specrel_builder &b) {
    const float t_min = -24.0f, t_max = 24.0f, t_loop = t_max - t_min, t_sync = -20.0f;
    demo_ground(b, 22.0f, 24.0f, 4, 4, -140.0, 120.0);
    float master[3] = {1.05f, 1.05f, 1.05f};
    double wm[9] = {0.0, (double)-0.7f, 8.0, 0, 0, 0, -110.0, 70.0, 2};
    add_tube(b, SPECREL_MESH_CUBE, master, SPECREL_WL_INERTIAL, wm, true, t_sync, 1, nullptr, 0);
    float lc[3] = {0.55f, 0.625f, 0.55f};
    double wl[9] = {0.0, (double)0.975f, 8.0, 0, 0, 0, -110.0, 70.0, 2};
    add_tube(b, SPECREL_MESH_CUBE, lc, SPECREL_WL_INERTIAL, wl, true, t_sync, 5, nullptr, 0);
    float orbiter[3] = {0.62f, 0.62f, 0.62f};
    const double tau_c = 6.283185307179586;
    const double omega = 2.0 * tau_c / (double)t_loop, radius = 0.80 / omega;
    for (int k = 0; k < 6; ++k) {
        const double phase = (double)k * tau_c / 6.0;
        double wo[11] = {0.0, -0.6, 8.0, radius, omega, (double)t_sync, phase, 0.25, -110.0, 70.0, 320};
        add_tube(b, SPECREL_MESH_CUBE, orbiter, SPECREL_WL_ORBIT, wo, true, t_sync, k % 2 == 0 ? 2 : 3, nullptr, 0);
    }
    b.light = v3(0.0f, 6.8f, 3.0f);
    b.stars = starfield(v3(0, 0, 0), 90, 60.0f, 170.0f, 7);
}

// gpu_matches_cpu_decision_age_with_frontier's scene (gpu.rs:1156-1185).
void gen_decision_age(specrel_builder &b) {
    float gp[5] = {22.0f, 22.0f, -2.2f, 4, 4};
    double gw[9] = {0, 0, 0, 0, 0, 0, -40.0, 40.0, 1};
    add_tube(b, SPECREL_MESH_GROUND, gp, SPECREL_WL_INERTIAL, gw, false, 0, 0, nullptr, 0);
    float cube[3] = {0.6f, 0.6f, 0.6f};
    for (int k = 0; k < 4; ++k) {
        double wp[9] = {(double)(-3.0f + 2.0f * (float)k), -1.0, 4.0, 0.0, 0.0, -0.3, -20.0, 0.0, 20};
        std::vector<float> dec(21);
        for (int i = 0; i < 21; ++i) dec[i] = (float)i + (float)k;
        add_tube(b, SPECREL_MESH_CUBE, cube, SPECREL_WL_INERTIAL, wp, false, 0, 1, dec.data(), 21);
    }
    b.light = v3(0.0f, 6.0f, 2.0f);
}

}  // namespace

extern "C" {

const char *specrel_scene_last_error(void) { return g_err.c_str(); }

int specrel_gen_defaults(const char *name, specrel_gen_params *out) {
    return guarded([&] {
        *out = defaults(name ? name : "");
        return 0;
    });
}

specrel_builder *specrel_builder_new(void) { return new specrel_builder(); }
void specrel_builder_free(specrel_builder *b) { delete b; }
uint32_t specrel_builder_num_simplices(const specrel_builder *b) { return (uint32_t)b->prims.size(); }

int specrel_builder_add_tube(specrel_builder *b, uint32_t mesh_kind, const float *mp, uint32_t wl_kind,
                             const double *wp, float sync_t, uint32_t do_sync, uint32_t material,
                             const float *decided, uint32_t n_decided) {
    return guarded([&] {
        add_tube(*b, mesh_kind, mp, wl_kind, wp, do_sync != 0, sync_t, material, decided, n_decided);
        return 0;
    });
}

int specrel_builder_add_simplex(specrel_builder *b, const float verts[16], const float uv[8],
                                const float tri_bary[12], const float tau[4], const float grad[8],
                                const float uvws[2], const float normal[4], const float four_vel[4],
                                const float decided[4], uint32_t material) {
    return guarded([&] {
        if (b->deferred) throw Error("explicit simplices cannot be mixed with deferred tubes");
        V4 vv[4];
        float uvv[4][2], gg[4][2];
        V3 tb[4];
        for (int i = 0; i < 4; ++i) {
            vv[i] = v4(verts[4 * i], verts[4 * i + 1], verts[4 * i + 2], verts[4 * i + 3]);
            uvv[i][0] = uv[2 * i];
            uvv[i][1] = uv[2 * i + 1];
            gg[i][0] = grad[2 * i];
            gg[i][1] = grad[2 * i + 1];
            tb[i] = v3(tri_bary[3 * i], tri_bary[3 * i + 1], tri_bary[3 * i + 2]);
        }
        float dd[4] = {DECIDED_AT_STATIC, DECIDED_AT_STATIC, DECIDED_AT_STATIC, DECIDED_AT_STATIC};
        if (decided)
            for (int i = 0; i < 4; ++i) dd[i] = decided[i];
        specrel_geom g;
        specrel_attr a;
        Box4 box;
        bool ok = make_simplex(vv, uvv, tb, tau, gg, uvws[0], uvws[1],
                               v4(normal[0], normal[1], normal[2], normal[3]),
                               v4(four_vel[0], four_vel[1], four_vel[2], four_vel[3]), dd, material, g, a, box);
        if (ok) {
            b->prims.geoms.push_back(g);
            b->prims.attrs.push_back(a);
            b->prims.boxes.push_back(box);
        }
        return ok ? 1 : 0;
    });
}

int specrel_builder_add_star(specrel_builder *b, const float *ev, uint32_t n_events, const float beta[3],
                             float temperature, float intensity) {
    return guarded([&] {
        if (n_events <= 1) {
            b->stars.push_back(star_moving(v3(ev[0], ev[1], ev[2]), beta ? v3(beta[0], beta[1], beta[2]) : v3(0, 0, 0),
                                           temperature, intensity));
        } else {
            std::vector<V4> e(n_events);
            for (uint32_t i = 0; i < n_events; ++i) e[i] = v4(ev[4 * i], ev[4 * i + 1], ev[4 * i + 2], ev[4 * i + 3]);
            b->stars.push_back(star_along(e.data(), n_events, temperature, intensity));
        }
        return 0;
    });
}

int specrel_builder_set_light(specrel_builder *b, float x, float y, float z) {
    b->light = v3(x, y, z);
    return 0;
}

int specrel_builder_set_materials(specrel_builder *b, const specrel_material *m, uint32_t n) {
    b->materials.assign(m, m + n);
    return 0;
}

int specrel_builder_finish(specrel_builder *b, specrel_host_scene **out) {
    return guarded([&] {
        if (!b->pending.empty()) throw Error("the builder has tubes waiting for their records");
        auto s = std::make_unique<specrel_host_scene>();
        finish_into(*b, *s);
        *out = s.release();
        return 0;
    });
}

namespace {
void generate_into(specrel_builder &b, const char *name, const specrel_gen_params *pp) {
        std::string n = name ? name : "";
        specrel_gen_params p = pp ? *pp : defaults(n);
        if (n == "cube") {
            // One 12-triangle cube (half extent 0.5) at (0,0,5) on an inertial worldline, v = 0.6c
            // along +x over t in [-20, 20], 64 steps: 12 * 64 * 3 = 2304 tetrahedra.
            float he[3] = {0.5f, 0.5f, 0.5f};
            double wp[9] = {0, 0, 5, 0.6, 0, 0, (double)p.t_min, (double)p.t_max, (double)p.slices};
            add_tube(b, SPECREL_MESH_CUBE, he, SPECREL_WL_INERTIAL, wp, false, 0, 6, nullptr, 0);
            b.light = v3(2.0f, 6.5f, 2.0f);
        } else if (n == "large") {
            gen_large(b, p, false);
        } else if (n == "accel_clocks") {
            gen_large(b, p, true);
        } else if (n == "twin") {
            gen_twin(b);
        } else if (n == "decision_age") {
            gen_decision_age(b);
        } else if (n == "barn") {
            gen_barn(b);
        } else if (n == "penrose") {
            gen_penrose(b);
        } else if (n == "train") {
            gen_train(b);
        } else if (n == "carousel") {
            gen_carousel(b);
        } else {
            throw Error("unknown scene name: " + n);
        }
}
}  // namespace

int specrel_scene_generate(const char *name, const specrel_gen_params *pp, specrel_host_scene **out) {
    return guarded([&] {
        specrel_builder b;
        generate_into(b, name, pp);
        auto s = std::make_unique<specrel_host_scene>();
        finish_into(b, *s);
        *out = s.release();
        return 0;
    });
}

int specrel_scene_generate_deferred(const char *name, const specrel_gen_params *pp, specrel_builder **out) {
    return guarded([&] {
        if (!out) throw Error("out is NULL");
        auto b = std::make_unique<specrel_builder>();
        b->deferred = true;
        generate_into(*b, name, pp);
        *out = b.release();
        return 0;
    });
}

uint32_t specrel_builder_pending_tubes(const specrel_builder *b, specrel_tube *out, uint32_t cap) {
    if (!b) return 0;
    uint32_t n = (uint32_t)b->pending.size();
    for (uint32_t i = 0; out && i < n && i < cap; ++i) out[i] = descriptor(b->pending[i]);
    return n;
}

namespace {
// Appends the records of every pending tube (counts[i] of tube i, in order) and closes them.
void take_records(specrel_builder &b, const specrel_geom *g, const specrel_attr *a, const float *boxes, uint64_t n,
                  const uint64_t *counts) {
    uint64_t total = 0;
    for (size_t i = 0; i < b.pending.size(); ++i) total += counts[i];
    if (total != n) throw Error("tube record counts do not add up to n");
    if (n && (!g || !a || !boxes)) throw Error("record arrays are NULL");
    uint64_t at = 0;
    for (size_t i = 0; i < b.pending.size(); ++i) {
        specrel_object &o = b.objects[b.pending[i].object];
        o.prim_first = (uint32_t)b.prims.size();
        o.prim_count = (uint32_t)counts[i];
        for (uint64_t k = at; k < at + counts[i]; ++k) {
            b.prims.geoms.push_back(g[k]);
            b.prims.attrs.push_back(a[k]);
            const float *x = boxes + 8 * k;
            b.prims.boxes.push_back(Box4{v4(x[0], x[1], x[2], x[3]), v4(x[4], x[5], x[6], x[7])});
        }
        at += counts[i];
    }
    b.pending.clear();
}
}  // namespace

int specrel_builder_set_tube_records(specrel_builder *b, const specrel_geom *geoms, const specrel_attr *attrs,
                                     const float *boxes, uint64_t n, const uint64_t *counts) {
    return guarded([&] {
        if (!b || (!counts && !b->pending.empty())) throw Error("builder or counts is NULL");
        take_records(*b, geoms, attrs, boxes, n, counts);
        return 0;
    });
}

int specrel_builder_extrude_pending(specrel_builder *b) {
    return guarded([&] {
        if (!b) throw Error("builder is NULL");
        std::vector<uint64_t> counts;
        Prims all;
        for (const PendingTube &t : b->pending) {
            size_t before = all.size();
            extrude_tube(descriptor(t), all);
            counts.push_back(all.size() - before);
        }
        std::vector<float> boxes(8 * all.size());
        for (size_t i = 0; i < all.size(); ++i) {
            put4(&boxes[8 * i], all.boxes[i].mn);
            put4(&boxes[8 * i + 4], all.boxes[i].mx);
        }
        take_records(*b, all.geoms.data(), all.attrs.data(), boxes.data(), all.size(), counts.data());
        return 0;
    });
}

int specrel_extrude_tubes_host(const specrel_tube *tubes, uint32_t n, specrel_geom *geoms, specrel_attr *attrs,
                               float *boxes, uint64_t cap, uint64_t *n_out, uint64_t *counts) {
    return guarded([&] {
        if ((n && !tubes) || !n_out) throw Error("tubes or n_out is NULL");
        // Tubes are independent: extrude them on a pool of threads, then concatenate in order.
        std::vector<Prims> parts(n);
        std::atomic<uint32_t> next{0};
        std::vector<std::string> errs(n);
        auto work = [&] {
            for (uint32_t i; (i = next.fetch_add(1)) < n;) {
                try {
                    extrude_tube(tubes[i], parts[i]);
                } catch (const std::exception &e) {
                    errs[i] = e.what();
                }
            }
        };
        unsigned nt = std::max(1u, std::min(n, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (unsigned i = 1; i < nt; ++i) pool.emplace_back(work);
        work();
        for (auto &th : pool) th.join();
        for (uint32_t i = 0; i < n; ++i)
            if (!errs[i].empty()) throw Error("tube " + std::to_string(i) + ": " + errs[i]);
        uint64_t total = 0;
        for (const Prims &p : parts) total += p.size();
        *n_out = total;
        if ((geoms || attrs || boxes) && total > cap)
            throw Error("output buffers hold " + std::to_string(cap) + " records, " + std::to_string(total) +
                        " needed");
        uint64_t at = 0;
        for (uint32_t i = 0; i < n; ++i) {
            const Prims &p = parts[i];
            if (geoms && p.size()) std::memcpy(geoms + at, p.geoms.data(), p.size() * sizeof(specrel_geom));
            if (attrs && p.size()) std::memcpy(attrs + at, p.attrs.data(), p.size() * sizeof(specrel_attr));
            if (boxes)
                for (size_t k = 0; k < p.size(); ++k) {
                    put4(boxes + 8 * (at + k), p.boxes[k].mn);
                    put4(boxes + 8 * (at + k) + 4, p.boxes[k].mx);
                }
            if (counts) counts[i] = p.size();
            at += p.size();
        }
        return 0;
    });
}

int specrel_scene_default_camera(const char *name, const specrel_gen_params *pp, specrel_camera *out) {
    return guarded([&] {
        std::string n = name ? name : "";
        specrel_gen_params p = pp ? *pp : defaults(n);
        std::memset(out, 0, sizeof *out);
        Tetrad T = tetrad_rest();
        V4 pos = v4(0, 0, 0, 0);
        if (n == "large" || n == "accel_clocks") {
            // Observer at the origin event, boosted by rapidity 1.0 in a seeded direction.
            Rng rng{p.seed ^ 0xC0FFEEull};
            double d[3];
            rng.unit(d);
            boost_local(T, v3((float)d[0], (float)d[1], (float)d[2]));
        } else if (n == "twin") {  // DemoScene::initial_camera_pos (scene.rs:92-101)
            pos = v4(-15.0f, 2.2f, 0.0f, -3.5f);
        } else if (n == "barn") {
            pos = v4(-4.5f, 0.0f, 0.1f, -3.2f);
        } else if (n == "penrose") {
            pos = v4(0.0f, 0.0f, 0.0f, -7.5f);
        } else if (n == "train") {
            pos = v4(-3.0f, 0.0f, 0.0f, -4.5f);
        } else if (n == "carousel") {
            pos = v4(-18.0f, 0.0f, 0.8f, -2.5f);
        } else if (n == "decision_age") {
            pos = v4(0.0f, 0.0f, 0.0f, -4.0f);
        }
        put4(out->position, pos);
        for (int i = 0; i < 4; ++i) put4(out->e[i], T.e[i]);
        out->proper_time = 0.0f;
        out->fov_y_rad = 65.0f * (3.14159265358979323846f / 180.0f);
        return 0;
    });
}

int specrel_scene_get_arrays(const specrel_host_scene *s, specrel_scene_arrays *o) {
    std::memset(o, 0, sizeof *o);
    o->nodes = s->nodes.data();
    o->num_nodes = (uint32_t)s->nodes.size();
    o->num_prims = (uint32_t)s->geoms.size();
    o->geoms = s->geoms.data();
    o->attrs = s->attrs.data();
    o->order = s->order.data();
    o->materials = s->materials.data();
    o->num_materials = (uint32_t)s->materials.size();
    o->num_stars = s->num_stars;
    o->star_words = s->star_words.data();
    o->num_star_words = (uint32_t)(s->star_words.size() / 4);
    for (int i = 0; i < 4; ++i) o->light_pos[i] = s->light[i];
    return 0;
}

uint32_t specrel_scene_objects(const specrel_host_scene *s, specrel_object *out, uint32_t cap) {
    uint32_t n = (uint32_t)s->objects.size();
    if (out) std::memcpy(out, s->objects.data(), sizeof(specrel_object) * std::min(n, cap));
    return n;
}

int specrel_scene_star_params(const specrel_host_scene *s, float *ti, uint32_t n) {
    if (n > s->num_stars) return -1;
    std::memcpy(ti, s->star_ti.data(), sizeof(float) * 2 * n);
    return 0;
}

void specrel_scene_free(specrel_host_scene *s) { delete s; }

int specrel_worldline_sample(uint32_t kind, const double *wp, specrel_wl_sample *out, uint32_t cap, uint32_t *n_out) {
    return guarded([&] {
        std::vector<Sample> s = sample_worldline(kind, wp);
        if (s.size() > cap) throw Error("sample buffer too small");
        for (size_t i = 0; i < s.size(); ++i) {
            specrel_wl_sample &o = out[i];
            std::memset(&o, 0, sizeof o);
            put4(o.center, s[i].center);
            o.proper_time = s[i].proper_time;
            o.phi[0] = s[i].phi.x;
            o.phi[1] = s[i].phi.y;
            o.phi[2] = s[i].phi.z;
            o.proper_accel_local[0] = s[i].accel_local.x;
            o.proper_accel_local[1] = s[i].accel_local.y;
            o.proper_accel_local[2] = s[i].accel_local.z;
            for (int k = 0; k < 4; ++k) put4(o.e[k], s[i].tetrad.e[k]);
        }
        *n_out = (uint32_t)s.size();
        return 0;
    });
}

static std::vector<Sample> from_c(const specrel_wl_sample *s, uint32_t n) {
    std::vector<Sample> v(n);
    for (uint32_t i = 0; i < n; ++i) {
        v[i].center = v4(s[i].center[0], s[i].center[1], s[i].center[2], s[i].center[3]);
        v[i].proper_time = s[i].proper_time;
        v[i].phi = v3(s[i].phi[0], s[i].phi[1], s[i].phi[2]);
        v[i].accel_local = v3(s[i].proper_accel_local[0], s[i].proper_accel_local[1], s[i].proper_accel_local[2]);
        for (int k = 0; k < 4; ++k) v[i].tetrad.e[k] = v4(s[i].e[k][0], s[i].e[k][1], s[i].e[k][2], s[i].e[k][3]);
    }
    return v;
}

int specrel_sync_proper_time(specrel_wl_sample *s, uint32_t n, float t_sync) {
    std::vector<Sample> v = from_c(s, n);
    sync_proper_time_at(v, t_sync);
    for (uint32_t i = 0; i < n; ++i) {
        s[i].proper_time = v[i].proper_time;
        s[i].phi[0] = v[i].phi.x;
        s[i].phi[1] = v[i].phi.y;
        s[i].phi[2] = v[i].phi.z;
    }
    return 0;
}

int specrel_extrude(uint32_t mesh_kind, const float *mp, const specrel_wl_sample *s, uint32_t n, uint32_t material,
                    specrel_geom *geoms, specrel_attr *attrs, uint32_t cap, uint32_t *n_out) {
    return guarded([&] {
        Mesh m = make_mesh(mesh_kind, mp);
        Prims p;
        extrude(m, from_c(s, n), nullptr, material, p);
        if (p.size() > cap) throw Error("output buffer too small");
        std::memcpy(geoms, p.geoms.data(), p.size() * sizeof(specrel_geom));
        std::memcpy(attrs, p.attrs.data(), p.size() * sizeof(specrel_attr));
        *n_out = (uint32_t)p.size();
        return 0;
    });
}

static Tetrad load_t(const float *e) {
    Tetrad T;
    for (int k = 0; k < 4; ++k) T.e[k] = v4(e[4 * k], e[4 * k + 1], e[4 * k + 2], e[4 * k + 3]);
    return T;
}
static void store_t(float *e, const Tetrad &T) {
    for (int k = 0; k < 4; ++k) put4(e + 4 * k, T.e[k]);
}
void specrel_tetrad_rest(float e[16]) { store_t(e, tetrad_rest()); }
void specrel_tetrad_from_beta(float e[16], const float beta[3]) {
    store_t(e, tetrad_from_beta(v3(beta[0], beta[1], beta[2])));
}
void specrel_tetrad_boost_local(float e[16], const float r[3]) {
    Tetrad T = load_t(e);
    boost_local(T, v3(r[0], r[1], r[2]));
    store_t(e, T);
}
void specrel_tetrad_rotate_local(float e[16], float yaw, float pitch, float roll) {
    Tetrad T = load_t(e);
    rotate_local(T, yaw, pitch, roll);
    store_t(e, T);
}
void specrel_tetrad_boost_to(float e[16], const float u[4]) {
    Tetrad T = load_t(e);
    boost_to(T, v4(u[0], u[1], u[2], u[3]));
    store_t(e, T);
}

int specrel_frame_build(const specrel_camera *cam, const specrel_render_config *cfg, const specrel_dynamics *dyn,
                        specrel_frame *f) {
    std::memset(f, 0, sizeof *f);
    if (cfg->width == 0 || cfg->height == 0) return -2;
    float aspect = (float)cfg->width / (float)cfg->height;
    float half_h = tanf_(0.5f * cam->fov_y_rad);
    float half_w = aspect * half_h;
    for (int i = 0; i < 4; ++i) f->cam_pos[i] = cam->position[i];
    for (int k = 0; k < 4; ++k)
        for (int i = 0; i < 4; ++i) f->e[k][i] = cam->e[k][i];
    f->half_wh[0] = half_w;
    f->half_wh[1] = half_h;
    f->size[0] = cfg->width;
    f->size[1] = cfg->height;
    f->mode = cfg->mode;
    f->shadow_mode = cfg->shadow_mode;
    f->slice_mode = cfg->slice_mode;
    f->prophecy_delta = cfg->prophecy_delta;
    bool has_frontier = dyn && dyn->has_frontier;
    const float *fo = has_frontier ? dyn->frontier_origin : cam->position;
    for (int i = 0; i < 4; ++i) f->frontier_origin[i] = fo[i];
    f->frontier_tau = has_frontier ? dyn->frontier_proper_time : cam->proper_time;
    f->frontier_view_offset = 0.14f + (dyn ? dyn->frontier_lag : 0.0f);
    uint32_t nfl = dyn ? std::min<uint32_t>(dyn->num_flashes, SPECREL_MAX_FLASHES) : 0;
    f->num_flashes = nfl;
    for (uint32_t i = 0; i < nfl; ++i) {
        const specrel_flash &fl = dyn->flashes[i];
        for (int c = 0; c < 4; ++c) f->flash_at[i][c] = fl.at[c];
        f->flash_color_width[i][0] = fl.color[0];
        f->flash_color_width[i][1] = fl.color[1];
        f->flash_color_width[i][2] = fl.color[2];
        f->flash_color_width[i][3] = fl.width;
        f->flash_intensity[i] = fl.intensity;
    }
    // haze_origins (renderer.rs:253-262)
    uint32_t nh = 0;
    uint32_t nrem = dyn ? std::min<uint32_t>(dyn->num_remote, 4) : 0;
    if (cfg->slice_mode == SPECREL_VIEW_PROPHECY) {
        for (int i = 0; i < 4; ++i) f->haze_origin[nh][i] = f->frontier_origin[i];
        nh++;
    }
    if (cfg->slice_mode == SPECREL_VIEW_PROPHECY || cfg->slice_mode == SPECREL_VIEW_SIMULTANEITY) {
        for (uint32_t r = 0; r < nrem; ++r) {
            for (int i = 0; i < 4; ++i) f->haze_origin[nh][i] = dyn->remote_origins[r][i];
            nh++;
        }
    }
    f->num_haze = nh;
    return 0;
}

uint32_t specrel_common_materials(specrel_material *out, uint32_t cap) {
    std::vector<specrel_material> m = common_materials();
    uint32_t n = (uint32_t)std::min<size_t>(cap, m.size());
    std::memcpy(out, m.data(), n * sizeof(specrel_material));
    return (uint32_t)m.size();
}

void specrel_build_tables(uint32_t *out) { build_tables(out); }

}  // extern "C"

// ---------------------------------------------------------------------------------------
// Scene file (include/specrel_file.h)
// ---------------------------------------------------------------------------------------
struct specrel_scene_file {
    sr::file::Mapped m;
};

int specrel_scene_file_write(const char *path, const specrel_scene_arrays *s, const specrel_object *objects,
                             uint32_t n_objects, uint64_t bvh_layout, uint64_t bvh_edit) {
    return guarded([&] {
        if (!s) throw std::invalid_argument("scene is NULL");
        try {
            sr::file::write(path, *s, objects, n_objects, bvh_layout, bvh_edit);
        } catch (const sr::file::Error &e) {
            throw std::runtime_error(e.msg);
        }
        return 0;
    });
}

int specrel_scene_file_open(const char *path, int verify, specrel_scene_file **out, specrel_scene_arrays *arrays,
                            specrel_file_info *info) {
    return guarded([&] {
        if (!out || !arrays) throw std::invalid_argument("out or arrays is NULL");
        *out = nullptr;
        std::unique_ptr<specrel_scene_file> f(new specrel_scene_file);
        try {
            sr::file::open(path, verify != 0, f->m, *arrays);
        } catch (const sr::file::Error &e) {
            throw std::runtime_error(e.msg);
        }
        if (info) {
            std::memset(info, 0, sizeof *info);
            info->version = f->m.h.version;
            info->flags = f->m.h.flags;
            info->bvh_layout = f->m.h.bvh_layout;
            info->bvh_edit = f->m.h.bvh_edit;
            info->file_bytes = f->m.h.file_bytes;
            info->num_objects = f->m.h.num_objects;
        }
        *out = f.release();
        return 0;
    });
}

const specrel_object *specrel_scene_file_objects(const specrel_scene_file *f, uint32_t *count) {
    if (count) *count = f ? f->m.h.num_objects : 0;
    return f ? (const specrel_object *)f->m.objects : nullptr;
}

void specrel_scene_file_close(specrel_scene_file *f) { delete f; }

int specrel_build_bvh_host(const float *boxes, uint64_t n, specrel_node *nodes, uint32_t *order, uint32_t *n_nodes) {
    return guarded([&] {
        if ((n && !boxes) || !nodes || !order || !n_nodes) throw Error("boxes, nodes, order or n_nodes is NULL");
        std::vector<Aabb> bx(n);
        for (uint64_t i = 0; i < n; ++i)
            bx[i] = Box4{v4(boxes[8 * i], boxes[8 * i + 1], boxes[8 * i + 2], boxes[8 * i + 3]),
                         v4(boxes[8 * i + 4], boxes[8 * i + 5], boxes[8 * i + 6], boxes[8 * i + 7])};
        std::vector<specrel_node> out;
        std::vector<uint32_t> ord;
        build_tree(bx.data(), n, out, ord);
        std::memcpy(nodes, out.data(), out.size() * sizeof(specrel_node));
        std::memcpy(order, ord.data(), ord.size() * 4);
        *n_nodes = (uint32_t)out.size();
        return 0;
    });
}

int specrel_builder_finish_with_tree(specrel_builder *b, const specrel_node *nodes, uint32_t n_nodes,
                                     const uint32_t *order, specrel_host_scene **out) {
    return guarded([&] {
        if (!b || !out || (n_nodes && (!nodes || !order))) throw Error("builder, tree or out is NULL");
        if (!b->pending.empty()) throw Error("the builder has tubes waiting for their records");
        size_t n = b->prims.size();
        std::vector<uint8_t> seen(n, 0);
        for (size_t i = 0; i < n; ++i) {
            if (order[i] >= n || seen[order[i]]) throw Error("order is not a permutation of the records");
            seen[order[i]] = 1;
        }
        auto s = std::make_unique<specrel_host_scene>();
        s->nodes.assign(nodes, nodes + n_nodes);
        s->order.assign(order, order + n);
        finish_rest(*b, *s);
        *out = s.release();
        return 0;
    });
}
