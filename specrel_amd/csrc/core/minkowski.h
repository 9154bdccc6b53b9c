// Minkowski-space arithmetic shared by the scene producer (host C++) and the HIP kernels.
//
// Every expression keeps the reference's evaluation order, because results must match the
// CPU renderer bit for bit (compiled with -ffp-contract=off, no fast-math):
//   Vec3::dot          ((x*x' + y*y') + z*z')                    minkowski.rs:23-25
//   Vec4::minkowski_dot ((-t*t' + x*x') + y*y') + z*z'          minkowski.rs:182-184
//   Vec4::euclidean_dot ((t*t' + x*x') + y*y') + z*z'           minkowski.rs:198-200
//   division by a scalar = multiply by its reciprocal           minkowski.rs:124-129, 345-350
//   normalized() guards len > 1e-12                              minkowski.rs:47-54
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SR_HD __host__ __device__ __forceinline__
#else
#define SR_HD inline
#endif

namespace sr {

// f32 transcendentals: evaluated in double and rounded once to float. The same definition is
// used on the host (glibc) and on the device (ocml), so both give the correctly rounded value
// except where the double result sits within a double ulp of a float rounding boundary.
SR_HD float expf_(float x) { return (float)exp((double)x); }
SR_HD float sinf_(float x) { return (float)sin((double)x); }
SR_HD float cosf_(float x) { return (float)cos((double)x); }
SR_HD float atan2f_(float y, float x) { return (float)atan2((double)y, (double)x); }
SR_HD float acosf_(float x) { return (float)acos((double)x); }
SR_HD float tanf_(float x) { return (float)tan((double)x); }
SR_HD float coshf_(float x) { return (float)cosh((double)x); }
SR_HD float sinhf_(float x) { return (float)sinh((double)x); }
SR_HD float powf_(float x, float y) { return (float)pow((double)x, (double)y); }

// f32::max / f32::min are IEEE maxNum/minNum (a NaN operand yields the other one).
SR_HD float fmax_(float a, float b) { return fmaxf(a, b); }
SR_HD float fmin_(float a, float b) { return fminf(a, b); }
// f32::clamp: NaN passes through (core::f32::clamp).
SR_HD float clamp_(float x, float lo, float hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
// f32::rem_euclid: r = self % rhs (exact fmod); r < 0 ? r + |rhs| : r.
SR_HD float rem_euclid_(float x, float m) {
    float r = fmodf(x, m);
    return r < 0.0f ? r + fabsf(m) : r;
}
// f32::round: half away from zero.
SR_HD float round_(float x) { return roundf(x); }
// `as i32`: saturating, NaN -> 0.
SR_HD int32_t f2i_sat(float x) {
    if (x != x) return 0;
    if (x >= 2147483647.0f) return 2147483647;
    if (x <= -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)x;
}
// `as usize`: saturating, negative and NaN -> 0.
SR_HD uint64_t f2u_sat(float x) {
    if (!(x > 0.0f)) return 0;
    if (x >= 18446744073709551616.0f) return ~(uint64_t)0;
    return (uint64_t)x;
}

struct V3 {
    float x, y, z;
};
struct V4 {
    float t, x, y, z;
};

SR_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
SR_HD V4 v4(float t, float x, float y, float z) { return V4{t, x, y, z}; }

SR_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
SR_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
SR_HD V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
SR_HD V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
SR_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
SR_HD V3 &operator+=(V3 &a, V3 b) {
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
    return a;
}
SR_HD V3 &operator-=(V3 &a, V3 b) {
    a.x -= b.x;
    a.y -= b.y;
    a.z -= b.z;
    return a;
}
SR_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SR_HD V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
SR_HD float length(V3 a) { return sqrtf(dot(a, a)); }
SR_HD V3 normalized(V3 a) {
    float len = length(a);
    if (len > 1e-12f) return a * (1.0f / len);
    return V3{0.0f, 0.0f, 0.0f};
}
SR_HD V3 div(V3 a, float s) { return a * (1.0f / s); }
SR_HD V3 clamp3(V3 a, float lo, float hi) {
    return V3{clamp_(a.x, lo, hi), clamp_(a.y, lo, hi), clamp_(a.z, lo, hi)};
}

SR_HD V4 operator+(V4 a, V4 b) { return V4{a.t + b.t, a.x + b.x, a.y + b.y, a.z + b.z}; }
SR_HD V4 operator-(V4 a, V4 b) { return V4{a.t - b.t, a.x - b.x, a.y - b.y, a.z - b.z}; }
SR_HD V4 operator*(V4 a, float s) { return V4{a.t * s, a.x * s, a.y * s, a.z * s}; }
SR_HD V4 operator-(V4 a) { return V4{-a.t, -a.x, -a.y, -a.z}; }
SR_HD V4 &operator+=(V4 &a, V4 b) {
    a.t += b.t;
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
    return a;
}
SR_HD V4 &operator-=(V4 &a, V4 b) {
    a.t -= b.t;
    a.x -= b.x;
    a.y -= b.y;
    a.z -= b.z;
    return a;
}
SR_HD V4 div(V4 a, float s) { return a * (1.0f / s); }
SR_HD V3 spatial(V4 a) { return V3{a.x, a.y, a.z}; }
SR_HD V4 from_ts(float t, V3 s) { return V4{t, s.x, s.y, s.z}; }
SR_HD V4 splat4(float v) { return V4{v, v, v, v}; }
SR_HD float mdot(V4 a, V4 b) { return -a.t * b.t + a.x * b.x + a.y * b.y + a.z * b.z; }
SR_HD float edot(V4 a, V4 b) { return a.t * b.t + a.x * b.x + a.y * b.y + a.z * b.z; }
SR_HD V4 vmin(V4 a, V4 b) {
    return V4{fmin_(a.t, b.t), fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)};
}
SR_HD V4 vmax(V4 a, V4 b) {
    return V4{fmax_(a.t, b.t), fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)};
}
SR_HD float axis(V4 a, int i) { return i == 0 ? a.t : i == 1 ? a.x : i == 2 ? a.y : a.z; }

// Vec4::cross4 (minkowski.rs:247-262).
SR_HD V4 cross4(V4 u, V4 v, V4 w) {
    float m_xy = v.x * w.y - v.y * w.x;
    float m_xz = v.x * w.z - v.z * w.x;
    float m_yz = v.y * w.z - v.z * w.y;
    float m_tx = v.t * w.x - v.x * w.t;
    float m_ty = v.t * w.y - v.y * w.t;
    float m_tz = v.t * w.z - v.z * w.t;
    float nt = u.x * m_yz - u.y * m_xz + u.z * m_xy;
    float nx = -(u.t * m_yz - u.y * m_tz + u.z * m_ty);
    float ny = u.t * m_xz - u.x * m_tz + u.z * m_tx;
    float nz = -(u.t * m_xy - u.x * m_ty + u.y * m_tx);
    return V4{nt, nx, ny, nz};
}

// four_velocity_from_beta (minkowski.rs:266-270).
SR_HD V4 four_velocity_from_beta(V3 beta) {
    float b2 = fmin_(dot(beta, beta), 0.999999f);
    float gamma = 1.0f / sqrtf(1.0f - b2);
    return from_ts(gamma, beta * gamma);
}
// beta_from_four_velocity (minkowski.rs:274-280).
SR_HD V3 beta_from_four_velocity(V4 u) {
    if (fabsf(u.t) > 1e-12f) return div(spatial(u), u.t);
    return V3{0.0f, 0.0f, 0.0f};
}
// project_to_rest_space / rest_frame_cos (minkowski.rs:364-380).
SR_HD V4 project_to_rest_space(V4 v, V4 u) { return v + u * mdot(v, u); }
SR_HD float rest_frame_cos(V4 v, V4 n, V4 u) {
    V4 vp = project_to_rest_space(v, u);
    float len_sq = mdot(vp, vp);
    if (len_sq > 1e-12f) return clamp_(mdot(vp, n) / sqrtf(len_sq), -1.0f, 1.0f);
    return 0.0f;
}
// lorentz_boost_u1_to_u2 (minkowski.rs:385-392).
SR_HD V4 lorentz_boost_u1_to_u2(V4 v, V4 u1, V4 u2) {
    float u1_dot_u2 = mdot(u1, u2);
    float denom = 1.0f - u1_dot_u2;
    V4 sum = u1 + u2;
    float sum_dot_v = mdot(sum, v);
    float u1_dot_v = mdot(u1, v);
    return v + sum * (sum_dot_v / denom) - u2 * (2.0f * u1_dot_v);
}

// Tetrad (minkowski.rs:399-609).
struct Tetrad {
    V4 e[4];
};
SR_HD Tetrad tetrad_rest() {
    return Tetrad{{V4{1, 0, 0, 0}, V4{0, 1, 0, 0}, V4{0, 0, 1, 0}, V4{0, 0, 0, 1}}};
}
SR_HD V4 local_spatial_to_world(const Tetrad &T, V3 l) {
    return T.e[1] * l.x + T.e[2] * l.y + T.e[3] * l.z;
}
SR_HD V4 local_to_world(const Tetrad &T, V4 l) {
    return T.e[0] * l.t + T.e[1] * l.x + T.e[2] * l.y + T.e[3] * l.z;
}
SR_HD V4 world_to_local(const Tetrad &T, V4 v) {
    return V4{-mdot(v, T.e[0]), mdot(v, T.e[1]), mdot(v, T.e[2]), mdot(v, T.e[3])};
}
SR_HD V3 world_to_local_spatial(const Tetrad &T, V4 v) {
    return V3{mdot(v, T.e[1]), mdot(v, T.e[2]), mdot(v, T.e[3])};
}
SR_HD V4 past_light_ray_direction(const Tetrad &T, V3 dir_local) {
    V3 n = normalized(dir_local);
    return -T.e[0] + T.e[1] * n.x + T.e[2] * n.y + T.e[3] * n.z;
}
SR_HD V4 future_light_ray_direction(const Tetrad &T, V3 dir_local) {
    V3 n = normalized(dir_local);
    return T.e[0] + T.e[1] * n.x + T.e[2] * n.y + T.e[3] * n.z;
}
SR_HD void reorthonormalize(Tetrad &T) {
    float n0 = sqrtf(fmax_(-mdot(T.e[0], T.e[0]), 1e-12f));
    float sign = T.e[0].t >= 0.0f ? 1.0f : -1.0f;
    T.e[0] = T.e[0] * (sign / n0);
    for (int a = 1; a <= 3; ++a) {
        V4 v = T.e[a];
        v += T.e[0] * mdot(v, T.e[0]);
        for (int b = 1; b < a; ++b) v -= T.e[b] * mdot(v, T.e[b]);
        float na = sqrtf(fmax_(mdot(v, v), 1e-12f));
        T.e[a] = v * (1.0f / na);
    }
}
SR_HD void boost_to(Tetrad &T, V4 new_u0) {
    float norm_sq = mdot(new_u0, new_u0);
    V4 u2;
    if (norm_sq < -1e-12f) {
        float inv = 1.0f / sqrtf(-norm_sq);
        float sign = new_u0.t >= 0.0f ? 1.0f : -1.0f;
        u2 = new_u0 * (sign * inv);
    } else {
        u2 = V4{1, 0, 0, 0};
    }
    V4 u1 = T.e[0];
    T.e[0] = u2;
    for (int a = 1; a <= 3; ++a) T.e[a] = lorentz_boost_u1_to_u2(T.e[a], u1, u2);
    reorthonormalize(T);
}
SR_HD Tetrad tetrad_from_beta(V3 beta) {
    Tetrad T = tetrad_rest();
    boost_to(T, four_velocity_from_beta(beta));
    return T;
}
SR_HD void boost_local(Tetrad &T, V3 d_rapidity) {
    float eta = length(d_rapidity);
    if (eta < 1e-8f) return;
    V3 dir = div(d_rapidity, eta);
    V4 s = local_spatial_to_world(T, dir);
    float ch = coshf_(eta);
    float sh = sinhf_(eta);
    boost_to(T, T.e[0] * ch + s * sh);
}
SR_HD void rotate_local(Tetrad &T, float yaw, float pitch, float roll) {
    if (fabsf(yaw) > 1e-8f) {
        float s = sinf_(yaw), c = cosf_(yaw);
        V4 e1 = T.e[1], e3 = T.e[3];
        T.e[1] = e1 * c - e3 * s;
        T.e[3] = e1 * s + e3 * c;
    }
    if (fabsf(pitch) > 1e-8f) {
        float s = sinf_(pitch), c = cosf_(pitch);
        V4 e2 = T.e[2], e3 = T.e[3];
        T.e[2] = e2 * c - e3 * s;
        T.e[3] = e2 * s + e3 * c;
    }
    if (fabsf(roll) > 1e-8f) {
        float s = sinf_(roll), c = cosf_(roll);
        V4 e1 = T.e[1], e2 = T.e[2];
        T.e[1] = e1 * c - e2 * s;
        T.e[2] = e1 * s + e2 * c;
    }
    reorthonormalize(T);
}

}  // namespace sr
