// Device evaluation of the f32 transcendental helpers the renderer's kernels call (core/minkowski.h:
// each the double routine of the device math library rounded once to float), over caller arrays,
// for the test that sweeps them against the host definitions. A diagnostics library of its own;
// the renderer does not link it.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../core/minkowski.h"
#include "../hip/device.hpp"

namespace {

__global__ void k_math(int fn, const float *a, const float *b, float *out, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        float r;
        switch (fn) {
            case 0: r = sr::expf_(a[i]); break;
            case 1: r = sr::sinf_(a[i]); break;
            case 2: r = sr::cosf_(a[i]); break;
            case 3: r = sr::atan2f_(a[i], b[i]); break;
            default: r = sr::acosf_(a[i]); break;
        }
        out[i] = r;
    }
}

// shade_sky (the renderer's, with its f32 fast decision of the grid lines) for n directions.
__global__ void k_sky(const float4 *dirs, float *out, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const float4 d = dirs[i];
        const sr::V3 c = dev::shade_sky(sr::V4{d.x, d.y, d.z, d.w});
        out[3 * i] = c.x, out[3 * i + 1] = c.y, out[3 * i + 2] = c.z;
    }
}

}  // namespace

extern "C" {

// The sky colour of n directions (t, x, y, z floats each) into out (3 floats each). 0 on success.
int specrel_skycheck(const float *dirs, float *out, uint64_t n) {
    if (!dirs || !out) return -2;
    if (n == 0) return 0;
    float *d = nullptr;
    if (hipMalloc(&d, n * 7 * sizeof(float)) != hipSuccess) return -3;
    int rc = 0;
    if (hipMemcpy(d, dirs, n * 16, hipMemcpyHostToDevice) != hipSuccess) rc = -4;
    if (!rc) {
        k_sky<<<1024, 256>>>((const float4 *)d, d + 4 * n, n);
        if (hipGetLastError() != hipSuccess || hipMemcpy(out, d + 4 * n, n * 12, hipMemcpyDeviceToHost) != hipSuccess)
            rc = -4;
    }
    (void)hipFree(d);
    return rc;
}

// fn: 0 exp, 1 sin, 2 cos, 3 atan2 (y = a, x = b), 4 acos; host arrays of n floats. 0 on success.
int specrel_mathcheck(int fn, const float *a, const float *b, float *out, uint64_t n) {
    if (fn < 0 || fn > 4 || !a || !out || (fn == 3 && !b)) return -2;
    if (n == 0) return 0;
    float *d = nullptr;
    const size_t bytes = n * sizeof(float);
    if (hipMalloc(&d, 3 * bytes) != hipSuccess) return -3;
    int rc = 0;
    if (hipMemcpy(d, a, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        (fn == 3 && hipMemcpy(d + n, b, bytes, hipMemcpyHostToDevice) != hipSuccess))
        rc = -4;
    if (!rc) {
        k_math<<<1024, 256>>>(fn, d, d + n, d + 2 * n, n);
        if (hipGetLastError() != hipSuccess || hipMemcpy(out, d + 2 * n, bytes, hipMemcpyDeviceToHost) != hipSuccess)
            rc = -4;
    }
    (void)hipFree(d);
    return rc;
}

}  // extern "C"
