// Device evaluation of the f32 transcendental helpers the renderer's kernels call (core/minkowski.h:
// each the double routine of the device math library rounded once to float), over caller arrays,
// for the test that sweeps them against the host definitions. A diagnostics library of its own;
// the renderer does not link it.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../core/minkowski.h"

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

}  // namespace

extern "C" {

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
