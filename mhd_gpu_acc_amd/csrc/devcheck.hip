// Device self-check of arithmetic building blocks, for the test suite only (not
// part of the product library): sqrt_cr of physics.hpp against sqrt(), lane by lane,
// on device buffers the caller owns.
#include <hip/hip_runtime.h>

#include "physics.hpp"

namespace {

__global__ __launch_bounds__(256) void sqrt_pair_kernel(const double* __restrict__ x, double* __restrict__ a,
                                                        double* __restrict__ b, long long n) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        a[k] = mhd::sqrt_cr(x[k]);
        b[k] = sqrt(x[k]);
    }
}

}  // namespace

extern "C" {

// a[k] = sqrt_cr(x[k]), b[k] = sqrt(x[k]) for k < n on the null stream; waits for
// completion.  Returns 0, or 1 on a HIP error.
int dc_sqrt_pair(const double* x, double* a, double* b, long long n) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(sqrt_pair_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, x, a, b, n);
    if (hipGetLastError() != hipSuccess) return 1;
    return hipDeviceSynchronize() == hipSuccess ? 0 : 1;
}

}  // extern "C"
