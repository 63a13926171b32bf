// Device self-check of arithmetic building blocks, for the test suite only (not
// part of the product library): sqrt_cr of physics.hpp against sqrt(), and FastDiv's
// shared-reciprocal quotients against division, lane by lane, on device buffers the
// caller owns.
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

// Two quotients by one denominator with FastDiv (one reciprocal), the same two with
// division, and whether FastDiv's range held for this lane's three operands.
__global__ __launch_bounds__(256) void div_pair_kernel(const double* __restrict__ a1, const double* __restrict__ a2,
                                                       const double* __restrict__ b, double* __restrict__ q,
                                                       double* __restrict__ e, int* __restrict__ ok, long long n) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        mhd::FastDiv fd;
        const mhd::Rcp r = fd.rcp(b[k]);
        q[2 * k] = fd.q(a1[k], r);
        q[2 * k + 1] = fd.q(a2[k], r);
        e[2 * k] = a1[k] / b[k];
        e[2 * k + 1] = a2[k] / b[k];
        ok[k] = (fd.lo >= -mhd::DIV_EXP && fd.hi <= mhd::DIV_EXP) ? 1 : 0;
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

// For k < n: q[2k], q[2k+1] = FastDiv's a1[k] / b[k], a2[k] / b[k] (one shared
// reciprocal); e[2k], e[2k+1] = the same by division; ok[k] = 1 when the three
// operands' frexp exponents lie in [-DIV_EXP, DIV_EXP] (0 for zero, inf, NaN).
// Null stream, waits for completion.  Returns 0, or 1 on a HIP error.
int dc_div_pair(const double* a1, const double* a2, const double* b, double* q, double* e, int* ok, long long n) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(div_pair_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, a1, a2, b, q, e, ok, n);
    if (hipGetLastError() != hipSuccess) return 1;
    return hipDeviceSynchronize() == hipSuccess ? 0 : 1;
}

}  // extern "C"
