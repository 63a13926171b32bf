// Initial states on the device for the cases of ic.hpp (Brio-Wu, shock-bubble,
// shock-cloud, the batch Orszag-Tang family): one thread per cell, ghost cells
// included, written straight into the SoA planes of up to three grids.  It replaces
// initialise_grid + upload_from_aos (src/init.cpp:34-43,
// head/gpu/grid_gpu.cuh:139-189) for those cases, and a whole batch is initialised
// by one launch from a few hundred bytes of parameters per run.  The cell
// arithmetic is ic.hpp, the source the host initialiser runs, so the state is
// bitwise the host state.
#include <hip/hip_runtime.h>

#include "ic.hpp"
#include "launch.hpp"
#include "sweeps.hpp"

#pragma clang fp contract(off)

namespace mhd {

bool ic_on_device(int case_id) {
    return case_id == MHD_CASE_SHOCK_BUBBLE || case_id == MHD_CASE_BRIO_WU || case_id == MHD_CASE_BRIO_WU_T ||
           case_id == MHD_CASE_SHOCK_CLOUD || case_id == MHD_CASE_OT_FAMILY;
}

namespace {

__device__ void ic_fill(const IcRun& R, long long c) {
    const int j = (int)(c / R.tnx), i = (int)(c % R.tnx);
    // cell centres, ghosts included (init.cpp:9-15), with the host's expression
    const double x = R.x_min + ((double)(i - NG) + 0.5) * R.dx;
    const double y = R.y_min + ((double)(j - NG) + 0.5) * R.dy;
    mhd_ic::State s;
    switch (R.case_id) {
        case MHD_CASE_SHOCK_BUBBLE: s = mhd_ic::shock_bubble(x, y); break;
        case MHD_CASE_BRIO_WU:
        case MHD_CASE_BRIO_WU_T: s = mhd_ic::brio_wu(x); break;
        case MHD_CASE_SHOCK_CLOUD: s = mhd_ic::shock_cloud(x, y, i, j, R.p.seed); break;
        default: s = mhd_ic::ot_family(x, y, R.p); break;
    }
    for (int g = 0; g < R.ngrids; ++g) {
        double* U = R.U[g];
#pragma unroll
        for (int f = 0; f < 9; ++f) U[f * R.ps + c] = s.q[f];
    }
}

__global__ __launch_bounds__(256) void ic_kernel_one(IcRun R) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < R.ps) ic_fill(R, c);
}

// Cells of all runs in one flat range; run r owns [runs[r].first, runs[r].first + ps).
__global__ __launch_bounds__(256) void ic_kernel_batch(const IcRun* __restrict__ runs, int nruns, long long total) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= total) return;
    int lo = 0, hi = nruns - 1;  // last run with first <= q
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (runs[mid].first <= q) lo = mid;
        else hi = mid - 1;
    }
    ic_fill(runs[lo], q - runs[lo].first);
}

}  // namespace

hipError_t launch_ic_one(const IcRun& R, hipStream_t s) {
    const long long nb = (R.ps + 255) / 256;
    hipLaunchKernelGGL(ic_kernel_one, dim3((unsigned)nb), dim3(256), 0, s, R);
    return hipGetLastError();
}

hipError_t launch_ic_batch(const IcRun* d_runs, int nruns, long long total, hipStream_t s) {
    const long long nb = (total + 255) / 256;
    hipLaunchKernelGGL(ic_kernel_batch, dim3((unsigned)nb), dim3(256), 0, s, d_runs, nruns, total);
    return hipGetLastError();
}

}  // namespace mhd
