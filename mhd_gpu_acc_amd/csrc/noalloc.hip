// mhd_advance_grids_noalloc: one step over caller-owned device grids with no
// allocation at all, for callers that cannot call hipMalloc (code under HIP graph
// capture, tight serving loops, harnesses that forbid it).
//
// The step's only device memory beyond its input and output grids is scratch: the
// 8-byte CFL accumulator and the frame of undamped psi (4 (nx + ny) doubles) that the
// closing ghost fill reads.  A run handle allocates both; here they live in one
// fixed file-scope device array, `scratch`, which each GPU holds for the life of the
// process.  Host code never passes the array's name to a kernel (on the host it is a
// stand-in, not a device address): the first call on each device looks up that
// device's address with hipGetSymbolAddress and keeps it, and every use goes through
// that address.  Every call on one GPU shares its array, so calls that overlap on one
// GPU race; the accumulator is cleared by the step's first kernel, and every frame
// entry the closing ghost fill reads is written by the same step's y sweep first.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>

#include "../../include/mhd_hip.h"
#include "launch.hpp"

namespace {

constexpr size_t SCRATCH_ACC_BYTES = 256;  // the accumulator, padded so the frame stays aligned
__device__ __align__(256) unsigned char scratch[MHD_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
struct DeviceScratch {
    unsigned char* ptr = nullptr;  // device address of `scratch` on this device
    int cus = 0;
};
DeviceScratch g_dev[MAX_DEVICES];
std::mutex g_mu;

// This device's scratch address and CU count, looked up on its first call.
int device_scratch(DeviceScratch* out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return mhd::set_error(MHD_ERR_HIP, "hipGetDevice failed");
    if (dev < 0 || dev >= MAX_DEVICES) return mhd::set_error(MHD_ERR_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    DeviceScratch& d = g_dev[dev];
    if (!d.ptr) {
        void* p = nullptr;
        if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess || !p)
            return mhd::set_error(MHD_ERR_HIP, "hipGetSymbolAddress(scratch) failed");
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
        d.cus = cus;
        d.ptr = static_cast<unsigned char*>(p);
    }
    *out = d;
    return MHD_OK;
}

}  // namespace

extern "C" int mhd_advance_grids_noalloc(const mhd_run_desc* d, double ch, const double* Uold, double* Utmp,
                                         double* Unew, double dt, double* max_speed) {
    if (!d || !Uold || !Utmp || !Unew) return mhd::set_error(MHD_ERR_INVALID, "null argument");
    if (Uold == Unew || Uold == Utmp || Utmp == Unew) return mhd::set_error(MHD_ERR_INVALID, "grids must be distinct");
    if (d->ng != 2 || d->nx < 2 || d->ny < 2) return mhd::set_error(MHD_ERR_INVALID, "bad grid size (ng must be 2)");
    if ((long long)(d->nx + 4) * (d->ny + 4) >= (1LL << 31)) return mhd::set_error(MHD_ERR_INVALID, "grid too large");
    if (d->solver < 0 || d->solver > 3 || d->limiter < 0 || d->limiter > 1)
        return mhd::set_error(MHD_ERR_INVALID, "unknown Riemann solver or limiter");
    for (int k = 0; k < 4; ++k)
        if (d->bc[k] < 0 || d->bc[k] > 3) return mhd::set_error(MHD_ERR_INVALID, "unknown boundary type");
    if (!(d->x_max > d->x_min) || !(d->y_max > d->y_min)) return mhd::set_error(MHD_ERR_INVALID, "empty domain");
    const size_t frame_bytes = 4 * ((size_t)d->nx + (size_t)d->ny) * sizeof(double);
    if (SCRATCH_ACC_BYTES + frame_bytes > (size_t)MHD_NOALLOC_SCRATCH_BYTES)
        return mhd::set_error(MHD_ERR_INVALID, "mhd_advance_grids_noalloc: nx + ny exceeds MHD_NOALLOC_MAX_NXNY "
                                               "(the library's fixed scratch); use mhd_advance_grids");
    if (max_speed && !mhd::fused_cfl())
        return mhd::set_error(MHD_ERR_INVALID, "max_speed needs the y sweep's fused CFL reduction (MHD_FUSED_CFL)");
    DeviceScratch s;
    if (int rc = device_scratch(&s)) return rc;
    const hipStream_t stream = static_cast<hipStream_t>(d->stream);
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(s.ptr);
    mhd::StepArgs a{};
    a.Uold = Uold;
    a.Utmp = Utmp;
    a.Unew = Unew;
    a.nx = d->nx;
    a.ny = d->ny;
    a.ps = (long long)(d->nx + 4) * (d->ny + 4);
    a.dx = (d->x_max - d->x_min) / static_cast<double>(d->nx);  // as mhd_run_create
    a.dy = (d->y_max - d->y_min) / static_cast<double>(d->ny);
    a.dt = dt;
    a.gamma = d->gamma;
    a.ch = ch;
    a.damp = std::exp(-dt * ch / d->cr);  // as mhd_advance (solver_cpu.cpp:264)
    a.solver = d->solver;
    a.limiter = d->limiter;
    std::memcpy(a.bc, d->bc, sizeof a.bc);
    a.seg = 0;
    a.xwg = 0;
    a.ywg = 0;
    a.xgrid = 8;  // the run handle's default launch shapes
    a.xtail = 50;
    a.addr_buf = true;
    a.cus = s.cus;
    a.frame = reinterpret_cast<double*>(s.ptr + SCRATCH_ACC_BYTES);
    a.acc = max_speed ? acc : nullptr;
    a.acc_host = nullptr;
    a.stream = stream;
    a.ev = nullptr;
    a.fill_tmp = true;  // Utmp is the caller's: all of its ghost cells are filled, as the reference's are
    hipError_t e = mhd::launch_step(a);
    if (e != hipSuccess) return mhd::set_error(MHD_ERR_HIP, hipGetErrorString(e));
    if (max_speed) {
        unsigned long long bits = 0;
        if ((e = hipMemcpyAsync(&bits, acc, sizeof bits, hipMemcpyDeviceToHost, stream)) != hipSuccess ||
            (e = hipStreamSynchronize(stream)) != hipSuccess)
            return mhd::set_error(MHD_ERR_HIP, hipGetErrorString(e));
        std::memcpy(max_speed, &bits, sizeof bits);
    }
    return MHD_OK;
}
