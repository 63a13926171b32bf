// Host-side launch interface of the step kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/mhd_hip.h"

namespace mhd {

struct StepArgs {
    const double* Uold;
    double* Utmp;
    double* Unew;
    int nx, ny;
    long long ps;            // plane stride in doubles
    double dx, dy, dt;
    double gamma, ch;
    double damp;             // exp(-dt ch / cr), computed on the host with libm
    int solver, limiter;
    int bc[4];
    int seg;                 // y-sweep rows per lane segment (0 = chosen from the grid size)
    int xwg;                 // x-sweep workgroup width: 64, 128 or 256 (0 = default)
    int ywg;                 // y-sweep workgroup width: 64, 128 or 256 (0 = default)
    int xgrid;               // x-sweep blocks per CU (persistent), 0 = one block per segment
    int xtail;               // per mille of the x sweep's segments in one-segment blocks after
                             // the persistent ones (with xgrid > 0)
    bool addr_buf;           // sweeps address through buffer resources where the grid allows (else flat)
    int cus;                 // compute units of the device
    double* frame;           // undamped psi of the interior frame (2*2*(nx+ny) doubles)
    unsigned long long* acc; // max signal speed accumulator (bit pattern of a double >= 0)
    unsigned long long* acc_host;  // null, or the device address of pinned host memory that
                                   // the last ghost fill copies acc into ([0]), then seq ([1])
    unsigned long long seq;        // this step's number, stored after the maximum
    hipStream_t stream;
    hipEvent_t* ev;          // null, or 5 events recorded around the 4 launches
    bool x_only;             // only the x sweep and its ghost fill (Utmp); Unew is not formed
    bool fill_tmp;           // fill all of Utmp's ghost cells (for callers that read Utmp); else
                             // only the ghost rows the y sweep reads, written by the x sweep
};

hipError_t launch_step(const StepArgs& a);

// One reverse step (adjoint.hip): lam holds the adjoint of the step's output on
// entry and of its input (Uold) on return; lam_tmp receives the adjoint of Utmp.
struct AdjArgs {
    const double* Uold;      // step input
    const double* Utmp;      // its x-swept intermediate (ghosts filled)
    double* lam;
    double* lam_tmp;         // ghost columns must be zero
    int nx, ny;
    long long ps;
    double dx, dy, dt;
    double gamma, ch;
    double damp;             // the step's psi damping factor
    int solver, limiter;     // solver 0 (HLL) or 2 (HLLD)
    int bc[4];
    bool addr_buf;           // address through buffer resources where the grid allows (else flat)
    hipStream_t stream;
};
hipError_t launch_adjoint(const AdjArgs& a);
// Counts the non-finite values of the 9 planes of a grid into *count (device).
hipError_t launch_count_nonfinite(const double* U, long long n, long long ps, unsigned int* count, hipStream_t s);
// Records a failure for mhd_last_error() on this thread and returns code.
int set_error(int code, const char* msg);
// Whether the y sweep reduces the next step's CFL maximum (else compute_dt launches it).
bool fused_cfl();
hipError_t launch_cfl(const double* U, int nx, int ny, long long ps, double gamma, double ch,
                      unsigned long long* acc, hipStream_t s);
hipError_t launch_bc(double* U, const double* prev, int nx, int ny, long long ps, const int* bc, hipStream_t s);
// Streaming kernel of known traffic on `bytes` (a multiple of 16) bytes: kind 0 copies a
// into b, 1 reads a, 2 writes b.
hipError_t launch_bandwidth(int kind, const double* a, double* b, long long bytes, int cus, hipStream_t s);
// *acc (zeroed by the caller) gets the bits of max |B| over the interior cells of U.
hipError_t launch_max_b(const double* U, int nx, int ny, long long ps, unsigned long long* acc, hipStream_t s);

// A grid set to initialise on the device (init.hip): the first ngrids of U get
// the same state.  Cells are numbered row-major over the tnx x tny grid, ghosts
// included; in a batch launch this run owns [first, first + ps).
struct IcRun {
    double* U[3];
    int ngrids;
    int case_id;
    int tnx, tny;
    long long ps;
    long long first;
    double x_min, y_min, dx, dy;
    mhd_ic_params p;
};
// Whether case_id has a device initial state (the cases of ic.hpp).
bool ic_on_device(int case_id);
hipError_t launch_ic_one(const IcRun& R, hipStream_t s);
hipError_t launch_ic_batch(const IcRun* d_runs, int nruns, long long total, hipStream_t s);

}  // namespace mhd
