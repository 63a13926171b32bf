// Private host-side state of a run handle (mhd_run of the C ABI), shared by the
// host layer (api.cpp) and the adjoint driver (tape.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mhd_hip.h"

struct mhd_run {
    mhd_run_desc d;
    int tnx, tny;
    size_t cells;  // (nx+4)(ny+4): plane stride of every grid
    double dx, dy;
    double* U[3] = {nullptr, nullptr, nullptr};  // current, intermediate (U*), next
    double* frame = nullptr;
    unsigned long long* acc = nullptr;
    unsigned long long* host_acc = nullptr;      // pinned, coherent: [0] CFL maximum, [1] step number
    unsigned long long* host_acc_dev = nullptr;  // its device address
    bool host_acc_fresh = false;  // the last step's ghost fill stores acc into host_acc
    unsigned long long seq = 0;   // number of the last step that hands its maximum over
    double ch = 0.0;                         // c_h, starts at 0 (physics.hpp:41)
    bool acc_valid = false;                  // acc holds the CFL max of U[0] with the current ch
    int seg = 0;
    int xwg = 0;
    int ywg = 0;
    int xgrid = 0;
    int xtail = 50;  // per mille of x segments in one-segment blocks (MHD_XTAIL)
    bool addr_buf = true;
    int cus = 1;
    bool profiling = false;
    std::vector<hipEvent_t> events;  // 5 per profiled step
    hipEvent_t order = nullptr;      // orders a stream change (mhd_run_set_stream)
    int prof_steps = 0;
    hipStream_t stream() const { return static_cast<hipStream_t>(d.stream); }
};

namespace mhd {

// Record msg for mhd_last_error() and return code.
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
// compute_dt on grid U (reuse: the fused CFL maximum of the last step is current).
int dt_from_state(mhd_run* r, const double* U, bool reuse, double* dt);
// One advance_second_order from Uold through Utmp into Unew with the run's c_h.
// want_cfl: the y sweep also reduces Unew's CFL maximum into the run's accumulator
// (the next compute_dt); a replay with recorded dt / c_h skips it.  x_only: only the x
// sweep and the ghost fill of Utmp (a replay that needs U* of its last step, not U_next).
// stream: where to launch (null: the run's stream; a replay on another stream records no
// profiling events).  The launches use the run's frame buffer, so two steps of one run
// must not be in flight on different streams at once.
// fill_tmp: fill all of Utmp's ghost cells (caller-owned grids, which may read any of
// them); else only the interior columns of its ghost rows, which the y sweep and the
// reverse y pass read (all of them through a Dirichlet bottom or top side).
int step_grids(mhd_run* r, const double* Uold, double* Utmp, double* Unew, double dt, bool want_cfl = true,
               bool x_only = false, hipStream_t stream = nullptr, bool fill_tmp = true);

}  // namespace mhd

#define MHD_HIP(call)                                          \
    do {                                                       \
        hipError_t e_ = (call);                                \
        if (e_ != hipSuccess) return mhd::hip_fail(e_, #call); \
    } while (0)
