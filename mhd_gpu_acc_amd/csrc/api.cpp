// Host layer of the C ABI (include/mhd_hip.h): run handles, state transfer,
// dt control and the driver loop, over the kernels of kernels.hip.
//
// The time-step control keeps the reference's host arithmetic: c_h ratchets to
// the largest signal speed (solver_cpu.cpp:283-312), dt = (cfl min(dx,dy)) / c_h,
// and the psi damping factor is exp(((-dt) c_h) / c_r) evaluated with the host
// libm (solver_cpu.cpp:264).  The maximum signal speed is reduced on the device
// (fused into the y sweep of the previous step, or by the stand-alone CFL kernel)
// and read back as 8 bytes per step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mhd_hip.h"
#include "launch.hpp"
#include "run.hpp"

namespace mhd {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
int hip_fail(hipError_t e, const char* what) {
    return fail(MHD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace mhd

using mhd::fail;
using mhd::g_err;
using mhd::hip_fail;

namespace {

int check_desc(const mhd_run_desc* d) {
    if (!d) return fail(MHD_ERR_INVALID, "null run description");
    if (d->ng != 2) return fail(MHD_ERR_INVALID, "ng must be 2");
    if (d->nx < 2 || d->ny < 2) return fail(MHD_ERR_INVALID, "nx and ny must be >= 2");
    if ((long long)(d->nx + 4) * (d->ny + 4) >= (1LL << 31)) return fail(MHD_ERR_INVALID, "grid too large");
    if (d->solver < 0 || d->solver > 3) return fail(MHD_ERR_INVALID, "unknown Riemann solver");
    if (d->limiter < 0 || d->limiter > 1) return fail(MHD_ERR_INVALID, "unknown limiter");
    for (int k = 0; k < 4; ++k)
        if (d->bc[k] < 0 || d->bc[k] > 3) return fail(MHD_ERR_INVALID, "unknown boundary type");
    if (!(d->x_max > d->x_min) || !(d->y_max > d->y_min)) return fail(MHD_ERR_INVALID, "empty domain");
    return MHD_OK;
}

// Host read of the device CFL maximum (a non-negative double as raw bits).  After a
// step the closing ghost fill stores it into host_acc[0] and then the step's number into
// host_acc[1]: the host polls the number instead of waiting for the stream, which saves
// the stream synchronisation's latency on every step.  The poll asks the stream every
// 1024 turns whether it failed or finished without the number arriving.  Otherwise the
// maximum is copied and the stream synchronised.
int read_acc(mhd_run* r, double* out) {
    if (r->host_acc_fresh) {
        for (unsigned spin = 1;; ++spin) {
            if (__atomic_load_n(r->host_acc + 1, __ATOMIC_ACQUIRE) == r->seq) break;
            if ((spin & 1023u) == 0) {
                const hipError_t q = hipStreamQuery(r->stream());
                if (q == hipSuccess) {
                    if (__atomic_load_n(r->host_acc + 1, __ATOMIC_ACQUIRE) == r->seq) break;
                    return fail(MHD_ERR_HIP, "compute_dt: the step finished without handing over its CFL maximum");
                }
                if (q != hipErrorNotReady) return hip_fail(q, "hipStreamQuery");
            }
            __builtin_ia32_pause();
        }
    } else {
        MHD_HIP(hipMemcpyAsync(r->host_acc, r->acc, sizeof(unsigned long long), hipMemcpyDeviceToHost, r->stream()));
        MHD_HIP(hipStreamSynchronize(r->stream()));
    }
    unsigned long long bits = __atomic_load_n(r->host_acc, __ATOMIC_ACQUIRE);
    std::memcpy(out, &bits, sizeof bits);
    return MHD_OK;
}

}  // namespace

namespace mhd {

// compute_dt on a grid: reduction on the device, ratchet and dt on the host.
int dt_from_state(mhd_run* r, const double* U, bool reuse, double* dt) {
    if (!reuse) {
        r->host_acc_fresh = false;
        MHD_HIP(hipMemsetAsync(r->acc, 0, sizeof(unsigned long long), r->stream()));
        MHD_HIP(mhd::launch_cfl(U, r->d.nx, r->d.ny, (long long)r->cells, r->d.gamma, r->ch, r->acc, r->stream()));
    }
    double mx = 0.0;
    if (int rc = read_acc(r, &mx)) return rc;
    if (reuse == false && U == r->U[0]) r->acc_valid = true;
    if (!(mx > 0.0)) return fail(MHD_ERR_NO_SIGNAL, "compute_dt: non-positive maximum wave speed.");
    r->ch = mx;
    *dt = r->d.cfl * std::min(r->dx, r->dy) / mx;
    return MHD_OK;
}

int step_grids(mhd_run* r, const double* Uold, double* Utmp, double* Unew, double dt, bool want_cfl, bool x_only,
               hipStream_t stream, bool fill_tmp) {
    mhd::StepArgs a{};
    a.Uold = Uold;
    a.Utmp = Utmp;
    a.Unew = Unew;
    a.nx = r->d.nx;
    a.ny = r->d.ny;
    a.ps = (long long)r->cells;
    a.dx = r->dx;
    a.dy = r->dy;
    a.dt = dt;
    a.gamma = r->d.gamma;
    a.ch = r->ch;
    a.damp = std::exp(-dt * r->ch / r->d.cr);
    a.solver = r->d.solver;
    a.limiter = r->d.limiter;
    std::memcpy(a.bc, r->d.bc, sizeof a.bc);
    a.seg = r->seg;
    a.xwg = r->xwg;
    a.ywg = r->ywg;
    a.xgrid = r->xgrid;
    a.xtail = r->xtail;
    a.addr_buf = r->addr_buf;
    a.cus = r->cus;
    a.frame = r->frame;
    a.acc = want_cfl ? r->acc : nullptr;
    a.acc_host = want_cfl ? r->host_acc_dev : nullptr;
    a.seq = (want_cfl && !x_only) ? ++r->seq : 0;
    a.stream = stream ? stream : r->stream();
    a.ev = nullptr;
    a.x_only = x_only;
    a.fill_tmp = fill_tmp;
    if (x_only) {
        a.acc = nullptr;
        a.acc_host = nullptr;
    }
    if (r->profiling && !x_only && !stream) {
        const size_t base = r->events.size();
        for (int k = 0; k < 5; ++k) {
            hipEvent_t ev;
            MHD_HIP(hipEventCreate(&ev));
            r->events.push_back(ev);
        }
        a.ev = r->events.data() + base;
        r->prof_steps += 1;
    }
    r->host_acc_fresh = false;
    MHD_HIP(mhd::launch_step(a));
    r->host_acc_fresh = want_cfl && !x_only && mhd::fused_cfl();
    return MHD_OK;
}

}  // namespace mhd

using mhd::dt_from_state;
using mhd::step_grids;

namespace mhd {
int set_error(int code, const char* msg) { return fail(code, msg); }
}  // namespace mhd

extern "C" {

const char* mhd_last_error(void) { return g_err.c_str(); }

int mhd_device_count(int* n) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
    return MHD_OK;
}

int mhd_run_create(const mhd_run_desc* desc, mhd_run** out) {
    if (!out) return fail(MHD_ERR_INVALID, "null output handle");
    *out = nullptr;
    if (int rc = check_desc(desc)) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(MHD_ERR_NOT_BUILT, "no HIP device");
    mhd_run* r = new mhd_run();
    r->d = *desc;
    r->tnx = desc->nx + 4;
    r->tny = desc->ny + 4;
    r->cells = (size_t)r->tnx * r->tny;
    r->dx = (desc->x_max - desc->x_min) / static_cast<double>(desc->nx);  // grid_cpu.hpp:25-26
    r->dy = (desc->y_max - desc->y_min) / static_cast<double>(desc->ny);
    const char* seg = std::getenv("MHD_YSEG");
    r->seg = seg ? std::atoi(seg) : 0;
    const char* xwg = std::getenv("MHD_XWG");
    r->xwg = xwg ? std::atoi(xwg) : 0;
    const char* ywg = std::getenv("MHD_YWG");
    r->ywg = ywg ? std::atoi(ywg) : 0;
    const char* xgrid = std::getenv("MHD_XGRID");
    r->xgrid = xgrid ? std::atoi(xgrid) : 8;  // persistent x sweep, 8 blocks per CU (measured best)
    const char* xtail = std::getenv("MHD_XTAIL");  // per mille of x segments in one-segment blocks
    r->xtail = xtail ? std::min(std::max(std::atoi(xtail), 0), 1000) : 50;  // 50: measured best at 4096^2
    const char* addr = std::getenv("MHD_ADDR");  // "flat": the sweeps' path for grids above 4 GiB
    r->addr_buf = !(addr && std::strcmp(addr, "flat") == 0);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (hipDeviceGetAttribute(&r->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || r->cus < 1)
        r->cus = 1;
    auto cleanup = [&](hipError_t e, const char* what) {
        mhd_run_destroy(r);
        return hip_fail(e, what);
    };
    for (int k = 0; k < 3; ++k) {
        hipError_t e = hipMalloc(&r->U[k], 9 * r->cells * sizeof(double));
        if (e != hipSuccess) return cleanup(e, "hipMalloc(grid)");
        e = hipMemsetAsync(r->U[k], 0, 9 * r->cells * sizeof(double), r->stream());
        if (e != hipSuccess) return cleanup(e, "hipMemset(grid)");
    }
    hipError_t e = hipMalloc(&r->frame, 4 * (size_t)(desc->nx + desc->ny) * sizeof(double));
    if (e != hipSuccess) return cleanup(e, "hipMalloc(frame)");
    if ((e = hipMalloc(&r->acc, sizeof(unsigned long long))) != hipSuccess) return cleanup(e, "hipMalloc(acc)");
    if ((e = hipHostMalloc(&r->host_acc, 2 * sizeof(unsigned long long), hipHostMallocMapped | hipHostMallocCoherent)) !=
        hipSuccess)
        return cleanup(e, "hipHostMalloc");
    if ((e = hipHostGetDevicePointer((void**)&r->host_acc_dev, r->host_acc, 0)) != hipSuccess)
        return cleanup(e, "hipHostGetDevicePointer");
    r->host_acc[0] = 0;
    r->host_acc[1] = 0;
    if ((e = hipEventCreateWithFlags(&r->order, hipEventDisableTiming)) != hipSuccess)
        return cleanup(e, "hipEventCreate");
    if ((e = hipStreamSynchronize(r->stream())) != hipSuccess) return cleanup(e, "hipStreamSynchronize");
    *out = r;
    return MHD_OK;
}

int mhd_run_destroy(mhd_run* r) {
    if (!r) return MHD_OK;
    (void)hipStreamSynchronize(r->stream());
    for (hipEvent_t ev : r->events) (void)hipEventDestroy(ev);
    if (r->order) (void)hipEventDestroy(r->order);
    for (double* p : r->U)
        if (p) (void)hipFree(p);
    if (r->frame) (void)hipFree(r->frame);
    if (r->acc) (void)hipFree(r->acc);
    if (r->host_acc) (void)hipHostFree(r->host_acc);
    delete r;
    return MHD_OK;
}

int mhd_run_set_stream(mhd_run* r, void* stream) {
    if (!r) return fail(MHD_ERR_INVALID, "null run");
    if (stream == r->d.stream) return MHD_OK;
    // The new stream waits for what is already queued on the old one (no host wait).
    MHD_HIP(hipEventRecord(r->order, r->stream()));
    MHD_HIP(hipStreamWaitEvent(static_cast<hipStream_t>(stream), r->order, 0));
    r->d.stream = stream;
    return MHD_OK;
}

size_t mhd_plane_stride(const mhd_run* r) { return r ? r->cells : 0; }
double* mhd_state_ptr(mhd_run* r) { return r ? r->U[0] : nullptr; }

int mhd_set_state(mhd_run* r, const double* src, int layout) {
    if (!r || !src) return fail(MHD_ERR_INVALID, "null argument");
    const size_t n = r->cells;
    if (layout == 0) {
        std::vector<double> soa(9 * n);
        for (size_t c = 0; c < n; ++c)
            for (int f = 0; f < 9; ++f) soa[f * n + c] = src[9 * c + f];
        MHD_HIP(hipMemcpy(r->U[0], soa.data(), 9 * n * sizeof(double), hipMemcpyHostToDevice));
    } else if (layout == 1) {
        MHD_HIP(hipMemcpy(r->U[0], src, 9 * n * sizeof(double), hipMemcpyHostToDevice));
    } else if (layout == 2) {
        MHD_HIP(hipMemcpyAsync(r->U[0], src, 9 * n * sizeof(double), hipMemcpyDeviceToDevice, r->stream()));
    } else {
        return fail(MHD_ERR_INVALID, "unknown layout");
    }
    // Utmp/Unew start as copies (scripts/cpu/main_cpu.cpp:173-174); only their
    // ghost cells are ever read before being written (Dirichlet sides).
    MHD_HIP(hipMemcpyAsync(r->U[1], r->U[0], 9 * n * sizeof(double), hipMemcpyDeviceToDevice, r->stream()));
    MHD_HIP(hipMemcpyAsync(r->U[2], r->U[0], 9 * n * sizeof(double), hipMemcpyDeviceToDevice, r->stream()));
    MHD_HIP(hipStreamSynchronize(r->stream()));
    r->acc_valid = false;
    return MHD_OK;
}

int mhd_run_init_state(mhd_run* r, int case_id, const mhd_ic_params* p) {
    if (!r) return fail(MHD_ERR_INVALID, "null run");
    if (!mhd::ic_on_device(case_id))
        return fail(MHD_ERR_INVALID, "this case has no device initial state (its host state calls the C library's "
                                     "sin/tanh/exp): use mhd_init_state_host and mhd_set_state");
    if (case_id == MHD_CASE_OT_FAMILY && !p) return fail(MHD_ERR_INVALID, "the OT family needs its parameters");
    mhd::IcRun R{};
    for (int g = 0; g < 3; ++g) R.U[g] = r->U[g];  // Utmp/Unew start as copies (main_cpu.cpp:173-174)
    R.ngrids = 3;
    R.case_id = case_id;
    R.tnx = r->tnx;
    R.tny = r->tny;
    R.ps = (long long)r->cells;
    R.x_min = r->d.x_min;
    R.y_min = r->d.y_min;
    R.dx = (r->d.x_max - r->d.x_min) / static_cast<double>(r->d.nx);  // as the host initialiser
    R.dy = (r->d.y_max - r->d.y_min) / static_cast<double>(r->d.ny);
    if (p) R.p = *p;
    MHD_HIP(mhd::launch_ic_one(R, r->stream()));
    MHD_HIP(hipStreamSynchronize(r->stream()));
    r->acc_valid = false;
    return MHD_OK;
}

int mhd_init_state_device(int case_id, const mhd_case_config* gm, const mhd_ic_params* p, double* soa, void* stream) {
    if (!gm || !soa || gm->nx < 1 || gm->ny < 1 || gm->ng != 2) return fail(MHD_ERR_INVALID, "bad geometry");
    if (!mhd::ic_on_device(case_id))
        return fail(MHD_ERR_INVALID, "this case has no device initial state (its host state calls the C library's "
                                     "sin/tanh/exp): use mhd_init_state_host");
    if (case_id == MHD_CASE_OT_FAMILY && !p) return fail(MHD_ERR_INVALID, "the OT family needs its parameters");
    mhd::IcRun R{};
    R.U[0] = soa;
    R.ngrids = 1;
    R.case_id = case_id;
    R.tnx = gm->nx + 4;
    R.tny = gm->ny + 4;
    R.ps = (long long)R.tnx * R.tny;
    R.x_min = gm->x_min;
    R.y_min = gm->y_min;
    R.dx = (gm->x_max - gm->x_min) / static_cast<double>(gm->nx);
    R.dy = (gm->y_max - gm->y_min) / static_cast<double>(gm->ny);
    if (p) R.p = *p;
    MHD_HIP(mhd::launch_ic_one(R, static_cast<hipStream_t>(stream)));
    return MHD_OK;
}

int mhd_get_state(mhd_run* r, double* dst, int layout) {
    if (!r || !dst) return fail(MHD_ERR_INVALID, "null argument");
    const size_t n = r->cells;
    MHD_HIP(hipStreamSynchronize(r->stream()));
    if (layout == 0) {
        std::vector<double> soa(9 * n);
        MHD_HIP(hipMemcpy(soa.data(), r->U[0], 9 * n * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t c = 0; c < n; ++c)
            for (int f = 0; f < 9; ++f) dst[9 * c + f] = soa[f * n + c];
    } else if (layout == 1) {
        MHD_HIP(hipMemcpy(dst, r->U[0], 9 * n * sizeof(double), hipMemcpyDeviceToHost));
    } else if (layout == 2) {
        MHD_HIP(hipMemcpyAsync(dst, r->U[0], 9 * n * sizeof(double), hipMemcpyDeviceToDevice, r->stream()));
        MHD_HIP(hipStreamSynchronize(r->stream()));
    } else {
        return fail(MHD_ERR_INVALID, "unknown layout");
    }
    return MHD_OK;
}

int mhd_set_ch(mhd_run* r, double ch) {
    if (!r) return fail(MHD_ERR_INVALID, "null run");
    r->ch = ch;
    r->acc_valid = false;
    return MHD_OK;
}

int mhd_get_ch(const mhd_run* r, double* ch) {
    if (!r || !ch) return fail(MHD_ERR_INVALID, "null argument");
    *ch = r->ch;
    return MHD_OK;
}

int mhd_compute_dt(mhd_run* r, double* dt) {
    if (!r || !dt) return fail(MHD_ERR_INVALID, "null argument");
    // After a step the y sweep has already reduced max(c_h, signal speed) over the
    // new state with the c_h a fresh compute_dt would use; a repeated call is
    // idempotent (max over cells of fmax(s, M) is M once c_h = M).
    return dt_from_state(r, r->U[0], r->acc_valid, dt);
}

int mhd_advance(mhd_run* r, double dt) {
    if (!r) return fail(MHD_ERR_INVALID, "null run");
    // U* is the run's own: only the ghost rows the y sweep reads are needed
    if (int rc = step_grids(r, r->U[0], r->U[1], r->U[2], dt, true, false, nullptr, false)) return rc;
    std::swap(r->U[0], r->U[2]);
    r->acc_valid = mhd::fused_cfl();
    return MHD_OK;
}

int mhd_step_to(mhd_run* r, double t_end, int max_steps, double* t_out, int* steps_out, double* dts, double* chs) {
    if (!r) return fail(MHD_ERR_INVALID, "null run");
    double t = 0.0;
    int steps = 0, rc = MHD_OK;
    while (t < t_end && steps < max_steps) {  // main_cpu.cpp:196
        double dt_raw = 0.0;
        if ((rc = mhd_compute_dt(r, &dt_raw)) != MHD_OK) break;
        const double dt = std::min(dt_raw, t_end - t);
        if (!std::isfinite(dt) || dt <= 0.0) { rc = fail(MHD_ERR_DT, "dt non-finite or <= 0"); break; }
        if (dt < 1.0e-8 * t_end) { rc = fail(MHD_ERR_DT_FLOOR, "dt below 1e-8 t_end: numerical blow-up"); break; }
        if ((rc = mhd_advance(r, dt)) != MHD_OK) break;
        if (dts) dts[steps] = dt;
        if (chs) chs[steps] = r->ch;
        t += dt;
        ++steps;
    }
    if (t_out) *t_out = t;
    if (steps_out) *steps_out = steps;
    return rc;
}

int mhd_apply_boundary(mhd_run* r) {
    if (!r) return fail(MHD_ERR_INVALID, "null run");
    // In place: Dirichlet sides read the grid's own (unchanged) ghosts.
    MHD_HIP(mhd::launch_bc(r->U[0], r->U[0], r->d.nx, r->d.ny, (long long)r->cells, r->d.bc, r->stream()));
    return MHD_OK;
}

int mhd_compute_dt_grid(mhd_run* r, const double* U, double* dt) {
    if (!r || !U || !dt) return fail(MHD_ERR_INVALID, "null argument");
    return dt_from_state(r, U, false, dt);
}

int mhd_advance_grids(mhd_run* r, const double* Uold, double* Utmp, double* Unew, double dt) {
    if (!r || !Uold || !Utmp || !Unew) return fail(MHD_ERR_INVALID, "null argument");
    if (Uold == Unew || Uold == Utmp || Utmp == Unew) return fail(MHD_ERR_INVALID, "grids must be distinct");
    r->acc_valid = false;
    return step_grids(r, Uold, Utmp, Unew, dt);
}

int mhd_profile(mhd_run* r, int enable) {
    if (!r) return fail(MHD_ERR_INVALID, "null run");
    MHD_HIP(hipStreamSynchronize(r->stream()));
    for (hipEvent_t ev : r->events) (void)hipEventDestroy(ev);
    r->events.clear();
    r->prof_steps = 0;
    r->profiling = enable != 0;
    return MHD_OK;
}

int mhd_profile_read(mhd_run* r, double* ms4, int* steps) {
    if (!r || !ms4) return fail(MHD_ERR_INVALID, "null argument");
    MHD_HIP(hipStreamSynchronize(r->stream()));
    for (int k = 0; k < 4; ++k) ms4[k] = 0.0;
    for (size_t s = 0; s + 5 <= r->events.size(); s += 5)
        for (int k = 0; k < 4; ++k) {
            float ms = 0.0f;
            MHD_HIP(hipEventElapsedTime(&ms, r->events[s + k], r->events[s + k + 1]));
            ms4[k] += ms;
        }
    if (steps) *steps = r->prof_steps;
    return MHD_OK;
}

int mhd_probe_bandwidth(const double* a, double* b, long long bytes, int reps, double* tbps3) {
    if (!a || !b || !tbps3) return fail(MHD_ERR_INVALID, "null argument");
    const char* pa = reinterpret_cast<const char*>(a);
    const char* pb = reinterpret_cast<const char*>(b);
    const bool overlap = bytes > 0 && pa < pb + bytes && pb < pa + bytes;
    if (overlap || bytes < (1LL << 20) || bytes % 16 != 0 || reps < 1)
        return fail(MHD_ERR_INVALID, "disjoint buffers of at least 1 MiB (a multiple of 16 bytes) and reps >= 1");
    int dev = 0, cus = 1;
    MHD_HIP(hipGetDevice(&dev));
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    MHD_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(MHD_ERR_HIP, "hipEventCreate failed");
    }
    int rc = MHD_OK;
    // write, read, then copy: b holds a copy of a on return
    const int order[3] = {2, 1, 0};
    for (int k : order) {
        hipError_t e = mhd::launch_bandwidth(k, a, b, bytes, cus, nullptr);  // warm-up
        if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
        for (int i = 0; i < reps && e == hipSuccess; ++i) e = mhd::launch_bandwidth(k, a, b, bytes, cus, nullptr);
        if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
            rc = fail(MHD_ERR_HIP, hipGetErrorString(e));
            break;
        }
        const double moved = (k == 0 ? 2.0 : 1.0) * (double)bytes;
        tbps3[k] = moved * reps / (ms * 1e-3) / 1e12;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int mhd_sync(mhd_run* r) {
    if (!r) return fail(MHD_ERR_INVALID, "null run");
    MHD_HIP(hipStreamSynchronize(r->stream()));
    return MHD_OK;
}

}  // extern "C"
