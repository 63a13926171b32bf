// Checkpointed discrete adjoint of a fixed number of steps (mhd_tape_* of the C ABI).
//
// Forward: the run advances nsteps steps exactly as mhd_step_to would (compute_dt
// with the c_h ratchet, then advance), the dt and c_h of every step are recorded,
// and the state before steps 0, K, 2K, ... is copied to a device checkpoint.  The
// last segment steps through the segment buffers themselves.
// Backward: for each segment, last first, the segment is recomputed from its
// checkpoint with the recorded dt and c_h (the same kernels, so the same bits),
// keeping every step's input U and intermediate U* -- except the last segment on
// the first backward after a forward, which the forward left there; then the
// reverse steps of adjoint.hip run from the segment's last step to its first.
// With two or more segments there are two sets of segment buffers: a second stream
// recomputes the next segment into one set while the run's stream reverses the
// current one from the other, so the recompute fills what the reverse sweeps leave
// idle (their tails, the launch gaps).  Events order the two streams; results are
// the same bits either way.
// Device memory: ceil(nsteps/K) checkpoints + (2K+1) segment grids per set + 2
// adjoint grids.  When the second set does not fit, the tape runs with one set.
//
// The adjoint treats the whole update as differentiable: the reference's revert of
// a cell whose update is non-finite (solver_cpu.cpp:386-392) cannot occur on a
// finite trajectory, and the backward checks that the recorded trajectory's first
// and last states are finite.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "launch.hpp"
#include "run.hpp"

using mhd::fail;

struct mhd_tape {
    mhd_run* run = nullptr;
    int nsteps = 0, K = 0, nck = 0;
    std::vector<double> dts, chs;
    bool recorded = false;
    bool last_kept = false;    // set 0 holds the last segment as the forward stepped it
    std::vector<double*> ck;   // checkpoints
    struct Set {
        std::vector<double*> S;  // segment states, K+1
        std::vector<double*> T;  // segment intermediates, K
    } set[2];                    // set[1] stays empty when the tape runs with one set
    bool two = false;            // two sets and a recompute stream
    hipStream_t side = nullptr;  // recompute stream (default priority)
    hipEvent_t ev_join = nullptr, ev_ready[2] = {nullptr, nullptr}, ev_free[2] = {nullptr, nullptr};
    double* lam[2] = {nullptr, nullptr};
    unsigned int* count = nullptr;
    unsigned int* host_count = nullptr;  // pinned
};

namespace {

size_t grid_bytes(const mhd_run* r) { return 9 * r->cells * sizeof(double); }

int alloc_grid(mhd_run* r, double** p) {
    MHD_HIP(hipMalloc(p, grid_bytes(r)));
    return MHD_OK;
}

// Device SoA copy of a state or adjoint given in one of the mhd_set_state layouts.
int upload(mhd_run* r, const double* src, int layout, double* dst) {
    const size_t n = r->cells;
    if (layout == 0) {
        std::vector<double> soa(9 * n);
        for (size_t c = 0; c < n; ++c)
            for (int f = 0; f < 9; ++f) soa[f * n + c] = src[9 * c + f];
        MHD_HIP(hipMemcpyAsync(dst, soa.data(), grid_bytes(r), hipMemcpyHostToDevice, r->stream()));
        MHD_HIP(hipStreamSynchronize(r->stream()));
    } else if (layout == 1) {
        MHD_HIP(hipMemcpyAsync(dst, src, grid_bytes(r), hipMemcpyHostToDevice, r->stream()));
        MHD_HIP(hipStreamSynchronize(r->stream()));
    } else if (layout == 2) {
        MHD_HIP(hipMemcpyAsync(dst, src, grid_bytes(r), hipMemcpyDeviceToDevice, r->stream()));
    } else {
        return fail(MHD_ERR_INVALID, "unknown layout");
    }
    return MHD_OK;
}

int download(mhd_run* r, const double* src, int layout, double* dst) {
    const size_t n = r->cells;
    if (layout == 0) {
        std::vector<double> soa(9 * n);
        MHD_HIP(hipMemcpyAsync(soa.data(), src, grid_bytes(r), hipMemcpyDeviceToHost, r->stream()));
        MHD_HIP(hipStreamSynchronize(r->stream()));
        for (size_t c = 0; c < n; ++c)
            for (int f = 0; f < 9; ++f) dst[9 * c + f] = soa[f * n + c];
    } else if (layout == 1) {
        MHD_HIP(hipMemcpyAsync(dst, src, grid_bytes(r), hipMemcpyDeviceToHost, r->stream()));
        MHD_HIP(hipStreamSynchronize(r->stream()));
    } else if (layout == 2) {
        MHD_HIP(hipMemcpyAsync(dst, src, grid_bytes(r), hipMemcpyDeviceToDevice, r->stream()));
        MHD_HIP(hipStreamSynchronize(r->stream()));
    } else {
        return fail(MHD_ERR_INVALID, "unknown layout");
    }
    return MHD_OK;
}

int check_adjoint_run(const mhd_run* r) {
    if (r->d.solver < MHD_HLL || r->d.solver > MHD_FORCE) return fail(MHD_ERR_INVALID, "adjoint: unknown Riemann solver");
    return MHD_OK;
}

int nonfinite(mhd_tape* t, const double* U, unsigned int* out) {
    mhd_run* r = t->run;
    MHD_HIP(hipMemsetAsync(t->count, 0, sizeof(unsigned int), r->stream()));
    MHD_HIP(mhd::launch_count_nonfinite(U, (long long)r->cells, (long long)r->cells, t->count, r->stream()));
    MHD_HIP(hipMemcpyAsync(t->host_count, t->count, sizeof(unsigned int), hipMemcpyDeviceToHost, r->stream()));
    MHD_HIP(hipStreamSynchronize(r->stream()));
    *out = *t->host_count;
    return MHD_OK;
}

int reverse_step(mhd_run* r, const double* Uold, const double* Utmp, double dt, double ch, double* lam,
                 double* lam_tmp) {
    mhd::AdjArgs a{};
    a.Uold = Uold;
    a.Utmp = Utmp;
    a.lam = lam;
    a.lam_tmp = lam_tmp;
    a.nx = r->d.nx;
    a.ny = r->d.ny;
    a.ps = (long long)r->cells;
    a.dx = r->dx;
    a.dy = r->dy;
    a.dt = dt;
    a.gamma = r->d.gamma;
    a.ch = ch;
    a.damp = std::exp(-dt * ch / r->d.cr);  // as the forward step (api.cpp step_grids)
    a.solver = r->d.solver;
    a.limiter = r->d.limiter;
    std::memcpy(a.bc, r->d.bc, sizeof a.bc);
    a.addr_buf = r->addr_buf;
    a.stream = r->stream();
    MHD_HIP(mhd::launch_adjoint(a));
    return MHD_OK;
}

}  // namespace

extern "C" {

int mhd_tape_destroy(mhd_tape* t) {
    if (!t) return MHD_OK;
    if (t->run) (void)hipStreamSynchronize(t->run->stream());
    if (t->side) (void)hipStreamSynchronize(t->side);
    for (double* p : t->ck) if (p) (void)hipFree(p);
    for (auto& st : t->set) {
        for (double* p : st.S) if (p) (void)hipFree(p);
        for (double* p : st.T) if (p) (void)hipFree(p);
    }
    for (double* p : t->lam) if (p) (void)hipFree(p);
    for (hipEvent_t e : {t->ev_join, t->ev_ready[0], t->ev_ready[1], t->ev_free[0], t->ev_free[1]})
        if (e) (void)hipEventDestroy(e);
    if (t->side) (void)hipStreamDestroy(t->side);
    if (t->count) (void)hipFree(t->count);
    if (t->host_count) (void)hipHostFree(t->host_count);
    delete t;
    return MHD_OK;
}

int mhd_tape_create(mhd_run* r, int nsteps, int K, mhd_tape** out) {
    if (!out) return fail(MHD_ERR_INVALID, "null output handle");
    *out = nullptr;
    if (!r) return fail(MHD_ERR_INVALID, "null run");
    if (nsteps < 1 || K < 1) return fail(MHD_ERR_INVALID, "nsteps and K must be >= 1");
    if (int rc = check_adjoint_run(r)) return rc;
    mhd_tape* t = new mhd_tape();
    t->run = r;
    t->nsteps = nsteps;
    t->K = std::min(K, nsteps);
    t->nck = (nsteps + t->K - 1) / t->K;
    t->ck.assign(t->nck, nullptr);
    t->set[0].S.assign(t->K + 1, nullptr);
    t->set[0].T.assign(t->K, nullptr);
    int rc = MHD_OK;
    for (auto& p : t->ck) if (!rc) rc = alloc_grid(r, &p);
    for (auto& p : t->set[0].S) if (!rc) rc = alloc_grid(r, &p);
    for (auto& p : t->set[0].T) if (!rc) rc = alloc_grid(r, &p);
    for (auto& p : t->lam) if (!rc) rc = alloc_grid(r, &p);
    if (!rc) {
        hipError_t e = hipEventCreateWithFlags(&t->ev_join, hipEventDisableTiming);
        for (int b = 0; b < 2 && e == hipSuccess; ++b) {
            e = hipEventCreateWithFlags(&t->ev_ready[b], hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&t->ev_free[b], hipEventDisableTiming);
        }
        if (e != hipSuccess) rc = mhd::hip_fail(e, "tape events");
    }
    // MHD_TAPE_SETS=1 (a measurement knob): one set, no recompute stream.
    const char* sets = std::getenv("MHD_TAPE_SETS");
    if (!rc && t->nck > 1 && !(sets && std::strcmp(sets, "1") == 0)) {
        // The second set and its stream are an optimisation: without the memory the tape
        // runs with one set.
        auto& st = t->set[1];
        st.S.assign(t->K + 1, nullptr);
        st.T.assign(t->K, nullptr);
        bool ok = hipStreamCreateWithFlags(&t->side, hipStreamNonBlocking) == hipSuccess;
        for (auto& p : st.S) ok = ok && hipMalloc(&p, grid_bytes(r)) == hipSuccess;
        for (auto& p : st.T) ok = ok && hipMalloc(&p, grid_bytes(r)) == hipSuccess;
        if (!ok) {
            (void)hipGetLastError();
            for (double* p : st.S) if (p) (void)hipFree(p);
            for (double* p : st.T) if (p) (void)hipFree(p);
            st.S.clear();
            st.T.clear();
            if (t->side) (void)hipStreamDestroy(t->side);
            t->side = nullptr;
        }
        t->two = ok;
    }
    if (!rc) {
        hipError_t e = hipMalloc(&t->count, sizeof(unsigned int));
        if (e == hipSuccess) e = hipHostMalloc(&t->host_count, sizeof(unsigned int), hipHostMallocDefault);
        // the y pass never writes the ghost columns of the U* adjoint: they stay zero (with a
        // Dirichlet side the transpose of the copies from U* sets them every step)
        if (e == hipSuccess) e = hipMemsetAsync(t->lam[1], 0, grid_bytes(r), r->stream());
        if (e == hipSuccess) e = hipStreamSynchronize(r->stream());
        if (e != hipSuccess) rc = mhd::hip_fail(e, "tape allocation");
    }
    if (rc) {
        mhd_tape_destroy(t);
        return rc;
    }
    *out = t;
    return MHD_OK;
}

int mhd_tape_forward(mhd_tape* t, double* dts, double* chs) {
    if (!t) return fail(MHD_ERR_INVALID, "null tape");
    mhd_run* r = t->run;
    t->recorded = false;
    t->last_kept = false;
    t->dts.assign(t->nsteps, 0.0);
    t->chs.assign(t->nsteps, 0.0);
    const int s_last = (t->nck - 1) * t->K;  // first step of the last segment
    const std::vector<double*>& S = t->set[0].S;
    const std::vector<double*>& T = t->set[0].T;
    for (int s = 0; s < t->nsteps; ++s) {
        if (s % t->K == 0)
            MHD_HIP(hipMemcpyAsync(t->ck[s / t->K], r->U[0], grid_bytes(r), hipMemcpyDeviceToDevice, r->stream()));
        double dt = 0.0;
        if (s < s_last) {
            if (int rc = mhd_compute_dt(r, &dt)) return rc;
            if (!std::isfinite(dt) || dt <= 0.0) return fail(MHD_ERR_DT, "dt non-finite or <= 0");
            if (int rc = mhd_advance(r, dt)) return rc;
        } else {
            // The last segment steps through the segment buffers the backward reverses,
            // with the arithmetic of mhd_compute_dt / mhd_advance, so it needs no recompute.
            const int q = s - s_last;
            if (q == 0)
                MHD_HIP(hipMemcpyAsync(S[0], r->U[0], grid_bytes(r), hipMemcpyDeviceToDevice, r->stream()));
            const bool reuse = q == 0 ? r->acc_valid : mhd::fused_cfl();
            if (int rc = mhd::dt_from_state(r, S[q], reuse, &dt)) return rc;
            if (!std::isfinite(dt) || dt <= 0.0) return fail(MHD_ERR_DT, "dt non-finite or <= 0");
            if (int rc = mhd::step_grids(r, S[q], T[q], S[q + 1], dt, true, false, nullptr, false)) return rc;
        }
        t->dts[s] = dt;
        t->chs[s] = r->ch;
        if (dts) dts[s] = dt;
        if (chs) chs[s] = r->ch;
    }
    const int m = t->nsteps - s_last;
    MHD_HIP(hipMemcpyAsync(r->U[0], S[m], grid_bytes(r), hipMemcpyDeviceToDevice, r->stream()));
    MHD_HIP(hipStreamSynchronize(r->stream()));
    r->acc_valid = mhd::fused_cfl();  // acc holds the CFL maximum of S[m], now U[0]
    t->recorded = true;
    t->last_kept = true;
    return MHD_OK;
}

int mhd_tape_backward(mhd_tape* t, const double* lam_final, double* lam0, int layout) {
    if (!t || !lam_final || !lam0) return fail(MHD_ERR_INVALID, "null argument");
    if (!t->recorded) return fail(MHD_ERR_INVALID, "tape: no recorded forward");
    mhd_run* r = t->run;
    // The recompute below sets the run's c_h step by step; the run itself (its state,
    // c_h ratchet and CFL accumulator) is left as the caller had it, which need not be
    // this tape's end (another tape's forward may have advanced the run since).
    const double ch_run = r->ch;
    const bool acc_valid = r->acc_valid, host_acc_fresh = r->host_acc_fresh;
    unsigned int bad = 0;
    if (int rc = nonfinite(t, t->ck[0], &bad)) return rc;
    if (bad) return fail(MHD_ERR_INVALID, "adjoint: the initial state is not finite");
    if (int rc = upload(r, lam_final, layout, t->lam[0])) return rc;
    int rc = MHD_OK;
    const bool kept = t->last_kept;
    const bool two = t->two;
    hipStream_t M = r->stream(), R = two ? t->side : M;
    // Segment k lives in set (nck-1-k) % 2 (with one set, set 0): the last segment in
    // set 0, where the forward left it.
    auto set_of = [&](int k) { return two ? (t->nck - 1 - k) % 2 : 0; };
    bool freed[2] = {false, false};  // ev_free[b] recorded: reverse steps read set b
    auto recompute = [&](int k) -> int {
        const int b = set_of(k);
        const std::vector<double*>& S = t->set[b].S;
        const std::vector<double*>& T = t->set[b].T;
        const int s0 = k * t->K, m = std::min(t->K, t->nsteps - s0);
        if (two && freed[b]) MHD_HIP(hipStreamWaitEvent(R, t->ev_free[b], 0));
        MHD_HIP(hipMemcpyAsync(S[0], t->ck[k], grid_bytes(r), hipMemcpyDeviceToDevice, R));
        // The reverse steps read U and U* of each step (U* at its interior columns only, so
        // the x sweep's pushed ghost rows stand in for U*'s ghost fill where the bottom and
        // top sides allow); the state after the segment is only needed for the last
        // segment's finiteness check, so elsewhere the last step stops after its x sweep.
        for (int q = 0; q < m; ++q) {
            r->ch = t->chs[s0 + q];
            const bool x_only = q == m - 1 && k != t->nck - 1;
            if (int e = mhd::step_grids(r, S[q], T[q], S[q + 1], t->dts[s0 + q], false, x_only, R, false))
                return e;
        }
        if (two) MHD_HIP(hipEventRecord(t->ev_ready[b], R));
        return MHD_OK;
    };
    if (two) {  // the recompute stream starts after what the run's stream holds
        MHD_HIP(hipEventRecord(t->ev_join, M));
        MHD_HIP(hipStreamWaitEvent(R, t->ev_join, 0));
    }
    t->last_kept = false;
    for (int k = t->nck - 1; k >= 0 && !rc; --k) {
        const int s0 = k * t->K, m = std::min(t->K, t->nsteps - s0), b = set_of(k);
        const bool here = kept && k == t->nck - 1;
        if (!here && (!two || k == t->nck - 1)) rc = recompute(k);
        if (!rc && two && k > 0) rc = recompute(k - 1);  // overlaps the reverse steps of k
        if (rc) break;
        if (two && !here) {
            const hipError_t e = hipStreamWaitEvent(M, t->ev_ready[b], 0);
            if (e != hipSuccess) { rc = mhd::hip_fail(e, "hipStreamWaitEvent"); break; }
        }
        const std::vector<double*>& S = t->set[b].S;
        const std::vector<double*>& T = t->set[b].T;
        if (k == t->nck - 1) {
            if ((rc = nonfinite(t, S[m], &bad))) break;
            if (bad) { rc = fail(MHD_ERR_INVALID, "adjoint: the recorded trajectory is not finite"); break; }
        }
        for (int q = m - 1; q >= 0 && !rc; --q)
            rc = reverse_step(r, S[q], T[q], t->dts[s0 + q], t->chs[s0 + q], t->lam[0], t->lam[1]);
        if (!rc && two) {
            const hipError_t e = hipEventRecord(t->ev_free[b], M);
            if (e != hipSuccess) rc = mhd::hip_fail(e, "hipEventRecord");
            freed[b] = true;
        }
    }
    // Set 0 still holds the last segment when no other segment went through it.
    if (!rc) t->last_kept = t->nck == 1 || (two && t->nck == 2);
    if (two) {  // whatever the recompute stream still holds precedes the run's next work
        const hipError_t e1 = hipEventRecord(t->ev_join, R);
        const hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(M, t->ev_join, 0) : e1;
        if (e2 != hipSuccess && !rc) rc = mhd::hip_fail(e2, "tape stream join");
    }
    // The recompute writes only the segment buffers and skips the CFL reduction, so the
    // run's U[0], accumulator and pinned copy are untouched: restore the flags as well.
    r->ch = ch_run;
    r->acc_valid = acc_valid;
    r->host_acc_fresh = host_acc_fresh;
    if (rc) return rc;
    return download(r, t->lam[0], layout, lam0);
}

int mhd_adjoint_step_grids(mhd_run* r, const double* Uold, const double* Utmp, double dt, double ch, double* lam,
                           double* lam_tmp) {
    if (!r || !Uold || !Utmp || !lam || !lam_tmp) return fail(MHD_ERR_INVALID, "null argument");
    if (int rc = check_adjoint_run(r)) return rc;
    MHD_HIP(hipMemsetAsync(lam_tmp, 0, grid_bytes(r), r->stream()));
    return reverse_step(r, Uold, Utmp, dt, ch, lam, lam_tmp);
}

}  // extern "C"
