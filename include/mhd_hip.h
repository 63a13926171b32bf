/*
 * mhd_hip.h -- C ABI of the MI355X (gfx950) GLM-MHD MUSCL-Hancock step.
 *
 * This is the drop-in boundary for the reference's solver interface.  Each
 * entry point names the reference interface it replaces.  No C++ types, no
 * exceptions and no torch types cross it: plain pointers, sizes, enums as
 * ints and an int status (MHD_OK == 0); mhd_last_error() returns the message
 * of the last failure on the calling thread.
 *
 * State layout.  A grid is 9 SoA planes of float64 in the reference field
 * order (rho, rho*u, rho*v, rho*w, Bx, By, Bz, E, psi) -- head/types.hpp:16-26.
 * Each plane holds (nx + 2 ng) x (ny + 2 ng) cells, row-major, ghost cells
 * included, index j * (nx + 2 ng) + i -- head/cpu/grid_cpu.hpp:85-87,
 * head/gpu/grid_gpu.cuh:102-110.  ng is 2.  Planes are `plane_stride`
 * doubles apart (mhd_plane_stride()).
 */
#ifndef MHD_HIP_H
#define MHD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------- */
enum {
    MHD_OK = 0,
    MHD_ERR_INVALID = 1,    /* bad argument, size mismatch, unknown enum */
    MHD_ERR_HIP = 2,        /* HIP runtime error (allocation, launch, copy) */
    MHD_ERR_NO_SIGNAL = 3,  /* compute_dt: non-positive maximum wave speed
                               (solver_cpu.cpp:304-306 throws here) */
    MHD_ERR_DT = 4,         /* step_to: dt non-finite or <= 0 (main_cpu.cpp:205-209) */
    MHD_ERR_DT_FLOOR = 5,   /* step_to: dt < 1e-8 t_end, blow-up stop (main_cpu.cpp:210-216) */
    MHD_ERR_NOT_BUILT = 6,  /* no HIP device or kernels unavailable */
    MHD_ERR_IO = 7          /* file I/O (snapshot CSVs, state dumps, driver log) */
};

/* ---- enums: same numeric order as the reference -------------------------------- */
/* RiemannSolver { HLL, HLLC, HLLD, FORCE } -- head/riemann.hpp:23 */
enum { MHD_HLL = 0, MHD_HLLC = 1, MHD_HLLD = 2, MHD_FORCE = 3 };
/* BoundaryType { Periodic, Transmissive, Reflecting, Dirichlet } -- head/test_cases.hpp:11-16 */
enum { MHD_BC_PERIODIC = 0, MHD_BC_TRANSMISSIVE = 1, MHD_BC_REFLECTING = 2, MHD_BC_DIRICHLET = 3 };
/* Slope limiter: minmod is the reference's (solver_cpu.cpp:26-43); van Leer is new. */
enum { MHD_LIMITER_MINMOD = 0, MHD_LIMITER_VANLEER = 1 };
/* CaseId { KelvinHelmholtz, ShockBubble, BrioWu, OrszagTang } -- head/test_cases.hpp:29-34,
 * then the cases this build adds. */
enum {
    MHD_CASE_KELVIN_HELMHOLTZ = 0,
    MHD_CASE_SHOCK_BUBBLE = 1,
    MHD_CASE_BRIO_WU = 2,
    MHD_CASE_ORSZAG_TANG = 3,
    MHD_CASE_ORSZAG_TANG_UNIT = 4,  /* [0,1]^2, rho = 25/(36 pi), p = 5/(12 pi), B / sqrt(4 pi) */
    MHD_CASE_SHOCK_CLOUD = 5,       /* [0,2]x[0,1] MHD shock-cloud with seeded bubble noise */
    MHD_CASE_BRIO_WU_T = 6,         /* Brio-Wu, transmissive in y, CFL 0.8 */
    MHD_CASE_OT_FAMILY = 7,         /* batch OT family: gamma, B scale, 4-mode density ripple */
    MHD_CASE_OT_BUMP = 8            /* unit OT plus a Gaussian density bump */
};

/* Per-run description (replaces the process-global phys::gamma / ch_glm / cr_glm of
 * head/physics.hpp:26-43 and the GpuWorkspace of head/gpu/solver_gpu.cuh:9-20). */
typedef struct {
    int nx, ny, ng;                      /* ng must be 2 */
    double x_min, x_max, y_min, y_max;   /* dx = (x_max - x_min) / nx, grid_cpu.hpp:25-26 */
    double gamma;                        /* adiabatic index */
    double cr;                           /* mixed-GLM c_r = c_p^2 / c_h (0.18) */
    double cfl;
    int solver;                          /* MHD_HLL .. MHD_FORCE */
    int limiter;                         /* MHD_LIMITER_* */
    int bc[4];                           /* left, right, bottom, top */
    void* stream;                        /* hipStream_t; NULL = the null stream */
} mhd_run_desc;

/* Case table entry (head/test_cases.hpp:38-55 CaseConfig). */
typedef struct {
    int nx, ny, ng;
    double x_min, x_max, y_min, y_max;
    double cfl, t_end, gamma;
    int bc[4];
} mhd_case_config;

/* Parameters of the analytic ICs this build adds (ignored by the reference cases). */
typedef struct {
    double gamma;
    double b_scale;
    uint64_t seed;
    int n_modes;
    double mode_amp[4], mode_kx[4], mode_ky[4], mode_phase[4];
    double pert_amp;
    double bump_amp, bump_sigma, bump_cx, bump_cy;
} mhd_ic_params;

typedef struct mhd_run mhd_run;

const char* mhd_last_error(void);
int mhd_device_count(int* n);

/* ---- problem set-up: src/test_cases.cpp:178-300, src/init.cpp:18-65 ---------- */
/* get_n_case_config(CaseId, n) -- test_cases.cpp:275-282 */
int mhd_case_config_get(int case_id, int n_scale, mhd_case_config* out);
/* initialise_grid -- init.cpp:34-43: fills every cell, ghosts included, of a host
 * AoS array (cells x 9, reference field order) from the analytic IC. */
int mhd_init_state_host(int case_id, const mhd_case_config* geom, const mhd_ic_params* params, double* aos);
/* initialise_grid + upload_from_aos (init.cpp:34-43, grid_gpu.cuh:139-189) on the
 * device: fills every cell, ghosts included, of a caller's device SoA grid (9 planes of
 * (nx+4)(ny+4) doubles) on `stream`.  Bitwise the mhd_init_state_host state.  Only the
 * cases whose state needs no C-library transcendental: MHD_CASE_BRIO_WU(_T),
 * MHD_CASE_SHOCK_BUBBLE, MHD_CASE_SHOCK_CLOUD and MHD_CASE_OT_FAMILY (whose sines
 * are a portable routine shared with the host); MHD_ERR_INVALID otherwise. */
int mhd_init_state_device(int case_id, const mhd_case_config* geom, const mhd_ic_params* params, double* soa,
                          void* stream);

/* ---- run handle: Grid2DGPU x 3 + GpuWorkspace (grid_gpu.cuh:17-223,
 *      solver_gpu.cuh:9-35): allocates Uold/Utmp/Unew on the device. ------------ */
int mhd_run_create(const mhd_run_desc* desc, mhd_run** out);
int mhd_run_destroy(mhd_run* run);
/* The stream the run's work is queued on from the next call on (the descriptor's
 * `stream`, NULL = the null stream).  The reference queues everything on the legacy
 * default stream (solver_gpu.cu:595,602,682); here a caller such as a framework's
 * current stream can take over a run: the new stream first waits, on the device, for
 * the work already queued on the old one. */
int mhd_run_set_stream(mhd_run* run, void* stream);
size_t mhd_plane_stride(const mhd_run* run);   /* doubles between planes */
/* Device base pointer of the current state (9 planes, plane_stride apart). */
double* mhd_state_ptr(mhd_run* run);

/* upload_from_aos / download_to_aos -- grid_gpu.cuh:120-177.
 * layout 0: host AoS (cells x 9); layout 1: host SoA (9 x cells, dense);
 * layout 2: device SoA (9 x cells, dense). */
int mhd_set_state(mhd_run* run, const double* src, int layout);
int mhd_get_state(mhd_run* run, double* dst, int layout);
/* mhd_init_state_device into the run's grids: the same state as mhd_init_state_host
 * followed by mhd_set_state, with no host work or upload. */
int mhd_run_init_state(mhd_run* run, int case_id, const mhd_ic_params* params);

/* set_gpu_physics_ch / the host mirror phys::ch_glm -- solver_gpu.cu:536-544 */
int mhd_set_ch(mhd_run* run, double ch);
int mhd_get_ch(const mhd_run* run, double* ch);

/* compute_dt(const Grid2D&, cfl) -- solver_cpu.hpp:49, solver_cpu.cpp:283-312;
 * compute_dt_gpu -- solver_gpu.cuh:40.  Ratchets c_h to max(c_h, max signal
 * speed) and returns dt = cfl min(dx, dy) / c_h.  MHD_ERR_NO_SIGNAL where the
 * reference throws.  Synchronises the run's stream. */
int mhd_compute_dt(mhd_run* run, double* dt);

/* advance_second_order(Uold, Utmp, Unew, dt, ws, solver, bc) -- solver_cpu.cpp:327-440,
 * advance_second_order_gpu -- solver_gpu.cuh:45-53.  x-sweep, BC, y-sweep, BC,
 * psi damping; then Uold <- Unew.  Asynchronous on the run's stream. */
int mhd_advance(mhd_run* run, double dt);

/* The driver loop of scripts/cpu/main_cpu.cpp:196-222 without output:
 * while (t < t_end && steps < max_steps) { dt = min(compute_dt, t_end - t);
 * guards; advance; t += dt }.  dts/chs (nullable, max_steps long) receive
 * each step's dt and c_h.  Returns MHD_OK, MHD_ERR_NO_SIGNAL, MHD_ERR_DT or
 * MHD_ERR_DT_FLOOR; *steps and *t always hold what was done. */
int mhd_step_to(mhd_run* run, double t_end, int max_steps, double* t, int* steps, double* dts, double* chs);

/* apply_boundary_gpu -- boundary_gpu.cuh:9 (boundary_cpu.hpp:45-149): BC on the
 * current state, in place (Dirichlet sides keep their values). */
int mhd_apply_boundary(mhd_run* run);

/* ---- grid-level entry points over caller-owned device grids ------------------------
 * Each grid is 9 dense SoA planes of (nx+4)(ny+4) doubles (plane stride = cells).
 * These are the Grid2DGPU-view forms of compute_dt_gpu / advance_second_order_gpu;
 * the run only supplies parameters, c_h and scratch. */
int mhd_compute_dt_grid(mhd_run* run, const double* U, double* dt);
int mhd_advance_grids(mhd_run* run, const double* Uold, double* Utmp, double* Unew, double dt);

/* mhd_advance_grids for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  It takes the run description and c_h directly instead of a run handle,
 * and calls no hipMalloc, hipFree or other allocating routine: the step's scratch
 * (CFL accumulator and the 4 (nx + ny)-double psi frame) lives in a fixed device array
 * of MHD_NOALLOC_SCRATCH_BYTES that each GPU holds until the process ends, so
 * nx + ny <= MHD_NOALLOC_MAX_NXNY (MHD_ERR_INVALID beyond).  Asynchronous on
 * desc->stream; max_speed (nullable) receives the CFL maximum of Unew, max over the
 * interior of fmax(signal speed, ch) -- the next compute_dt's c_h -- and then the call
 * synchronises the stream.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams, other threads) race. */
#define MHD_NOALLOC_SCRATCH_BYTES (1 << 20)
#define MHD_NOALLOC_MAX_NXNY ((MHD_NOALLOC_SCRATCH_BYTES - 256) / 32)
int mhd_advance_grids_noalloc(const mhd_run_desc* desc, double ch, const double* Uold, double* Utmp, double* Unew,
                              double dt, double* max_speed);

/* Synchronise the run's stream (errors from asynchronous launches surface here). */
int mhd_sync(mhd_run* run);

/* ---- batched form (no counterpart in the reference, which runs one case per
 *      process with process-global physics, head/physics.hpp:26-43) ------------------
 * n independent runs -- any grid sizes, gamma, c_r, CFL and BCs; one Riemann solver
 * and one limiter per batch -- advance together.  Per step, one launch per sweep
 * covers every run (a persistent grid over a tile list spanning all runs), one
 * launch fills every run's ghosts, and the whole step is a hipGraph replayed each
 * step.  Each run keeps its own c_h ratchet, dt from its own CFL maximum and psi
 * damping, computed with the arithmetic of mhd_compute_dt / mhd_advance, so each
 * run is bitwise the run mhd_step_to produces for it alone. */
typedef struct mhd_batch mhd_batch;
int mhd_batch_create(const mhd_run_desc* descs, int n, mhd_batch** out);
int mhd_batch_destroy(mhd_batch* batch);
int mhd_batch_size(const mhd_batch* batch, int* n);
int mhd_batch_set_state(mhd_batch* batch, int run, const double* src, int layout);
/* Every run's initial state on the device in one launch (params: one per run, or NULL
 * for cases without parameters): as mhd_init_state_host + mhd_batch_set_state per run,
 * bitwise, with a few hundred bytes uploaded per run.  Cases as mhd_init_state_device. */
int mhd_batch_init_states(mhd_batch* batch, int case_id, const mhd_ic_params* params);
int mhd_batch_get_state(mhd_batch* batch, int run, double* dst, int layout);
int mhd_batch_get_ch(const mhd_batch* batch, int run, double* ch);
int mhd_batch_set_ch(mhd_batch* batch, int run, double ch);
/* Every run takes nsteps steps (compute_dt + advance).  dts (nullable) receives
 * nsteps x n values, step-major.  MHD_ERR_NO_SIGNAL if some run has no signal
 * speed; *done = steps completed by all runs. */
int mhd_batch_step(mhd_batch* batch, int nsteps, double* dts, int* done);
/* Autotuner: times every candidate launch shape (x-sweep width and blocks per CU,
 * y-sweep rows per tile and blocks per CU) by graph replays from a saved copy of
 * the states, restores the states and keeps the fastest.  report (nullable)
 * receives the timing table as text. */
int mhd_batch_tune(mhd_batch* batch, int reps, char* report, size_t report_len);
int mhd_batch_set_shape(mhd_batch* batch, int xwg, int xgrid, int yseg, int ygrid);
/* Parameters of the Orszag-Tang family batch: the multiset {128 x32, 256 x16,
 * 512 x12, 1024 x4} (n = 64) shuffled with the seed, per run gamma = 1.4 + 0.6 u,
 * b_scale = 0.5 + u and 4 density modes (amplitude 2u - 1, integer |k| <= 4,
 * k != 0, phase 2 pi u), u from one SplitMix64 stream.  For n != 64 the sizes
 * cycle through 128, 256, 512.  sizes[n], params[n]. */
int mhd_batch_plan_ot_family(uint64_t seed, int n, int* sizes, mhd_ic_params* params);

/* ---- discrete adjoint (no counterpart in the reference) ------------------------
 * A tape differentiates nsteps steps of a run: U_N = step^N(U_0), each step
 * compute_dt + advance_second_order.  dt and c_h are frozen at the recorded
 * values and every branch (Riemann-solver region and fallbacks, limiter
 * selections, physical-state guards) at the one the recorded trajectory takes.
 * Supported: every Riemann solver (HLL, HLLC, HLLD, FORCE) and limiter, and every boundary
 * type; through a Dirichlet side the ghost cells copied from the previous grid carry
 * their adjoint back to it (boundary_cpu.hpp:12-32, solver_cpu.cpp:396-398,434-436).  The
 * gradient is with respect to every value of U_0, ghost cells included. */
typedef struct mhd_tape mhd_tape;
/* Device memory: ceil(nsteps/K) checkpoints, 2K+1 segment grids, 2 adjoint grids; with
 * two or more segments a second set of 2K+1 segment grids and a second stream, so that
 * the backward recomputes the next segment while it reverses the current one (one set
 * when the second does not fit, or with MHD_TAPE_SETS=1 in the environment). */
int mhd_tape_create(mhd_run* run, int nsteps, int K, mhd_tape** out);
int mhd_tape_destroy(mhd_tape* tape);
/* Advances the run nsteps steps from its current state (as mhd_step_to, without
 * the t_end clip), keeping the state before every K-th step.  dts and chs
 * (nullable, nsteps each) receive each step's dt and c_h. */
int mhd_tape_forward(mhd_tape* tape, double* dts, double* chs);
/* lam_final = dL/dU_N in, lam0 = dL/dU_0 out, both in an mhd_set_state layout.
 * Recomputes each K-step segment from its checkpoint, then runs the reverse
 * steps (both ordered after the run's stream, and the run's stream after them).
 * MHD_ERR_INVALID when the recorded trajectory is not finite. */
int mhd_tape_backward(mhd_tape* tape, const double* lam_final, double* lam0, int layout);
/* One reverse step over caller device grids: Uold a step's input, Utmp its
 * intermediate U* (ghosts filled), lam = dL/d(step output) in and dL/dUold out;
 * lam_tmp is scratch (receives dL/dU*). */
int mhd_adjoint_step_grids(mhd_run* run, const double* Uold, const double* Utmp, double dt, double ch, double* lam,
                           double* lam_tmp);

/* ---- run driver and output formats: scripts/cpu/main_cpu.cpp:36-276,
 *      scripts/gpu/main_gpu.cu:198-280, src/test_cases.cpp:206-264 --------------- */
/* CaseConfig::snapshot_times / snapshot_tags (test_cases.hpp:50-54) of a case: the
 * shock-bubble's nine Haas & Sturtevant times (test_cases.cpp:206-232) and
 * Orszag-Tang's pi, 2 pi (:263-264); none for the other cases.  *count gets the number;
 * with max_count >= it, times[k] and the NUL-terminated tags[k * tag_len] are filled. */
int mhd_case_snapshots(int case_id, int n_scale, double* times, char* tags, int tag_len, int max_count, int* count);
/* write_all_fields (main_cpu.cpp:66-91): <dir>/<prefix>_<f>.csv for f in rho, u, v, Bx,
 * By, Bz, E, psi, p (u = rho u / rho, p from cons_to_prim), interior cells, rows = j from
 * the top, values as "%.8e", each written to <file>.tmp and renamed; dir is created.
 * _host takes a host SoA grid (9 planes of (nx+4)(ny+4) doubles). */
int mhd_write_fields_csv(mhd_run* run, const char* dir, const char* prefix);
int mhd_write_fields_csv_host(const double* soa, int nx, int ny, double gamma, const char* dir, const char* prefix);
/* Binary FP64 state dump (no counterpart in the reference): a 128-byte header (magic
 * "MHDSTATE", u32 version 1, u32 header bytes, i32 nx, ny, ng, 9 fields, f64 t, i64 step,
 * f64 gamma, f64 c_h, u32 layout 1, reserved), then the 9 SoA planes of the run's
 * current state, ghosts included, little-endian.  _read checks the grid against the run,
 * loads it as mhd_set_state does and restores c_h, so the run continues as it would have. */
int mhd_write_state_bin(mhd_run* run, const char* path, double t, long long step);
int mhd_read_state_bin(mhd_run* run, const char* path, double* t, long long* step);

enum {
    MHD_DRIVE_DONE = 0,       /* reached t_end */
    MHD_DRIVE_BAD_DT = 1,     /* dt non-finite or <= 0 (main_cpu.cpp:205-209) */
    MHD_DRIVE_DT_FLOOR = 2,   /* dt < 1e-8 t_end (main_cpu.cpp:210-216) */
    MHD_DRIVE_MAX_STEPS = 3   /* max_steps taken first */
};
typedef struct {
    const char* case_name;            /* file prefix <case>_<platform>_<tag> */
    const char* platform;             /* "gpu" (main_gpu.cu:198) or "cpu"; NULL = "gpu" */
    const char* out_dir;              /* NULL: no files (--no-out) */
    const char* log_path;             /* progress / [TIMING] lines appended here; NULL = stdout */
    int n_scale;                      /* the [TIMING] line's n */
    double t_end;
    int n_snapshots;                  /* mhd_case_snapshots */
    const double* snapshot_times;
    const char* const* snapshot_tags;
    double print_interval;            /* progress line when t passes it (main_cpu.cpp:221-237); <= 0: none */
    int max_steps;                    /* 0: no limit */
    int verbose;                      /* write the progress, [snap], Done and [TIMING] lines */
    double t_start;                   /* restart (mhd_read_state_bin): time and step of the state; */
    long long step_start;             /* with step_start > 0 no t0 files, passed snapshots skipped */
} mhd_drive_opts;
typedef struct {
    double t;
    long long steps;
    double elapsed_s;
    double last_dt;
    int status;                       /* MHD_DRIVE_* */
} mhd_drive_result;
/* The time loop of main_cpu.cpp:176-276 on the run: writes <case>_<platform>_t0, then
 * per step dt = min(compute_dt, next snapshot or t_end - t), the two dt guards, advance,
 * a progress line with max |B| every print_interval of t, the snapshots reached, and at
 * the end the [TIMING] line and, without a schedule, the final state tagged
 * t<round(100 t_end)>.  One deliberate difference: the snapshot index advances without
 * output too, so a run with out_dir = NULL reaches t_end (the reference stops there with
 * dt = 0).  Returns MHD_OK with result->status, or an error (compute_dt's, I/O); the
 * result holds the t and step reached either way. */
int mhd_drive(mhd_run* run, const mhd_drive_opts* opts, mhd_drive_result* result);

/* ---- measurement (no counterpart in the reference, whose only timer is the
 *      steady_clock around the loop, main_cpu.cpp:191,254-269) --------------------
 * While enabled, every step records HIP events on the run's stream around its
 * launches.  mhd_profile_read synchronises and returns the summed milliseconds of
 * each launch kind since enabling -- ms[0] x sweep, ms[1] ghost fill after it,
 * ms[2] y sweep, ms[3] ghost fill after it -- and the number of steps. */
int mhd_profile(mhd_run* run, int enable);
int mhd_profile_read(mhd_run* run, double* ms4, int* steps);

/* The HBM bandwidth this device reaches on streaming kernels of known traffic (16-byte
 * lanes, 4 workgroups of 256 lanes per CU, grid-stride), the practical ceiling against
 * which the sweeps' bytes per second are read: tbps3[0] a copy of a into b (2 * bytes
 * moved per launch), tbps3[1] a read of a, tbps3[2] a write of b, in TB/s, each timed over
 * reps launches after one warm-up on the null stream (synchronous).  a and b are distinct
 * device buffers of `bytes` bytes (a multiple of 16, at least 1 MiB); on return b holds a
 * copy of a. */
int mhd_probe_bandwidth(const double* a, double* b, long long bytes, int reps, double* tbps3);

#ifdef __cplusplus
}
#endif
#endif /* MHD_HIP_H */
