// CPU oracle for the 2-D Dedner GLM-MHD MUSCL-Hancock step.
//
// TEST INFRASTRUCTURE ONLY.  This is a C++17/OpenMP restatement of the
// reference CPU path (head/types.hpp, head/physics.hpp, head/riemann.hpp,
// src/cpu/solver_cpu.cpp, head/cpu/boundary_cpu.hpp, head/cpu/grid_cpu.hpp,
// src/test_cases.cpp, src/init.cpp, scripts/cpu/main_cpu.cpp:196-222),
// operation for operation, so that it is bitwise equal to the reference on
// every input.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline may load it; the HIP product never does.
//
// Pinning (each function and how it is checked):
//   point physics, Riemann solvers, GLM flux, minmod, reconstruct, BCs,
//   compute_dt, advance, run loop, reference case table and ICs
//       -> bitwise against golden vectors written by the compiled reference
//          (tests/golden/make_golden.py) and against oracle/_ref when built.
//   van Leer limiter (absent from the reference) -> closed forms
//          (symmetry, vl(a,a)=a, sign/zero rules, harmonic-mean identity).
//   ICs absent from the reference (unit-box Orszag-Tang, MHD shock-cloud,
//   batch OT family, Gaussian bump) -> closed-form point checks
//          (values at chosen points, periodicity, pressure balance).
//   SplitMix64 -> the published test vector for seed 1234567.
//   tangent-linear model (Dual scalar) -> central finite differences of the
//          double instantiation.
//   frozen-branch differences (Shadow scalar) -> the double instantiation for
//          the unperturbed copy (bitwise) and the tangent model (to O(eps^2)).
//
// Field order everywhere is the reference Conserved order
// (rho, rho*u, rho*v, rho*w, Bx, By, Bz, E, psi); primitives are
// (rho, u, v, w, Bx, By, Bz, p, psi).  States are AoS, row-major with ghost
// cells, index (j * (nx + 2 ng) + i) * 9 + k, as in head/cpu/grid_cpu.hpp.
#pragma once

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nx, ny, ng;
    double x_min, x_max, y_min, y_max;
} orc_geom;

// Parameters of the analytic initial conditions that the reference does not
// have.  Unused fields are ignored by a case.
typedef struct {
    double   gamma;          // adiabatic index (batch OT family)
    double   b_scale;        // magnetic amplitude factor (batch OT family)
    uint64_t seed;           // shock-cloud noise seed
    int      n_modes;        // density-perturbation modes (batch OT family)
    double   mode_amp[4];
    double   mode_kx[4];
    double   mode_ky[4];
    double   mode_phase[4];
    double   pert_amp;       // total relative perturbation amplitude (0.05)
    double   bump_amp;       // Gaussian density bump (adjoint target)
    double   bump_sigma;
    double   bump_cx, bump_cy;
} orc_ic_params;

// ---- point functions --------------------------------------------------------------
void   orc_cons_to_prim(const double* U, double gamma, double* V);
void   orc_prim_to_cons(const double* V, double gamma, double* U);
void   orc_flux(const double* U, int dir, double ch, double gamma, double* F);
double orc_fast_speed(const double* V, int dir, double gamma);
double orc_max_signal_speed(const double* V, int dir, double ch, double gamma);
void   orc_solver_flux(int kind, const double* UL, const double* UR, int dir, double ch, double gamma,
                       double* F);
void   orc_riemann_flux(const double* UL, const double* UR, int dir, int solver, double ch, double gamma,
                        double* F);
void   orc_apply_glm_flux(double* F, const double* UL, const double* UR, int dir, double ch);
double orc_limiter(double a, double b, int limiter);
int    orc_is_physical(const double* V);
void   orc_reconstruct(const double* Wm, const double* Wc, const double* Wp, double dt_over_d, int dir,
                       double ch, double gamma, int limiter, double* UL_star, double* UR_star);
uint64_t orc_splitmix64(uint64_t* state);
double   orc_cell_uniform(uint64_t seed, int i, int j);
// The portable sine of the batch family's initial state (ic.hpp, shared with the library).
double   orc_ic_sin(double a);

// ---- grids -----------------------------------------------------------------------------
// Returns 0, or 1 when no interior cell has a finite positive signal speed.
int  orc_compute_dt(const double* U, const orc_geom* g, double cfl, double gamma, double* ch_inout,
                    double* dt);
int  orc_advance(const double* Uold, double* Utmp, double* Unew, const orc_geom* g, double dt, double ch,
                 double cr, double gamma, int solver, int limiter, const int* bc);
void orc_apply_boundary(double* U, const orc_geom* g, const int* bc);
void orc_copy_ghost_cells(const double* src, double* dst, const orc_geom* g);

// Driver loop of scripts/cpu/main_cpu.cpp:196-222 with no snapshot schedule.
// U holds the newest state on return.  Stop codes as in the product's
// mhd_step_to: 0 reached t_end or max_steps, 1 compute_dt found no signal
// speed, 2 dt non-finite or <= 0, 3 dt below 1e-8 t_end.
int orc_run_loop(double* U, const orc_geom* g, double cfl, double t_end, int max_steps, double gamma,
                 double cr, int solver, int limiter, const int* bc, double* ch_inout, double* dts,
                 double* chs, int* nsteps, double* t_out);

// orc_run_loop with the time of the step loop alone (steady clock, as
// main_cpu.cpp:191,254-257 times it) returned in *seconds.
int orc_run_loop_timed(double* U, const orc_geom* g, double cfl, double t_end, int max_steps, double gamma,
                       double cr, int solver, int limiter, const int* bc, double* ch_inout, double* dts,
                       double* chs, int* nsteps, double* t_out, double* seconds);
int orc_num_threads(void);

// Case table: ids 0..3 are the reference's (src/test_cases.cpp:178-269), 4..6 are
// the unit-box Orszag-Tang, the MHD shock-cloud and Brio-Wu with transmissive
// y and CFL 0.8.  cfg_out = nx, ny, ng, x0, x1, y0, y1, cfl, t_end, gamma.
int  orc_case_config(int case_id, int n, double* cfg_out, int* bc_out);
// Fills every cell (ghosts included) from the analytic IC, as src/init.cpp:34-43.
// case_id 7 = batch OT family, 8 = unit-box OT plus Gaussian bump.
int  orc_init_state(int case_id, const orc_geom* g, const orc_ic_params* p, double* U);

// ---- tangent-linear model (forward mode, branches and dt/c_h frozen) ---------------
// Advances (U, dU) through nsteps steps with the given dt and c_h sequences.
// On return U is the primal state (bitwise equal to orc_advance) and dU the
// directional derivative d U_final / d U_0 . dU_0.
int orc_tangent_run(double* U, double* dU, const orc_geom* g, int nsteps, const double* dts,
                    const double* chs, double cr, double gamma, int solver, int limiter, const int* bc);

// Frozen-branch central differences of nsteps steps: the copies U +- eps e_q of
// each probe q = (row, column, field) follow every branch of the unperturbed
// trajectory (dt / c_h given).  rho0: final density of the unperturbed run
// (tny x tnx); rho_pm: [q][+,-][tny][tnx].  nprobe <= 32; returns 1 otherwise.
int orc_frozen_fd_run(const double* U, const int* probes, int nprobe, double eps, const orc_geom* g, int nsteps,
                      const double* dts, const double* chs, double cr, double gamma, int solver, int limiter,
                      const int* bc, double* rho0, double* rho_pm);

// Point tangents (value and directional derivative, same branch rules) of the
// functions above, for the dot-product tests of the hand-written adjoint.
void orc_tan_cons_to_prim(const double* U, const double* dU, double gamma, double* V, double* dV);
void orc_tan_prim_to_cons(const double* V, const double* dV, double gamma, double* U, double* dU);
void orc_tan_flux(const double* U, const double* dU, int dir, double ch, double gamma, double* F, double* dF);
void orc_tan_fast_speed(const double* V, const double* dV, int dir, double gamma, double* c, double* dc);
void orc_tan_limiter(double a, double da, double b, double db, int limiter, double* v, double* dv);
void orc_tan_riemann(const double* UL, const double* dUL, const double* UR, const double* dUR, int dir, int solver,
                     double ch, double gamma, double* F, double* dF);
void orc_tan_reconstruct(const double* Wm, const double* dWm, const double* Wc, const double* dWc, const double* Wp,
                         const double* dWp, double dt_over_d, int dir, double ch, double gamma, int limiter,
                         double* L, double* dL, double* R, double* dR);

#ifdef __cplusplus
}
#endif
