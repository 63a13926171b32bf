// C-ABI shim over the reference CPU solver, compiled from the reference's own
// sources where they lie (see oracle/Makefile).  Test infrastructure only: it
// is used to generate golden vectors and to pin the CPU restatement
// (oracle/mhd_oracle.cpp); nothing in the product links it.
//
// The reference's anonymous-namespace helpers (minmod, reconstruct,
// is_physical) are reached by compiling src/cpu/solver_cpu.cpp inside this
// translation unit.
#include "cpu/solver_cpu.cpp"   // reference src/cpu/solver_cpu.cpp (via -I<ref>/src)

#include <cstring>
#include <exception>
#include <map>
#include <utility>

#include "cpu/boundary_cpu.hpp"
#include "cpu/grid_cpu.hpp"
#include "init.hpp"
#include "test_cases.hpp"

namespace {

inline Conserved load_c(const double* a) {
    return Conserved(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8]);
}
inline Primitive load_p(const double* a) {
    return Primitive(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8]);
}
inline void store_c(const Conserved& c, double* a) {
    a[0] = c.rho; a[1] = c.rhou; a[2] = c.rhov; a[3] = c.rhow;
    a[4] = c.Bx;  a[5] = c.By;   a[6] = c.Bz;   a[7] = c.E; a[8] = c.psi;
}
inline void store_p(const Primitive& p, double* a) {
    a[0] = p.rho; a[1] = p.u; a[2] = p.v; a[3] = p.w;
    a[4] = p.Bx;  a[5] = p.By; a[6] = p.Bz; a[7] = p.p; a[8] = p.psi;
}
inline Direction dir_of(int d) { return d == 0 ? Direction::X : Direction::Y; }

BoundaryConfig bc_of(const int* bc) {
    BoundaryConfig b;
    b.left   = static_cast<BoundaryType>(bc[0]);
    b.right  = static_cast<BoundaryType>(bc[1]);
    b.bottom = static_cast<BoundaryType>(bc[2]);
    b.top    = static_cast<BoundaryType>(bc[3]);
    return b;
}

std::map<std::pair<int, int>, CpuWorkspace> g_ws;

CpuWorkspace& workspace_for(int nx, int ny) {
    CpuWorkspace& ws = g_ws[{nx, ny}];
    if (!ws.is_initialized()) ws.init(nx, ny);
    return ws;
}

}  // namespace

extern "C" {

void   ref_set_gamma(double g) { phys::gamma = g; }
double ref_get_gamma(void) { return phys::gamma; }
void   ref_set_ch(double ch) { phys::ch_glm = ch; }
double ref_get_ch(void) { return phys::ch_glm; }
void   ref_set_cr(double cr) { phys::cr_glm = cr; }

// ---- point functions (head/physics.hpp, head/riemann.hpp) --------------------
void ref_cons_to_prim(const double* U, double* V) { store_p(phys::cons_to_prim(load_c(U)), V); }
void ref_prim_to_cons(const double* V, double* U) { store_c(phys::prim_to_cons(load_p(V)), U); }
void ref_flux(const double* U, int dir, double ch, double* F) {
    store_c(dir == 0 ? phys::flux_x(load_c(U), ch) : phys::flux_y(load_c(U), ch), F);
}
double ref_fast_speed(const double* V, int dir) {
    return dir == 0 ? phys::fast_speed_x(load_p(V)) : phys::fast_speed_y(load_p(V));
}
double ref_max_signal_speed(const double* V, int dir, double ch) {
    return dir == 0 ? phys::max_signal_speed_x(load_p(V), ch)
                    : phys::max_signal_speed_y(load_p(V), ch);
}
// kind: 0 hll, 1 hllc, 2 hlld, 3 force (raw solver, no GLM overwrite)
void ref_solver_flux(int kind, const double* UL, const double* UR, int dir, double ch, double* F) {
    const Conserved a = load_c(UL), b = load_c(UR);
    const Direction d = dir_of(dir);
    Conserved f;
    switch (kind) {
        case 0: f = hll_flux(a, b, d, ch); break;
        case 1: f = hllc_flux(a, b, d, ch); break;
        case 2: f = hlld_flux(a, b, d, ch); break;
        default: f = force_flux(a, b, d, ch); break;
    }
    store_c(f, F);
}
void ref_riemann_flux(const double* UL, const double* UR, int dir, int solver, double ch, double* F) {
    store_c(riemann_flux(load_c(UL), load_c(UR), dir_of(dir), static_cast<RiemannSolver>(solver), ch), F);
}
void ref_apply_glm_flux(double* F, const double* UL, const double* UR, int dir, double ch) {
    Conserved f = load_c(F);
    apply_glm_flux(f, load_c(UL), load_c(UR), dir_of(dir), ch);
    store_c(f, F);
}

// ---- solver_cpu.cpp anonymous-namespace helpers --------------------------------
double ref_minmod(double a, double b) { return minmod_scalar(a, b); }
int    ref_is_physical(const double* V) { return is_physical(load_p(V)) ? 1 : 0; }
// Uses phys::ch_glm for the half-step flux, as the reference does.
void ref_reconstruct(const double* Wm, const double* Wc, const double* Wp, double dt_over_d,
                     int dir, double* UL_star, double* UR_star) {
    Conserved a, b;
    reconstruct_cell_muscl_hancock(load_p(Wm), load_p(Wc), load_p(Wp), dt_over_d, dir_of(dir), a, b);
    store_c(a, UL_star);
    store_c(b, UR_star);
}

// ---- grids ------------------------------------------------------------------------
void* ref_grid_new(int nx, int ny, int ng, double x0, double x1, double y0, double y1) {
    return new Grid2D(nx, ny, ng, x0, x1, y0, y1);
}
void   ref_grid_free(void* g) { delete static_cast<Grid2D*>(g); }
double ref_grid_dx(void* g) { return static_cast<Grid2D*>(g)->dx(); }
double ref_grid_dy(void* g) { return static_cast<Grid2D*>(g)->dy(); }
// AoS copy in reference field order, all cells including ghosts.
void ref_grid_get(void* g, double* out) {
    const auto& d = static_cast<Grid2D*>(g)->data();
    for (std::size_t k = 0; k < d.size(); ++k) store_c(d[k], out + 9 * k);
}
void ref_grid_set(void* g, const double* in) {
    auto& d = static_cast<Grid2D*>(g)->data();
    for (std::size_t k = 0; k < d.size(); ++k) d[k] = load_c(in + 9 * k);
}
void ref_grid_copy(void* src, void* dst) { *static_cast<Grid2D*>(dst) = *static_cast<Grid2D*>(src); }

// Case table (src/test_cases.cpp): cfg_out = nx, ny, ng, x0, x1, y0, y1, cfl, t_end, gamma
// bc_out = left, right, bottom, top.  Returns the number of snapshot times.
int ref_case_config(int case_id, int n, double* cfg_out, int* bc_out, double* snaps_out) {
    try {
        const CaseConfig c = get_n_case_config(static_cast<CaseId>(case_id), n);
        cfg_out[0] = c.nx; cfg_out[1] = c.ny; cfg_out[2] = c.ng;
        cfg_out[3] = c.x_min; cfg_out[4] = c.x_max; cfg_out[5] = c.y_min; cfg_out[6] = c.y_max;
        cfg_out[7] = c.cfl; cfg_out[8] = c.t_end; cfg_out[9] = c.gamma;
        bc_out[0] = static_cast<int>(c.bc.left);   bc_out[1] = static_cast<int>(c.bc.right);
        bc_out[2] = static_cast<int>(c.bc.bottom); bc_out[3] = static_cast<int>(c.bc.top);
        for (std::size_t k = 0; k < c.snapshot_times.size(); ++k) snaps_out[k] = c.snapshot_times[k];
        return static_cast<int>(c.snapshot_times.size());
    } catch (const std::exception&) {
        return -1;
    }
}
// make_n_grid: sets phys::gamma and fills every cell (ghosts included) analytically.
void* ref_make_n_grid(int case_id, int n) {
    return new Grid2D(make_n_grid(static_cast<CaseId>(case_id), n));
}
void ref_initial_state_at(int case_id, double x, double y, double* U) {
    store_c(initial_state_at(static_cast<CaseId>(case_id), x, y), U);
}

// ---- solver_cpu public API --------------------------------------------------------
// Returns 0 on success, 1 if compute_dt threw (non-positive max speed).
int ref_compute_dt(void* g, double cfl, double* dt) {
    try {
        *dt = compute_dt(*static_cast<Grid2D*>(g), cfl);
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}
int ref_advance(void* uold, void* utmp, void* unew, double dt, int solver, const int* bc) {
    Grid2D& o = *static_cast<Grid2D*>(uold);
    try {
        advance_second_order(o, *static_cast<Grid2D*>(utmp), *static_cast<Grid2D*>(unew), dt,
                             workspace_for(o.nx(), o.ny()), static_cast<RiemannSolver>(solver), bc_of(bc));
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}
void ref_apply_boundary(void* g, const int* bc) { apply_boundary(*static_cast<Grid2D*>(g), bc_of(bc)); }
void ref_copy_ghost_cells(void* src, void* dst) {
    copy_ghost_cells(*static_cast<Grid2D*>(src), *static_cast<Grid2D*>(dst));
}

// The driver loop of scripts/cpu/main_cpu.cpp:196-222 without output: no
// snapshot schedule, t_next = t_end.  Runs at most max_steps steps.  Records dt
// and c_h per step.  Returns a stop code: 0 reached t_end or max_steps,
// 1 compute_dt threw, 2 dt non-finite or <= 0, 3 dt below the 1e-8*t_end floor.
int ref_run_loop(void* uold, void* utmp, void* unew, double cfl, double t_end, int solver,
                 const int* bc, int max_steps, double* dts, double* chs, int* nsteps, double* t_out) {
    Grid2D* o = static_cast<Grid2D*>(uold);
    Grid2D* n = static_cast<Grid2D*>(unew);
    double t = 0.0;
    int step = 0, code = 0;
    while (t < t_end && step < max_steps) {
        double dt_raw;
        if (ref_compute_dt(o, cfl, &dt_raw) != 0) { code = 1; break; }
        const double dt = std::min(dt_raw, t_end - t);
        if (!std::isfinite(dt) || dt <= 0.0) { code = 2; break; }
        if (dt < 1.0e-8 * t_end) { code = 3; break; }
        if (ref_advance(o, utmp, n, dt, solver, bc) != 0) { code = 4; break; }
        std::swap(o, n);
        t += dt;
        dts[step] = dt;
        chs[step] = phys::ch_glm;
        ++step;
    }
    // The caller's uold handle must hold the newest state.
    if (o != static_cast<Grid2D*>(uold)) std::swap(*static_cast<Grid2D*>(uold), *static_cast<Grid2D*>(unew));
    *nsteps = step;
    *t_out = t;
    return code;
}

}  // extern "C"
