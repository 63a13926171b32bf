// Host-only AddressSanitizer / UndefinedBehaviorSanitizer run of the CPU oracle
// (test infrastructure, built by `make -C oracle sanitize` into oracle/_san/).
//
// Drives every entry point the tests use over small grids: each Riemann solver and
// limiter on the unit-box Orszag-Tang vortex, every boundary type on each side, the
// failure paths of the reference driver (a NaN cell, an all-unphysical grid whose
// compute_dt finds no signal speed, HLLC Brio-Wu at CFL 0.4 running into the dt
// floor), every initial condition, the tangent-linear and frozen-branch runs, and the
// point functions.  Any out-of-bounds access, use of uninitialised heap memory the
// sanitizers see, signed overflow or other undefined behaviour aborts the run with a
// report; a clean run prints "sanitize: ok" and exits 0.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mhd_oracle.hpp"

namespace {

int fails = 0;

void expect(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "sanitize: %s\n", what);
        ++fails;
    }
}

std::vector<double> state(int case_id, const orc_geom& g, const orc_ic_params* p = nullptr) {
    std::vector<double> U((size_t)(g.nx + 2 * g.ng) * (g.ny + 2 * g.ng) * 9);
    expect(orc_init_state(case_id, &g, p, U.data()) == 0, "init_state");
    return U;
}

int run(std::vector<double>& U, const orc_geom& g, double cfl, double t_end, int steps, double gamma, int solver,
        int limiter, const int* bc, int* nsteps) {
    std::vector<double> dts(steps), chs(steps);
    double ch = 0.0, t = 0.0;
    return orc_run_loop(U.data(), &g, cfl, t_end, steps, gamma, 0.18, solver, limiter, bc, &ch, dts.data(),
                        chs.data(), nsteps, &t);
}

}  // namespace

int main() {
    const orc_geom ot{24, 20, 2, 0.0, 1.0, 0.0, 1.0};
    const double gamma = 5.0 / 3.0;

    // every solver x limiter, periodic
    const int periodic[4] = {0, 0, 0, 0};
    for (int s = 0; s < 4; ++s)
        for (int l = 0; l < 2; ++l) {
            std::vector<double> U = state(4, ot);
            int n = 0;
            const int code = run(U, ot, 0.8, 0.5, 6, gamma, s, l, periodic, &n);
            expect(code == 0 && n == 6, "OT run");
        }

    // boundary mixes: transmissive, reflecting, Dirichlet on each side
    const int mixes[4][4] = {{1, 1, 1, 1}, {2, 2, 2, 2}, {3, 1, 0, 0}, {1, 3, 2, 3}};
    for (const auto& bc : mixes) {
        std::vector<double> U = state(4, ot);
        int n = 0;
        run(U, ot, 0.6, 0.5, 4, gamma, 2, 1, bc, &n);
        std::vector<double> V = U;
        orc_apply_boundary(V.data(), &ot, bc);
        orc_copy_ghost_cells(U.data(), V.data(), &ot);
    }

    // a NaN cell: fallbacks in reconstruct, Riemann and the update revert
    {
        std::vector<double> U = state(4, ot);
        U[((size_t)7 * (ot.nx + 4) + 9) * 9 + 0] = std::nan("");
        int n = 0;
        run(U, ot, 0.8, 0.5, 3, gamma, 2, 1, periodic, &n);
    }
    // every cell unphysical: compute_dt finds no signal speed
    {
        std::vector<double> U = state(4, ot);
        for (size_t c = 0; c < U.size(); c += 9) U[c] = -1.0;
        double ch = 0.0, dt = 0.0;
        expect(orc_compute_dt(U.data(), &ot, 0.8, gamma, &ch, &dt) != 0, "no-signal compute_dt");
    }
    // HLLC Brio-Wu at CFL 0.4 stops at the dt floor
    {
        double cfg[10];
        int bc[4];
        expect(orc_case_config(2, 1, cfg, bc) == 0, "case_config");
        const orc_geom bw{(int)cfg[0], (int)cfg[1], (int)cfg[2], cfg[3], cfg[4], cfg[5], cfg[6]};
        std::vector<double> U = state(2, bw);
        int n = 0;
        const int code = run(U, bw, cfg[7], cfg[8], 2000, cfg[9], 1, 0, bc, &n);
        expect(code == 3, "HLLC Brio-Wu dt floor");
    }

    // every initial condition, on a small grid of its own domain
    orc_ic_params p{};
    p.gamma = 1.7; p.b_scale = 1.2; p.seed = 20261203; p.n_modes = 4; p.pert_amp = 0.05;
    for (int q = 0; q < 4; ++q) {
        p.mode_amp[q] = 0.3 * (q + 1); p.mode_kx[q] = q; p.mode_ky[q] = 2 - q; p.mode_phase[q] = 0.7 * q;
    }
    p.bump_amp = 0.02; p.bump_sigma = 0.05; p.bump_cx = 0.4; p.bump_cy = 0.6;
    for (int c = 0; c <= 8; ++c) {
        double cfg[10];
        int bc[4];
        const int cc = c <= 6 ? c : 4;
        expect(orc_case_config(cc, 1, cfg, bc) == 0, "case_config");
        const orc_geom g{16, 12, 2, cfg[3], cfg[4], cfg[5], cfg[6]};
        std::vector<double> U = state(c, g, &p);
    }

    // tangent-linear and frozen-branch runs
    {
        std::vector<double> U = state(4, ot);
        std::vector<double> dts(3), chs(3);
        double ch = 0.0, t = 0.0;
        int n = 0;
        std::vector<double> W = U;
        orc_run_loop(W.data(), &ot, 0.8, 0.5, 3, gamma, 0.18, 2, 1, periodic, &ch, dts.data(), chs.data(), &n, &t);
        std::vector<double> dU(U.size(), 1e-3);
        std::vector<double> T = U;
        expect(orc_tangent_run(T.data(), dU.data(), &ot, 3, dts.data(), chs.data(), 0.18, gamma, 2, 1, periodic) == 0,
               "tangent run");
        const int probes[6] = {5, 6, 0, 10, 11, 7};
        const size_t plane = (size_t)(ot.nx + 4) * (ot.ny + 4);
        std::vector<double> rho0(plane), rho_pm(2 * 2 * plane);
        expect(orc_frozen_fd_run(U.data(), probes, 2, 1e-6, &ot, 3, dts.data(), chs.data(), 0.18, gamma, 2, 1,
                                 periodic, rho0.data(), rho_pm.data()) == 0,
               "frozen-branch run");
    }

    // point functions
    {
        const double UL[9] = {1.0, 0.3, -0.2, 0.1, 0.75, 1.0, -0.3, 2.5, 0.01};
        const double UR[9] = {0.125, -0.05, 0.02, 0.0, 0.75, -1.0, 0.2, 0.3, -0.02};
        double V[9], F[9], W[9];
        orc_cons_to_prim(UL, gamma, V);
        orc_prim_to_cons(V, gamma, W);
        for (int dir = 0; dir < 2; ++dir) {
            orc_flux(UL, dir, 0.5, gamma, F);
            (void)orc_fast_speed(V, dir, gamma);
            (void)orc_max_signal_speed(V, dir, 0.5, gamma);
            for (int s = 0; s < 4; ++s) {
                orc_riemann_flux(UL, UR, dir, s, 0.5, gamma, F);
                orc_solver_flux(s, UL, UR, dir, 0.5, gamma, F);
            }
            orc_apply_glm_flux(F, UL, UR, dir, 0.5);
            double L[9], R[9];
            orc_reconstruct(V, V, V, 0.1, dir, 0.5, gamma, 1, L, R);
        }
        (void)orc_limiter(0.3, 0.7, 1);
        (void)orc_is_physical(V);
        (void)orc_ic_sin(12.5);
        (void)orc_cell_uniform(7, 3, 4);
    }

    if (fails) return 1;
    std::printf("sanitize: ok\n");
    return 0;
}
