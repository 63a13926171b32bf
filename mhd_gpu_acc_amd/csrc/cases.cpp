// Problem set-up of the drop-in interface: the case table (src/test_cases.cpp:178-286)
// and the analytic initial states (src/test_cases.cpp:24-148, src/init.cpp:34-43),
// plus the cases this build adds (unit-box Orszag-Tang, MHD shock-cloud,
// transmissive Brio-Wu, the batch Orszag-Tang family and the Gaussian bump).
//
// The reference's cases are evaluated on the host with the C library's sin, tanh,
// exp and sqrt, as the reference does, so a state built here is bitwise the
// state the reference (or any CPU code calling the same libm) builds.  The cases
// in ic.hpp (Brio-Wu, shock-bubble, shock-cloud, the batch family) are the same
// source the device initialiser runs.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/mhd_hip.h"
#include "ic.hpp"

namespace {

using mhd_ic::kPi;
using mhd_ic::State;
using mhd_ic::from_primitives;
using mhd_ic::splitmix_next;

// Kelvin-Helmholtz (test_cases.cpp:43-63)
State kelvin_helmholtz(double x, double y) {
    const double ux = 5.0 * (std::tanh(20.0 * (y + 0.5)) - (std::tanh(20.0 * (y - 0.5)) + 1.0));
    const double uy = 0.25 * std::sin(2.0 * kPi * x) *
                      (std::exp(-100.0 * (y + 0.5) * (y + 0.5)) - std::exp(-100.0 * (y - 0.5) * (y - 0.5)));
    return from_primitives(1.0, ux, uy, 0.0, 1.0, 0.0, 0.0, 50.0, 1.4);
}

// Orszag-Tang on [0, 2 pi]^2 (test_cases.cpp:139-148)
State orszag_tang(double x, double y) {
    const double gamma = 5.0 / 3.0;
    return from_primitives(gamma * gamma, -std::sin(y), std::sin(x), 0.0, -std::sin(y), std::sin(2.0 * x), 0.0,
                           gamma, gamma);
}

// Orszag-Tang on [0,1]^2: rho = 25/(36 pi), p = 5/(12 pi), u = -sin 2 pi y,
// v = sin 2 pi x, B = (-sin 2 pi y, sin 4 pi x) / sqrt(4 pi); with `bump`,
// rho += bump_amp exp(-r^2 / (2 sigma^2)) (p held, E recomputed).
State orszag_tang_unit(double x, double y, const mhd_ic_params* p, bool bump) {
    double rho = 25.0 / (36.0 * kPi);
    const double pr = 5.0 / (12.0 * kPi);
    const double b0 = 1.0 / std::sqrt(4.0 * kPi);
    if (bump) {
        const double r2 = (x - p->bump_cx) * (x - p->bump_cx) + (y - p->bump_cy) * (y - p->bump_cy);
        rho = rho + p->bump_amp * std::exp(-r2 / (2.0 * p->bump_sigma * p->bump_sigma));
    }
    const double sy = std::sin(2.0 * kPi * y);
    const double sx = std::sin(2.0 * kPi * x);
    return from_primitives(rho, -sy, sx, 0.0, -sy * b0, std::sin(4.0 * kPi * x) * b0, 0.0, pr, 5.0 / 3.0);
}

}  // namespace

extern "C" {

int mhd_case_config_get(int case_id, int n, mhd_case_config* c) {
    if (!c || n < 1) return MHD_ERR_INVALID;
    const double two_pi = 2.0 * kPi;
    auto set = [&](int nx, int ny, double x0, double x1, double y0, double y1, double cfl, double t_end, double g,
                   int l, int r, int b, int t) {
        c->nx = nx * n; c->ny = ny * n; c->ng = 2;
        c->x_min = x0; c->x_max = x1; c->y_min = y0; c->y_max = y1;
        c->cfl = cfl; c->t_end = t_end; c->gamma = g;
        c->bc[0] = l; c->bc[1] = r; c->bc[2] = b; c->bc[3] = t;
    };
    switch (case_id) {
        case MHD_CASE_KELVIN_HELMHOLTZ: set(200, 400, 0.0, 1.0, -1.0, 1.0, 0.3, 0.5, 1.4, 0, 0, 0, 0); break;
        case MHD_CASE_SHOCK_BUBBLE: {
            // t_end = last snapshot of the Haas-Sturtevant schedule (test_cases.cpp:199-225)
            const double vs = 1.22 * std::sqrt(1.4 * 1.01325e5 / 1.29);
            const double t0 = (0.035 - 0.025 - 0.005) / vs, tau = 0.025 / vs;
            set(500, 197, 0.0, 0.225, 0.0, 0.089, 0.4, t0 + 19.0 * tau, 1.4, 1, 1, 1, 1);
            break;
        }
        case MHD_CASE_BRIO_WU: set(800, 4, 0.0, 1.0, 0.0, 4.0 / 800.0, 0.4, 0.1, 2.0, 1, 1, 0, 0); break;
        case MHD_CASE_ORSZAG_TANG: set(192, 192, 0.0, two_pi, 0.0, two_pi, 0.4, two_pi, 5.0 / 3.0, 0, 0, 0, 0); break;
        case MHD_CASE_ORSZAG_TANG_UNIT:
        case MHD_CASE_OT_FAMILY:
        case MHD_CASE_OT_BUMP: set(1024, 1024, 0.0, 1.0, 0.0, 1.0, 0.8, 0.5, 5.0 / 3.0, 0, 0, 0, 0); break;
        case MHD_CASE_SHOCK_CLOUD: set(2048, 1024, 0.0, 2.0, 0.0, 1.0, 0.8, 1.0, 5.0 / 3.0, 1, 1, 1, 1); break;
        case MHD_CASE_BRIO_WU_T: set(800, 4, 0.0, 1.0, 0.0, 4.0 / 800.0, 0.8, 0.1, 2.0, 1, 1, 1, 1); break;
        default: return MHD_ERR_INVALID;
    }
    return MHD_OK;
}

int mhd_init_state_host(int case_id, const mhd_case_config* gm, const mhd_ic_params* p, double* aos) {
    if (!gm || !aos || gm->nx < 1 || gm->ny < 1 || gm->ng != 2) return MHD_ERR_INVALID;
    if ((case_id == MHD_CASE_OT_FAMILY || case_id == MHD_CASE_OT_BUMP) && !p) return MHD_ERR_INVALID;
    const int tnx = gm->nx + 2 * gm->ng, tny = gm->ny + 2 * gm->ng;
    const double dx = (gm->x_max - gm->x_min) / static_cast<double>(gm->nx);
    const double dy = (gm->y_max - gm->y_min) / static_cast<double>(gm->ny);
    for (int j = 0; j < tny; ++j) {
        const double y = gm->y_min + (static_cast<double>(j - gm->ng) + 0.5) * dy;  // init.cpp:13-15
        for (int i = 0; i < tnx; ++i) {
            const double x = gm->x_min + (static_cast<double>(i - gm->ng) + 0.5) * dx;
            State s;
            switch (case_id) {
                case MHD_CASE_KELVIN_HELMHOLTZ: s = kelvin_helmholtz(x, y); break;
                case MHD_CASE_SHOCK_BUBBLE: s = mhd_ic::shock_bubble(x, y); break;
                case MHD_CASE_BRIO_WU:
                case MHD_CASE_BRIO_WU_T: s = mhd_ic::brio_wu(x); break;
                case MHD_CASE_ORSZAG_TANG: s = orszag_tang(x, y); break;
                case MHD_CASE_ORSZAG_TANG_UNIT: s = orszag_tang_unit(x, y, nullptr, false); break;
                case MHD_CASE_SHOCK_CLOUD: s = mhd_ic::shock_cloud(x, y, i, j, p ? p->seed : 0); break;
                case MHD_CASE_OT_FAMILY: s = mhd_ic::ot_family(x, y, *p); break;
                case MHD_CASE_OT_BUMP: s = orszag_tang_unit(x, y, p, true); break;
                default: return MHD_ERR_INVALID;
            }
            std::memcpy(aos + 9 * (static_cast<size_t>(j) * tnx + i), s.q, sizeof s.q);
        }
    }
    return MHD_OK;
}

int mhd_batch_plan_ot_family(uint64_t seed, int n, int* sizes, mhd_ic_params* params) {
    if (n < 1 || !sizes || !params) return MHD_ERR_INVALID;
    uint64_t st = seed;
    auto uniform = [&]() { return static_cast<double>(splitmix_next(st) >> 11) * 0x1.0p-53; };
    if (n == 64) {
        int k = 0;
        for (int c = 0; c < 32; ++c) sizes[k++] = 128;
        for (int c = 0; c < 16; ++c) sizes[k++] = 256;
        for (int c = 0; c < 12; ++c) sizes[k++] = 512;
        for (int c = 0; c < 4; ++c) sizes[k++] = 1024;
        for (int i = 63; i >= 1; --i) {  // Fisher-Yates
            const int j = static_cast<int>(splitmix_next(st) % static_cast<uint64_t>(i + 1));
            const int t = sizes[i];
            sizes[i] = sizes[j];
            sizes[j] = t;
        }
    } else {
        for (int r = 0; r < n; ++r) sizes[r] = 128 << (r % 3);
    }
    for (int r = 0; r < n; ++r) {
        mhd_ic_params& p = params[r];
        std::memset(&p, 0, sizeof p);
        p.gamma = 1.4 + 0.6 * uniform();
        p.b_scale = 0.5 + uniform();
        p.n_modes = 4;
        p.pert_amp = 0.05;
        p.seed = seed;
        for (int m = 0; m < 4; ++m) {
            p.mode_amp[m] = 2.0 * uniform() - 1.0;
            int kx, ky;
            do {
                kx = static_cast<int>(splitmix_next(st) % 9) - 4;
                ky = static_cast<int>(splitmix_next(st) % 9) - 4;
            } while ((kx == 0 && ky == 0) || kx * kx + ky * ky > 16);
            p.mode_kx[m] = kx;
            p.mode_ky[m] = ky;
            p.mode_phase[m] = 2.0 * kPi * uniform();
        }
    }
    return MHD_OK;
}

}  // extern "C"
