// Analytic initial states evaluated from one source on the host (cases.cpp) and
// on the device (init.hip).  Both sides compile it without contraction and use
// only +, -, *, /, sqrt (all correctly rounded) and ic_sin below, so a state
// built on the device is bitwise the state built on the host.
//
// The reference's own cases that call the C library's sin / tanh / exp
// (test_cases.cpp:43-63, 139-148) stay host-only in cases.cpp: a device
// function cannot return the C library's bits, which are not correctly rounded.
#pragma once

#include <cstdint>

#if defined(__HIP__)
#define IC_HD __host__ __device__ inline
#else
#define IC_HD inline
#endif

namespace mhd_ic {

constexpr double kPi = 3.14159265358979323846;  // M_PI (test_cases.cpp:15-17)

struct State {
    double q[9];
};

// make_mhd (test_cases.cpp:24-35): conserved state from primitives, psi = 0.
IC_HD State from_primitives(double rho, double u, double v, double w, double bx, double by, double bz, double p,
                            double gamma) {
    const double b2 = bx * bx + by * by + bz * bz;
    const double ke = 0.5 * rho * (u * u + v * v + w * w);
    const double e = p / (gamma - 1.0) + ke + 0.5 * b2;
    return State{{rho, rho * u, rho * v, rho * w, bx, by, bz, e, 0.0}};
}

// SplitMix64 (Steele, Lea & Flood 2014).
IC_HD uint64_t splitmix_next(uint64_t& s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Counter-based draw for cell (i, j): one mixing step of seed ^ (j << 32 | i); the
// top 53 bits give a uniform double in [0, 1).
IC_HD double uniform_at(uint64_t seed, int i, int j) {
    uint64_t s = seed ^ ((static_cast<uint64_t>(static_cast<uint32_t>(j)) << 32) |
                         static_cast<uint64_t>(static_cast<uint32_t>(i)));
    return static_cast<double>(splitmix_next(s) >> 11) * 0x1.0p-53;
}

// sin(a) for |a| < 2^20 * pi/2, within one ulp.  Cody-Waite reduction
// r = a - k pi/2 with pi/2 split as P1 + P2 + P3 (P1 and P2 carry 33 significant
// bits, so k P1 and k P2 are exact for |k| < 2^20); r is kept as rh + rl (the
// rounding error of the second subtraction recovered by TwoSum).  Then the Taylor
// series of sin or cos on |r| <= pi/4, whose truncation error there is below 1e-21
// relative, with rl and the rounding of 1 - r^2/2 folded into the last addition.
// The coefficients are 1/n! rounded to double.
IC_HD double ic_sin(double a) {
    constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
    constexpr double P1 = 0x1.921fb544p+0, P2 = 0x1.0b4611a6p-34, P3 = 0x1.3198a2e037073p-69;
    const double k = __builtin_rint(a * kTwoOverPi);
    const double t = a - k * P1;  // exact
    const double kp2 = k * P2;    // exact
    const double u = t - kp2;     // TwoSum(t, -kp2) = u + e
    const double bv = u - t;
    const double e = (t - (u - bv)) + (-kp2 - bv);
    const double rl0 = e - k * P3;
    const double rh = u + rl0;
    const double rl = rl0 - (rh - u);
    const double r2 = rh * rh;
    const double hr = 0.5 * r2;
    const long long q = static_cast<long long>(k) & 3;
    double s;
    if ((q & 1) == 0) {
        double p = -0x1.2f49b46814157p-57;  // -1/19!
        p = 0x1.952c77030ad4ap-49 + r2 * p;  // 1/17!
        p = -0x1.ae7f3e733b81fp-41 + r2 * p;
        p = 0x1.6124613a86d09p-33 + r2 * p;
        p = -0x1.ae64567f544e4p-26 + r2 * p;
        p = 0x1.71de3a556c734p-19 + r2 * p;
        p = -0x1.a01a01a01a01ap-13 + r2 * p;
        p = 0x1.1111111111111p-7 + r2 * p;
        p = -0x1.5555555555555p-3 + r2 * p;  // -1/3!
        // sin(rh + rl) = sin rh + rl cos rh, cos rh ~ 1 - r^2/2
        s = rh + ((rl - rl * hr) + rh * (r2 * p));
    } else {
        double p = 0x1.e542ba4020225p-62;  // 1/20!
        p = -0x1.6827863b97d97p-53 + r2 * p;  // -1/18!
        p = 0x1.ae7f3e733b81fp-45 + r2 * p;
        p = -0x1.93974a8c07c9dp-37 + r2 * p;
        p = 0x1.1eed8eff8d898p-29 + r2 * p;
        p = -0x1.27e4fb7789f5cp-22 + r2 * p;
        p = 0x1.a01a01a01a01ap-16 + r2 * p;
        p = -0x1.6c16c16c16c17p-10 + r2 * p;
        p = 0x1.5555555555555p-5 + r2 * p;  // 1/4!
        // cos(rh + rl) = cos rh - rl sin rh; w = 1 - r^2/2 rounded, (1 - w) - hr its
        // rounding error (both differences exact, as w lies in [1/2, 1])
        const double w = 1.0 - hr;
        s = w + ((((1.0 - w) - hr) + r2 * (r2 * p)) - rh * rl);
    }
    return (q & 2) ? -s : s;
}

// Brio-Wu (test_cases.cpp:117-125)
IC_HD State brio_wu(double x) {
    if (x < 0.5) return from_primitives(1.0, 0.0, 0.0, 0.0, 0.75, 1.0, 0.0, 1.0, 2.0);
    return from_primitives(0.125, 0.0, 0.0, 0.0, 0.75, -1.0, 0.0, 0.1, 2.0);
}

// Mach 1.22 shock-bubble, air/helium (test_cases.cpp:76-106)
IC_HD State shock_bubble(double x, double y) {
    const double gamma = 1.4, rho_air = 1.29, rho_he = 0.214, p0 = 1.01325e5, ms = 1.22;
    const double c1 = __builtin_sqrt(gamma * p0 / rho_air);
    const double vs = ms * c1;
    const double rho2 = rho_air * (gamma + 1.0) * ms * ms / ((gamma - 1.0) * ms * ms + 2.0);
    const double p2 = p0 * (2.0 * gamma * ms * ms - (gamma - 1.0)) / (gamma + 1.0);
    const double u2 = vs * (1.0 - rho_air / rho2);
    if (x < 0.005) return from_primitives(rho2, u2, 0.0, 0.0, 0.0, 0.0, 0.0, p2, gamma);
    const double r2 = (x - 0.035) * (x - 0.035) + (y - 0.0445) * (y - 0.0445);
    if (r2 < 0.025 * 0.025) return from_primitives(rho_he, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, p0, gamma);
    return from_primitives(rho_air, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, p0, gamma);
}

// MHD shock-cloud on [0,2]x[0,1]: post-shock slab x < 0.6 moving right into an
// ambient state; a dense bubble r < 0.15 at (0.8, 0.5) in pressure balance with
// +-1 % seeded density noise.
IC_HD State shock_cloud(double x, double y, int i, int j, uint64_t seed) {
    const double gamma = 5.0 / 3.0;
    if (x < 0.6) return from_primitives(3.86859, 11.2536, 0.0, 0.0, 0.0, 2.1826, -2.1826, 167.345, gamma);
    const double r2 = (x - 0.8) * (x - 0.8) + (y - 0.5) * (y - 0.5);
    if (r2 < 0.15 * 0.15) {
        const double rho = 10.0 * (1.0 + 0.01 * (2.0 * uniform_at(seed, i, j) - 1.0));
        return from_primitives(rho, 0.0, 0.0, 0.0, 0.0, 0.56418, 0.56418, 1.0, gamma);
    }
    return from_primitives(1.0, 0.0, 0.0, 0.0, 0.0, 0.56418, 0.56418, 1.0, gamma);
}

// The batch Orszag-Tang family on [0,1]^2: rho = 25/(36 pi) (1 + pert_amp sum_m
// a_m sin(2 pi (kx_m x + ky_m y) + phi_m) / sum_m |a_m|), p = 5/(12 pi),
// u = -sin 2 pi y, v = sin 2 pi x, B = b_scale (-sin 2 pi y, sin 4 pi x) / sqrt(4 pi),
// with gamma and the modes from P (mhd_ic_params or the oracle's copy).  Every sine
// is ic_sin.
template <class P>
IC_HD State ot_family(double x, double y, const P& p) {
    double rho = 25.0 / (36.0 * kPi);
    const double pr = 5.0 / (12.0 * kPi);
    const double b0 = p.b_scale / __builtin_sqrt(4.0 * kPi);
    if (p.n_modes > 0) {
        double s = 0.0, norm = 0.0;
        for (int m = 0; m < p.n_modes; ++m) {
            s += p.mode_amp[m] * ic_sin(2.0 * kPi * (p.mode_kx[m] * x + p.mode_ky[m] * y) + p.mode_phase[m]);
            norm += __builtin_fabs(p.mode_amp[m]);
        }
        rho = rho * (1.0 + p.pert_amp * (s / norm));
    }
    const double sy = ic_sin(2.0 * kPi * y);
    const double sx = ic_sin(2.0 * kPi * x);
    return from_primitives(rho, -sy, sx, 0.0, -sy * b0, ic_sin(4.0 * kPi * x) * b0, 0.0, pr, p.gamma);
}

}  // namespace mhd_ic
