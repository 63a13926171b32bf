// CPU oracle for the 2-D Dedner GLM-MHD MUSCL-Hancock step.  TEST
// INFRASTRUCTURE ONLY (see mhd_oracle.hpp for what pins each function).
//
// Every arithmetic expression keeps the evaluation order of the reference
// (C++ left-to-right association, `a/s` on a 9-vector as `(1/s)*a`,
// std::min/std::max NaN semantics), because the goal is bitwise equality with
// the reference CPU path.  The physics is templated on the scalar type so that
// the same source, instantiated with a forward-mode dual number, gives the
// tangent-linear model with branches frozen at the primal values.
#include "mhd_oracle.hpp"

#include "../mhd_gpu_acc_amd/csrc/ic.hpp"  // the batch family's initial state (ic_sin)

#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPi = 3.14159265358979323846;  // M_PI, src/test_cases.cpp:15-17

// ------------------------------------------------------------------------------------
// Scalar layer: double and a forward-mode dual number.
// ------------------------------------------------------------------------------------
struct Dual {
    double v, d;
    Dual() : v(0.0), d(0.0) {}
    Dual(double x) : v(x), d(0.0) {}  // NOLINT: implicit from constants
    Dual(double x, double dx) : v(x), d(dx) {}
};
inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(Dual a, Dual b) {
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
inline bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
inline bool operator>=(Dual a, Dual b) { return a.v >= b.v; }

inline double val(double x) { return x; }
inline double val(Dual x) { return x.v; }

inline double s_sqrt(double x) { return std::sqrt(x); }
// The derivative is taken as 0 where sqrt is 0 (the clamped discriminant of the
// fast speed), the convention the hand-written adjoint shares.
inline Dual s_sqrt(Dual x) {
    const double r = std::sqrt(x.v);
    return Dual(r, r > 0.0 ? x.d / (2.0 * r) : 0.0);
}
inline double s_fabs(double x) { return std::fabs(x); }
inline Dual s_fabs(Dual x) { return x.v < 0.0 ? -x : (x.v > 0.0 ? x : Dual(std::fabs(x.v), 0.0)); }
// C fmin/fmax: a NaN argument yields the other one.
inline double s_fmin(double a, double b) { return std::fmin(a, b); }
inline double s_fmax(double a, double b) { return std::fmax(a, b); }
inline Dual s_fmin(Dual a, Dual b) {
    const double m = std::fmin(a.v, b.v);
    return (m == a.v && !(std::isnan(a.v))) ? a : b;
}
inline Dual s_fmax(Dual a, Dual b) {
    const double m = std::fmax(a.v, b.v);
    return (m == a.v && !(std::isnan(a.v))) ? a : b;
}
// Frozen-branch finite differences: a value of the recorded trajectory (v) and N
// perturbed copies (p) carried through the same operations.  Every comparison,
// selection and finiteness test reads v, so the copies follow the branches of the
// recorded trajectory exactly: (p_+ - p_-) / 2 eps is a central difference with
// every branch frozen, which converges to the tangent of the Dual instantiation.
template <int N> struct Shadow {
    double v;
    double p[N];
    Shadow() : v(0.0) { for (double& x : p) x = 0.0; }
    Shadow(double x) : v(x) { for (double& y : p) y = x; }  // NOLINT: implicit from constants
};
#define MHD_SHADOW_BINOP(op)                                                    \
    template <int N> inline Shadow<N> operator op(const Shadow<N>& a, const Shadow<N>& b) { \
        Shadow<N> r;                                                            \
        r.v = a.v op b.v;                                                       \
        for (int k = 0; k < N; ++k) r.p[k] = a.p[k] op b.p[k];                  \
        return r;                                                               \
    }
MHD_SHADOW_BINOP(+)
MHD_SHADOW_BINOP(-)
MHD_SHADOW_BINOP(*)
MHD_SHADOW_BINOP(/)
#undef MHD_SHADOW_BINOP
template <int N> inline Shadow<N> operator-(const Shadow<N>& a) {
    Shadow<N> r;
    r.v = -a.v;
    for (int k = 0; k < N; ++k) r.p[k] = -a.p[k];
    return r;
}
template <int N> inline bool operator<(const Shadow<N>& a, const Shadow<N>& b) { return a.v < b.v; }
template <int N> inline bool operator>(const Shadow<N>& a, const Shadow<N>& b) { return a.v > b.v; }
template <int N> inline bool operator<=(const Shadow<N>& a, const Shadow<N>& b) { return a.v <= b.v; }
template <int N> inline bool operator>=(const Shadow<N>& a, const Shadow<N>& b) { return a.v >= b.v; }
template <int N> inline double val(const Shadow<N>& x) { return x.v; }
template <int N> inline Shadow<N> s_sqrt(const Shadow<N>& x) {
    Shadow<N> r;
    r.v = std::sqrt(x.v);
    for (int k = 0; k < N; ++k) r.p[k] = std::sqrt(x.p[k]);
    return r;
}
template <int N> inline Shadow<N> s_fabs(const Shadow<N>& x) {
    if (x.v < 0.0) return -x;
    if (x.v > 0.0) return x;
    Shadow<N> r;
    r.v = std::fabs(x.v);
    for (int k = 0; k < N; ++k) r.p[k] = std::fabs(x.p[k]);
    return r;
}
template <int N> inline Shadow<N> s_fmin(const Shadow<N>& a, const Shadow<N>& b) {
    const double m = std::fmin(a.v, b.v);
    return (m == a.v && !(std::isnan(a.v))) ? a : b;
}
template <int N> inline Shadow<N> s_fmax(const Shadow<N>& a, const Shadow<N>& b) {
    const double m = std::fmax(a.v, b.v);
    return (m == a.v && !(std::isnan(a.v))) ? a : b;
}

// std::min / std::max: (b < a) ? b : a and (a < b) ? b : a.
template <class R> inline R s_min(R a, R b) { return (b < a) ? b : a; }
template <class R> inline R s_max(R a, R b) { return (a < b) ? b : a; }
template <class R> inline bool s_finite(R x) { return std::isfinite(val(x)); }

// ------------------------------------------------------------------------------------
// 9-vectors (head/types.hpp:16-182)
// ------------------------------------------------------------------------------------
enum : int { RHO = 0, MX = 1, MY = 2, MZ = 3, BXI = 4, BYI = 5, BZI = 6, EN = 7, PSI = 8 };
// Primitive slots: 1..3 velocities, 7 pressure.
enum : int { VU = 1, VV = 2, VW = 3, PR = 7 };

template <class R> struct V9 {
    R q[9];
    R& operator[](int k) { return q[k]; }
    const R& operator[](int k) const { return q[k]; }
};

// Axis-dependent slots: normal / tangential momentum (velocity) and field.
struct Axis {
    int n, t, bn, bt;
};
constexpr Axis kAxisX{MX, MY, BXI, BYI};
constexpr Axis kAxisY{MY, MX, BYI, BXI};
inline Axis axis_of(int dir) { return dir == 0 ? kAxisX : kAxisY; }

template <class R> inline V9<R> load9(const double* a) {
    V9<R> v;
    for (int k = 0; k < 9; ++k) v[k] = R(a[k]);
    return v;
}
template <class R> inline void store9(const V9<R>& v, double* a) {
    for (int k = 0; k < 9; ++k) a[k] = val(v[k]);
}

// ------------------------------------------------------------------------------------
// Physics (head/physics.hpp)
// ------------------------------------------------------------------------------------
// physics.hpp:69-78
template <class R> inline V9<R> cons_to_prim(const V9<R>& U, double gamma) {
    const R inv_rho = R(1.0) / U[RHO];
    const R ux = U[MX] * inv_rho;
    const R uy = U[MY] * inv_rho;
    const R uz = U[MZ] * inv_rho;
    const R b2 = U[BXI] * U[BXI] + U[BYI] * U[BYI] + U[BZI] * U[BZI];
    const R ke = R(0.5) * U[RHO] * (ux * ux + uy * uy + uz * uz);
    const R p = R(gamma - 1.0) * (U[EN] - ke - R(0.5) * b2);
    V9<R> V;
    V[RHO] = U[RHO]; V[VU] = ux; V[VV] = uy; V[VW] = uz;
    V[BXI] = U[BXI]; V[BYI] = U[BYI]; V[BZI] = U[BZI]; V[PR] = p; V[PSI] = U[PSI];
    return V;
}

// physics.hpp:83-89
template <class R> inline V9<R> prim_to_cons(const V9<R>& V, double gamma) {
    const R b2 = V[BXI] * V[BXI] + V[BYI] * V[BYI] + V[BZI] * V[BZI];
    const R ke = R(0.5) * V[RHO] * (V[VU] * V[VU] + V[VV] * V[VV] + V[VW] * V[VW]);
    const R E = V[PR] / R(gamma - 1.0) + ke + R(0.5) * b2;
    V9<R> U;
    U[RHO] = V[RHO]; U[MX] = V[RHO] * V[VU]; U[MY] = V[RHO] * V[VV]; U[MZ] = V[RHO] * V[VW];
    U[BXI] = V[BXI]; U[BYI] = V[BYI]; U[BZI] = V[BZI]; U[EN] = E; U[PSI] = V[PSI];
    return U;
}

// physics.hpp:105-164, written once for both axes: the x and y fluxes are the
// same expressions with (normal, tangential) slots exchanged.
template <class R> inline V9<R> flux(const V9<R>& U, Axis a, R ch, double gamma) {
    const R inv_rho = R(1.0) / U[RHO];
    const R ux = U[MX] * inv_rho;
    const R uy = U[MY] * inv_rho;
    const R uz = U[MZ] * inv_rho;
    const R b2 = U[BXI] * U[BXI] + U[BYI] * U[BYI] + U[BZI] * U[BZI];
    const R ke = R(0.5) * U[RHO] * (ux * ux + uy * uy + uz * uz);
    const R p = R(gamma - 1.0) * (U[EN] - ke - R(0.5) * b2);
    const R p_tot = p + R(0.5) * b2;
    const R bdotu = U[BXI] * ux + U[BYI] * uy + U[BZI] * uz;
    const R un = (a.n == MX) ? ux : uy;
    const R ut = (a.n == MX) ? uy : ux;
    const R bn = U[a.bn], bt = U[a.bt];
    V9<R> F;
    F[RHO] = U[a.n];
    F[a.n] = U[a.n] * un + p_tot - bn * bn;
    F[a.t] = U[a.n] * ut - bn * bt;
    F[MZ] = U[a.n] * uz - bn * U[BZI];
    F[a.bn] = U[PSI];
    F[a.bt] = bt * un - bn * ut;
    F[BZI] = U[BZI] * un - bn * uz;
    F[EN] = (U[EN] + p_tot) * un - bn * bdotu;
    F[PSI] = ch * ch * bn;
    return F;
}

// physics.hpp:175-190
template <class R> inline R fast_speed(const V9<R>& V, Axis a, double gamma) {
    const R a2 = R(gamma) * V[PR] / V[RHO];
    const R b2 = (V[BXI] * V[BXI] + V[BYI] * V[BYI] + V[BZI] * V[BZI]) / V[RHO];
    const R bn2 = V[a.bn] * V[a.bn] / V[RHO];
    const R disc = (a2 + b2) * (a2 + b2) - R(4.0) * a2 * bn2;
    return s_sqrt(R(0.5) * (a2 + b2 + s_sqrt(disc > R(0.0) ? disc : R(0.0))));
}

// physics.hpp:202-210
template <class R> inline R max_signal_speed(const V9<R>& V, Axis a, R ch, double gamma) {
    const R cf = fast_speed(V, a, gamma);
    return s_fmax(s_fabs(V[a.n]) + cf, ch);
}

// riemann.hpp:54-57
template <class R> inline bool prim_physical(const V9<R>& V) {
    return s_finite(V[RHO]) && s_finite(V[PR]) && V[RHO] > R(0.0) && V[PR] > R(0.0);
}

template <class R> inline V9<R> half_sum(const V9<R>& a, const V9<R>& b) {
    V9<R> r;
    for (int k = 0; k < 9; ++k) r[k] = R(0.5) * (a[k] + b[k]);
    return r;
}

// Star state in rotated components (riemann.hpp:219-230, 417-437).
template <class R>
inline V9<R> build_state(Axis a, R rhos, R uns, R uts, R uws, R bns, R bts, R bws, R es, R psis) {
    V9<R> U;
    U[RHO] = rhos;
    U[a.n] = rhos * uns;
    U[a.t] = rhos * uts;
    U[MZ] = rhos * uws;
    U[a.bn] = bns;
    U[a.bt] = bts;
    U[BZI] = bws;
    U[EN] = es;
    U[PSI] = psis;
    return U;
}

// riemann.hpp:70-107
template <class R> V9<R> hll_flux(const V9<R>& UL, const V9<R>& UR, Axis a, R ch, double gamma) {
    const V9<R> VL = cons_to_prim(UL, gamma);
    const V9<R> VR = cons_to_prim(UR, gamma);
    const V9<R> FL = flux(UL, a, ch, gamma);
    const V9<R> FR = flux(UR, a, ch, gamma);
    if (!prim_physical(VL) || !prim_physical(VR)) return half_sum(FL, FR);
    const R cfL = fast_speed(VL, a, gamma);
    const R cfR = fast_speed(VR, a, gamma);
    const R unL = VL[a.n], unR = VR[a.n];
    const R SL = s_fmin(s_fmin(unL - cfL, unR - cfR), -ch);
    const R SR = s_fmax(s_fmax(unL + cfL, unR + cfR), ch);
    if (SL >= R(0.0)) return FL;
    if (SR <= R(0.0)) return FR;
    const R denom = SR - SL;
    if (s_fabs(denom) < R(1.0e-14)) return half_sum(FL, FR);
    const R inv = R(1.0) / denom;
    V9<R> F;
    for (int k = 0; k < 9; ++k) F[k] = inv * (SR * FL[k] - SL * FR[k] + (SL * SR) * (UR[k] - UL[k]));
    return F;
}

// riemann.hpp:122-241
template <class R> V9<R> hllc_flux(const V9<R>& UL, const V9<R>& UR, Axis a, R ch, double gamma) {
    const V9<R> VL = cons_to_prim(UL, gamma);
    const V9<R> VR = cons_to_prim(UR, gamma);
    if (!prim_physical(VL) || !prim_physical(VR)) return hll_flux(UL, UR, a, ch, gamma);
    const R rhoL = VL[RHO], rhoR = VR[RHO];
    const R pL = VL[PR], pR = VR[PR];
    const R unL = VL[a.n], unR = VR[a.n], utL = VL[a.t], utR = VR[a.t];
    const R uwL = VL[VW], uwR = VR[VW];
    const R BnL = VL[a.bn], BnR = VR[a.bn], BtL = VL[a.bt], BtR = VR[a.bt];
    const R BwL = VL[BZI], BwR = VR[BZI];
    const R cfL = fast_speed(VL, a, gamma);
    const R cfR = fast_speed(VR, a, gamma);
    const R SL = s_fmin(unL - cfL, unR - cfR);
    const R SR = s_fmax(unL + cfL, unR + cfR);
    const V9<R> FL = flux(UL, a, ch, gamma);
    const V9<R> FR = flux(UR, a, ch, gamma);
    if (SL >= R(0.0)) return FL;
    if (SR <= R(0.0)) return FR;
    const R bL2 = BnL * BnL + BtL * BtL + BwL * BwL;
    const R bR2 = BnR * BnR + BtR * BtR + BwR * BwR;
    const R ptL = pL + R(0.5) * bL2;
    const R ptR = pR + R(0.5) * bR2;
    const R numerSM = (SR - unR) * rhoR * unR - (SL - unL) * rhoL * unL + ptL - ptR;
    const R denomSM = (SR - unR) * rhoR - (SL - unL) * rhoL;
    if (s_fabs(denomSM) < R(1.0e-14)) return hll_flux(UL, UR, a, ch, gamma);
    const R SM = numerSM / denomSM;
    const R rhoLs = rhoL * (SL - unL) / (SL - SM);
    const R rhoRs = rhoR * (SR - unR) / (SR - SM);
    if (rhoLs <= R(0.0) || rhoRs <= R(0.0) || !s_finite(rhoLs) || !s_finite(rhoRs))
        return hll_flux(UL, UR, a, ch, gamma);
    const R ptLs = ptL + rhoL * (SL - unL) * (SM - unL);
    const R ptRs = ptR + rhoR * (SR - unR) * (SM - unR);
    const R psiM = R(0.5) * (VL[PSI] + VR[PSI]) - R(0.5) * ch * (BnR - BnL);
    const R bdotuL = BnL * unL + BtL * utL + BwL * uwL;
    const R bdotuLs = BnL * SM + BtL * utL + BwL * uwL;
    const R bdotuR = BnR * unR + BtR * utR + BwR * uwR;
    const R bdotuRs = BnR * SM + BtR * utR + BwR * uwR;
    const R ELs = ((SL - unL) * UL[EN] - ptL * unL + ptLs * SM + BnL * (bdotuLs - bdotuL)) / (SL - SM);
    const R ERs = ((SR - unR) * UR[EN] - ptR * unR + ptRs * SM + BnR * (bdotuRs - bdotuR)) / (SR - SM);
    V9<R> F;
    if (SM >= R(0.0)) {
        const V9<R> Us = build_state(a, rhoLs, SM, utL, uwL, BnL, BtL, BwL, ELs, psiM);
        for (int k = 0; k < 9; ++k) F[k] = FL[k] + SL * (Us[k] - UL[k]);
    } else {
        const V9<R> Us = build_state(a, rhoRs, SM, utR, uwR, BnR, BtR, BwR, ERs, psiM);
        for (int k = 0; k < 9; ++k) F[k] = FR[k] + SR * (Us[k] - UR[k]);
    }
    return F;
}

// riemann.hpp:253-480 (Miyoshi-Kusano HLLD with GLM: B_n may differ left/right)
template <class R> V9<R> hlld_flux(const V9<R>& UL, const V9<R>& UR, Axis a, R ch, double gamma) {
    const V9<R> VL = cons_to_prim(UL, gamma);
    const V9<R> VR = cons_to_prim(UR, gamma);
    if (!prim_physical(VL) || !prim_physical(VR)) return hll_flux(UL, UR, a, ch, gamma);
    const R rhoL = VL[RHO], rhoR = VR[RHO];
    const R pL = VL[PR], pR = VR[PR];
    const R unL = VL[a.n], unR = VR[a.n], utL = VL[a.t], utR = VR[a.t];
    const R uwL = VL[VW], uwR = VR[VW];
    const R BnL = VL[a.bn], BnR = VR[a.bn], BtL = VL[a.bt], BtR = VR[a.bt];
    const R BwL = VL[BZI], BwR = VR[BZI];
    const R cfL = fast_speed(VL, a, gamma);
    const R cfR = fast_speed(VR, a, gamma);
    const R SL = s_fmin(unL - cfL, unR - cfR);
    const R SR = s_fmax(unL + cfL, unR + cfR);
    if (s_fabs(SR - SL) < R(1.0e-14)) return hll_flux(UL, UR, a, ch, gamma);
    const R bL2 = BnL * BnL + BtL * BtL + BwL * BwL;
    const R bR2 = BnR * BnR + BtR * BtR + BwR * BwR;
    const R ptL = pL + R(0.5) * bL2;
    const R ptR = pR + R(0.5) * bR2;
    const R numerSM = (SR - unR) * rhoR * unR - (SL - unL) * rhoL * unL + ptL - ptR;
    const R denomSM = (SR - unR) * rhoR - (SL - unL) * rhoL;
    if (s_fabs(denomSM) < R(1.0e-14)) return hll_flux(UL, UR, a, ch, gamma);
    const R SM = numerSM / denomSM;
    const R rhoLs = rhoL * (SL - unL) / (SL - SM);
    const R rhoRs = rhoR * (SR - unR) / (SR - SM);
    if (rhoLs <= R(0.0) || rhoRs <= R(0.0) || !s_finite(rhoLs) || !s_finite(rhoRs))
        return hll_flux(UL, UR, a, ch, gamma);
    const R ptstar = ptL + rhoL * (SL - unL) * (SM - unL);
    const R sqL = s_sqrt(rhoLs);
    const R sqR = s_sqrt(rhoRs);
    const R SLs = SM - s_fabs(BnL) / sqL;
    const R SRs = SM + s_fabs(BnR) / sqR;

    R utLs, uwLs, BtLs, BwLs;
    {
        const R d = rhoL * (SL - unL) * (SL - SM) - BnL * BnL;
        if (s_fabs(d) < R(1.0e-14)) {
            utLs = utL; uwLs = uwL; BtLs = BtL; BwLs = BwL;
        } else {
            utLs = utL - BnL * BtL * (SM - unL) / d;
            uwLs = uwL - BnL * BwL * (SM - unL) / d;
            BtLs = BtL * (rhoL * (SL - unL) * (SL - unL) - BnL * BnL) / d;
            BwLs = BwL * (rhoL * (SL - unL) * (SL - unL) - BnL * BnL) / d;
        }
    }
    const R bdotuL = BnL * unL + BtL * utL + BwL * uwL;
    const R bdotuLs = BnL * SM + BtLs * utLs + BwLs * uwLs;
    const R ELs = (UL[EN] * (SL - unL) - ptL * unL + ptstar * SM + BnL * (bdotuLs - bdotuL)) / (SL - SM);

    R utRs, uwRs, BtRs, BwRs;
    {
        const R d = rhoR * (SR - unR) * (SR - SM) - BnR * BnR;
        if (s_fabs(d) < R(1.0e-14)) {
            utRs = utR; uwRs = uwR; BtRs = BtR; BwRs = BwR;
        } else {
            utRs = utR - BnR * BtR * (SM - unR) / d;
            uwRs = uwR - BnR * BwR * (SM - unR) / d;
            BtRs = BtR * (rhoR * (SR - unR) * (SR - unR) - BnR * BnR) / d;
            BwRs = BwR * (rhoR * (SR - unR) * (SR - unR) - BnR * BnR) / d;
        }
    }
    const R bdotuR = BnR * unR + BtR * utR + BwR * uwR;
    const R bdotuRs = BnR * SM + BtRs * utRs + BwRs * uwRs;
    const R ERs = (UR[EN] * (SR - unR) - ptR * unR + ptstar * SM + BnR * (bdotuRs - bdotuR)) / (SR - SM);

    // Sign of B_n is taken from the left state (riemann.hpp:376).
    const R signBn = (BnL >= R(0.0)) ? R(1.0) : R(-1.0);
    R utss, uwss, BtLss, BwLss, BtRss, BwRss, utLss, uwLss, utRss, uwRss, ELss, ERss;
    if (s_fabs(sqL + sqR) < R(1.0e-14)) {
        utLss = utLs; uwLss = uwLs; BtLss = BtLs; BwLss = BwLs; ELss = ELs;
        utRss = utRs; uwRss = uwRs; BtRss = BtRs; BwRss = BwRs; ERss = ERs;
    } else {
        const R inv_d2 = R(1.0) / (sqL + sqR);
        utss = (sqL * utLs + sqR * utRs + signBn * (BtRs - BtLs)) * inv_d2;
        uwss = (sqL * uwLs + sqR * uwRs + signBn * (BwRs - BwLs)) * inv_d2;
        const R Btss = (sqL * BtRs + sqR * BtLs + signBn * sqL * sqR * (utRs - utLs)) * inv_d2;
        const R Bwss = (sqL * BwRs + sqR * BwLs + signBn * sqL * sqR * (uwRs - uwLs)) * inv_d2;
        utLss = utRss = utss;
        uwLss = uwRss = uwss;
        BtLss = BtRss = Btss;
        BwLss = BwRss = Bwss;
        // riemann.hpp:398-403 assigns E** once more and then overwrites it
        // (:406-412); only the final form is kept.
        const R bdotuLss = BnL * SM + Btss * utss + Bwss * uwss;
        ELss = ELs - sqL * signBn * (bdotuLss - bdotuLs);
        const R bdotuRss = BnR * SM + Btss * utss + Bwss * uwss;
        ERss = ERs + sqR * signBn * (bdotuRss - bdotuRs);
    }

    const R psiM = R(0.5) * (VL[PSI] + VR[PSI]) - R(0.5) * ch * (BnR - BnL);
    const V9<R> FL = flux(UL, a, ch, gamma);
    const V9<R> FR = flux(UR, a, ch, gamma);
    if (SL >= R(0.0)) return FL;
    if (SR <= R(0.0)) return FR;
    V9<R> F;
    if (SM >= R(0.0)) {
        const V9<R> Us = build_state(a, rhoLs, SM, utLs, uwLs, BnL, BtLs, BwLs, ELs, psiM);
        if (SLs >= R(0.0)) {
            for (int k = 0; k < 9; ++k) F[k] = FL[k] + SL * (Us[k] - UL[k]);
        } else {
            const V9<R> Uss = build_state(a, rhoLs, SM, utLss, uwLss, BnL, BtLss, BwLss, ELss, psiM);
            for (int k = 0; k < 9; ++k) F[k] = FL[k] + SL * (Us[k] - UL[k]) + SLs * (Uss[k] - Us[k]);
        }
    } else {
        const V9<R> Us = build_state(a, rhoRs, SM, utRs, uwRs, BnR, BtRs, BwRs, ERs, psiM);
        if (SRs <= R(0.0)) {
            for (int k = 0; k < 9; ++k) F[k] = FR[k] + SR * (Us[k] - UR[k]);
        } else {
            const V9<R> Uss = build_state(a, rhoRs, SM, utRss, uwRss, BnR, BtRss, BwRss, ERss, psiM);
            for (int k = 0; k < 9; ++k) F[k] = FR[k] + SR * (Us[k] - UR[k]) + SRs * (Uss[k] - Us[k]);
        }
    }
    return F;
}

// riemann.hpp:486-525
template <class R> V9<R> force_flux(const V9<R>& UL, const V9<R>& UR, Axis a, R ch, double gamma) {
    const V9<R> VL = cons_to_prim(UL, gamma);
    const V9<R> VR = cons_to_prim(UR, gamma);
    const V9<R> FL = flux(UL, a, ch, gamma);
    const V9<R> FR = flux(UR, a, ch, gamma);
    if (!prim_physical(VL) || !prim_physical(VR)) return half_sum(FL, FR);
    const R cfL = fast_speed(VL, a, gamma);
    const R cfR = fast_speed(VR, a, gamma);
    const R unL = VL[a.n], unR = VR[a.n];
    const R alpha = s_fmax(s_fmax(s_fabs(unL) + cfL, s_fabs(unR) + cfR), ch);
    if (alpha < R(1.0e-14)) return half_sum(FL, FR);
    V9<R> Flf, Uri;
    const R ha = R(0.5) * alpha;
    const R hia = R(0.5) * (R(1.0) / alpha);
    for (int k = 0; k < 9; ++k) {
        Flf[k] = R(0.5) * (FL[k] + FR[k]) - ha * (UR[k] - UL[k]);
        Uri[k] = R(0.5) * (UL[k] + UR[k]) - hia * (FR[k] - FL[k]);
    }
    if (!prim_physical(cons_to_prim(Uri, gamma))) return Flf;
    const V9<R> Fri = flux(Uri, a, ch, gamma);
    return half_sum(Flf, Fri);
}

// riemann.hpp:544-563
template <class R> void apply_glm(V9<R>& F, const V9<R>& UL, const V9<R>& UR, Axis a, R ch) {
    if (ch <= R(0.0)) return;
    const R BnL = UL[a.bn], BnR = UR[a.bn];
    const R psiL = UL[PSI], psiR = UR[PSI];
    const R bn_star = R(0.5) * (BnL + BnR) - R(0.5) / ch * (psiR - psiL);
    const R psi_star = R(0.5) * (psiL + psiR) - R(0.5) * ch * (BnR - BnL);
    F[a.bn] = psi_star;
    F[PSI] = ch * ch * bn_star;
}

// riemann.hpp:568-592
template <class R>
V9<R> riemann(const V9<R>& UL, const V9<R>& UR, Axis a, int solver, R ch, double gamma) {
    if (solver == 0) return hll_flux(UL, UR, a, ch, gamma);
    V9<R> F = (solver == 1)   ? hllc_flux(UL, UR, a, ch, gamma)
              : (solver == 2) ? hlld_flux(UL, UR, a, ch, gamma)
                              : force_flux(UL, UR, a, ch, gamma);
    // The finiteness test looks at rho, E, Bx and psi, whatever the axis.
    if (!s_finite(F[RHO]) || !s_finite(F[EN]) || !s_finite(F[BXI]) || !s_finite(F[PSI]))
        return hll_flux(UL, UR, a, ch, gamma);
    apply_glm(F, UL, UR, a, ch);
    return F;
}

// ------------------------------------------------------------------------------------
// MUSCL-Hancock (src/cpu/solver_cpu.cpp)
// ------------------------------------------------------------------------------------
constexpr double kFloor = 1.0e-12;  // solver_cpu.cpp:19-20

// solver_cpu.cpp:26-29 (minmod).  Van Leer is not in the reference: it is the
// harmonic form 2ab/(a+b) on same-sign slopes, with the minmod zero rule.
template <class R> inline R limit(R a, R b, int limiter) {
    if (a * b <= R(0.0)) return R(0.0);
    if (limiter == 0) return (a > R(0.0)) ? s_min(a, b) : s_max(a, b);
    return R(2.0) * a * b / (a + b);
}

// solver_cpu.cpp:49-56
template <class R> inline bool is_physical(const V9<R>& V) {
    return s_finite(V[RHO]) && V[RHO] > R(kFloor) && s_finite(V[PR]) && V[PR] > R(kFloor) &&
           s_finite(V[VU]) && s_finite(V[VV]) && s_finite(V[VW]) && s_finite(V[BXI]) &&
           s_finite(V[BYI]) && s_finite(V[BZI]) && s_finite(V[PSI]);
}

// solver_cpu.cpp:88-119
// Derivative convention (not the value): when both one-sided differences of a
// component are below rounding level -- |a| <= 2^-44 (|Wc| + |Wm|) and likewise b --
// their signs are noise (a uniform field recovered from conserved variables), the
// limiter's derivative would depend only on the noise ratio, and no finite
// difference can see it; the slope is then differentiated as the zero slope.
constexpr double kSlopeNoise = 0x1p-44;
inline bool slope_noise(double a, double b, double wm, double wc, double wp) {
    return std::fabs(a) <= kSlopeNoise * (std::fabs(wc) + std::fabs(wm)) &&
           std::fabs(b) <= kSlopeNoise * (std::fabs(wp) + std::fabs(wc));
}
inline double freeze(double s) { return s; }
inline Dual freeze(Dual s) { return Dual(s.v, 0.0); }
template <int N> inline Shadow<N> freeze(const Shadow<N>& s) { return Shadow<N>(s.v); }

template <class R>
inline void reconstruct(const V9<R>& Wm, const V9<R>& Wc, const V9<R>& Wp, R dtd, Axis a, R ch,
                        double gamma, int limiter, V9<R>& Ul_star, V9<R>& Ur_star) {
    V9<R> Wl, Wr;
    for (int k = 0; k < 9; ++k) {
        R s = limit(Wc[k] - Wm[k], Wp[k] - Wc[k], limiter);
        if (slope_noise(val(Wc[k] - Wm[k]), val(Wp[k] - Wc[k]), val(Wm[k]), val(Wc[k]), val(Wp[k])))
            s = freeze(s);
        Wl[k] = Wc[k] - R(0.5) * s;
        Wr[k] = Wc[k] + R(0.5) * s;
    }
    if (!is_physical(Wl)) Wl = Wc;
    if (!is_physical(Wr)) Wr = Wc;
    const V9<R> Ul = prim_to_cons(Wl, gamma);
    const V9<R> Ur = prim_to_cons(Wr, gamma);
    const V9<R> Fl = flux(Ul, a, ch, gamma);
    const V9<R> Fr = flux(Ur, a, ch, gamma);
    const R hd = R(0.5) * dtd;
    V9<R> cl, cr;
    for (int k = 0; k < 9; ++k) {
        const R h = hd * (Fr[k] - Fl[k]);
        cl[k] = Ul[k] - h;
        cr[k] = Ur[k] - h;
    }
    Ul_star = is_physical(cons_to_prim(cl, gamma)) ? cl : Ul;
    Ur_star = is_physical(cons_to_prim(cr, gamma)) ? cr : Ur;
}

// ------------------------------------------------------------------------------------
// Grids (head/cpu/grid_cpu.hpp): AoS, row-major, ghost cells of width ng.
// ------------------------------------------------------------------------------------
struct Geo {
    int nx, ny, ng, tnx, tny;
    double dx, dy;
    int ib, ie, jb, je;
    explicit Geo(const orc_geom& g)
        : nx(g.nx), ny(g.ny), ng(g.ng), tnx(g.nx + 2 * g.ng), tny(g.ny + 2 * g.ng),
          dx((g.x_max - g.x_min) / static_cast<double>(g.nx)),
          dy((g.y_max - g.y_min) / static_cast<double>(g.ny)),
          ib(g.ng), ie(g.ng + g.nx), jb(g.ng), je(g.ng + g.ny) {}
    std::size_t at(int i, int j) const { return static_cast<std::size_t>(j) * tnx + i; }
    std::size_t cells() const { return static_cast<std::size_t>(tnx) * tny; }
};

template <class R> using Field = std::vector<V9<R>>;

// boundary_cpu.hpp:12-32
template <class R> void copy_ghost(const Field<R>& src, Field<R>& dst, const Geo& g) {
    for (int j = 0; j < g.tny; ++j)
        for (int h = 0; h < g.ng; ++h) {
            dst[g.at(g.ib - 1 - h, j)] = src[g.at(g.ib - 1 - h, j)];
            dst[g.at(g.ie + h, j)] = src[g.at(g.ie + h, j)];
        }
    for (int i = g.ib; i < g.ie; ++i)
        for (int h = 0; h < g.ng; ++h) {
            dst[g.at(i, g.jb - 1 - h)] = src[g.at(i, g.jb - 1 - h)];
            dst[g.at(i, g.je + h)] = src[g.at(i, g.je + h)];
        }
}

// boundary_cpu.hpp:45-149.  Types: 0 periodic, 1 transmissive, 2 reflecting,
// 3 Dirichlet (left untouched).  Order: left, right (all rows), bottom, top
// (all columns).
template <class R> void boundary(Field<R>& U, const Geo& g, const int* bc) {
    auto reflect = [](V9<R> m, int mom, int fld) {
        m[mom] = -m[mom];
        m[fld] = -m[fld];
        return m;
    };
    for (int j = 0; j < g.tny; ++j)
        for (int h = 0; h < g.ng; ++h) {
            V9<R>& dl = U[g.at(g.ib - 1 - h, j)];
            if (bc[0] == 0) dl = U[g.at(g.ie - 1 - h, j)];
            else if (bc[0] == 1) dl = U[g.at(g.ib, j)];
            else if (bc[0] == 2) dl = reflect(U[g.at(g.ib + h, j)], MX, BXI);
        }
    for (int j = 0; j < g.tny; ++j)
        for (int h = 0; h < g.ng; ++h) {
            V9<R>& dr = U[g.at(g.ie + h, j)];
            if (bc[1] == 0) dr = U[g.at(g.ib + h, j)];
            else if (bc[1] == 1) dr = U[g.at(g.ie - 1, j)];
            else if (bc[1] == 2) dr = reflect(U[g.at(g.ie - 1 - h, j)], MX, BXI);
        }
    for (int i = 0; i < g.tnx; ++i)
        for (int h = 0; h < g.ng; ++h) {
            V9<R>& db = U[g.at(i, g.jb - 1 - h)];
            if (bc[2] == 0) db = U[g.at(i, g.je - 1 - h)];
            else if (bc[2] == 1) db = U[g.at(i, g.jb)];
            else if (bc[2] == 2) db = reflect(U[g.at(i, g.jb + h)], MY, BYI);
        }
    for (int i = 0; i < g.tnx; ++i)
        for (int h = 0; h < g.ng; ++h) {
            V9<R>& dt = U[g.at(i, g.je + h)];
            if (bc[3] == 0) dt = U[g.at(i, g.jb + h)];
            else if (bc[3] == 1) dt = U[g.at(i, g.je - 1)];
            else if (bc[3] == 2) dt = reflect(U[g.at(i, g.je - 1 - h)], MY, BYI);
        }
}

// solver_cpu.cpp:283-312.  ch is read as the previous c_h and replaced by the
// new maximum (a ratchet: the new value is never below the old one).
template <class R> int compute_dt(const Field<R>& U, const Geo& g, double cfl, double gamma, double& ch,
                                  double& dt) {
    double mx = 0.0;
    const double chp = ch;
#pragma omp parallel for collapse(2) reduction(max : mx) schedule(static)
    for (int j = g.jb; j < g.je; ++j)
        for (int i = g.ib; i < g.ie; ++i) {
            V9<double> Ud;
            for (int k = 0; k < 9; ++k) Ud[k] = val(U[g.at(i, j)][k]);
            const V9<double> V = cons_to_prim(Ud, gamma);
            if (!std::isfinite(V[RHO]) || !std::isfinite(V[PR]) || V[RHO] <= 0.0 || V[PR] <= 0.0) continue;
            const double sx = max_signal_speed(V, kAxisX, chp, gamma);
            const double sy = max_signal_speed(V, kAxisY, chp, gamma);
            if (std::isfinite(sx) && std::isfinite(sy)) mx = s_max(mx, s_max(sx, sy));
        }
    if (mx <= 0.0) return 1;
    ch = mx;
    dt = cfl * s_min(g.dx, g.dy) / mx;
    return 0;
}

template <class R> struct Workspace {
    Field<R> prim, fx, fy;
};

// One directional sweep: prim cache over all cells (solver_cpu.cpp:365-371),
// face fluxes on interior rows/columns (:198-261), conservative update with the
// non-finite revert (:375-394 / :415-432).
template <class R>
void sweep(const Field<R>& Uin, Field<R>& Uout, const Geo& g, Axis a, R dt, R ch, double gamma, int solver,
           int limiter, Workspace<R>& ws) {
    const bool xdir = (a.n == MX);
    const R dtd = dt / R(xdir ? g.dx : g.dy);
    ws.prim.resize(g.cells());
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 0; j < g.tny; ++j)
        for (int i = 0; i < g.tnx; ++i) ws.prim[g.at(i, j)] = cons_to_prim(Uin[g.at(i, j)], gamma);

    // Face f between cells c and c+1 along the axis; c runs from first-1 to last-1.
    const int nf_along = (xdir ? g.nx : g.ny) + 1;
    const int n_across = xdir ? g.ny : g.nx;
    Field<R>& fc = xdir ? ws.fx : ws.fy;
    fc.resize(static_cast<std::size_t>(nf_along) * n_across);
    const int di = xdir ? 1 : 0, dj = xdir ? 0 : 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int c = 0; c < n_across; ++c)
        for (int f = 0; f < nf_along; ++f) {
            const int i = xdir ? (g.ib - 1 + f) : (g.ib + c);
            const int j = xdir ? (g.jb + c) : (g.jb - 1 + f);
            const auto W = [&](int s) -> const V9<R>& { return ws.prim[g.at(i + s * di, j + s * dj)]; };
            V9<R> l0, r0, l1, r1;
            reconstruct(W(-1), W(0), W(1), dtd, a, ch, gamma, limiter, l0, r0);
            reconstruct(W(0), W(1), W(2), dtd, a, ch, gamma, limiter, l1, r1);
            fc[static_cast<std::size_t>(c) * nf_along + f] = riemann(r0, l1, a, solver, ch, gamma);
        }
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = g.jb; j < g.je; ++j)
        for (int i = g.ib; i < g.ie; ++i) {
            const int c = xdir ? (j - g.jb) : (i - g.ib);
            const int f = xdir ? (i - g.ib) : (j - g.jb);  // face below/left of the cell is f, above/right f+1
            const V9<R>& Fm = fc[static_cast<std::size_t>(c) * nf_along + f];
            const V9<R>& Fp = fc[static_cast<std::size_t>(c) * nf_along + f + 1];
            const V9<R>& U0 = Uin[g.at(i, j)];
            V9<R> u;
            for (int k = 0; k < 9; ++k) u[k] = U0[k] - dtd * (Fp[k] - Fm[k]);
            if (!s_finite(u[RHO]) || !s_finite(u[EN])) u = U0;
            Uout[g.at(i, j)] = u;
        }
}

// solver_cpu.cpp:327-440
template <class R>
void advance(const Field<R>& Uold, Field<R>& Utmp, Field<R>& Unew, const Geo& g, R dt, R ch, double cr,
             double gamma, int solver, int limiter, const int* bc, Workspace<R>& ws) {
    sweep(Uold, Utmp, g, kAxisX, dt, ch, gamma, solver, limiter, ws);
    copy_ghost(Uold, Utmp, g);
    boundary(Utmp, g, bc);
    sweep(Utmp, Unew, g, kAxisY, dt, ch, gamma, solver, limiter, ws);
    copy_ghost(Utmp, Unew, g);
    boundary(Unew, g, bc);
    // solver_cpu.cpp:263-275: damping after the BC, interior only.
    const R factor = R(std::exp(-val(dt) * val(ch) / cr));
    if (val(factor) >= 1.0) return;
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = g.jb; j < g.je; ++j)
        for (int i = g.ib; i < g.ie; ++i) Unew[g.at(i, j)][PSI] = Unew[g.at(i, j)][PSI] * factor;
}

Field<double> to_field(const double* U, const Geo& g) {
    Field<double> f(g.cells());
    std::memcpy(static_cast<void*>(f.data()), U, g.cells() * 9 * sizeof(double));
    return f;
}
void from_field(const Field<double>& f, double* U) {
    std::memcpy(U, static_cast<const void*>(f.data()), f.size() * 9 * sizeof(double));
}

// ------------------------------------------------------------------------------------
// Initial conditions (src/test_cases.cpp:24-148) and the cases the reference lacks.
// ------------------------------------------------------------------------------------
// test_cases.cpp:24-35
V9<double> make_mhd(double rho, double ux, double uy, double uz, double Bx, double By, double Bz, double p,
                    double gamma) {
    const double b2 = Bx * Bx + By * By + Bz * Bz;
    const double ke = 0.5 * rho * (ux * ux + uy * uy + uz * uz);
    const double E = p / (gamma - 1.0) + ke + 0.5 * b2;
    V9<double> U;
    U[RHO] = rho; U[MX] = rho * ux; U[MY] = rho * uy; U[MZ] = rho * uz;
    U[BXI] = Bx; U[BYI] = By; U[BZI] = Bz; U[EN] = E; U[PSI] = 0.0;
    return U;
}

uint64_t splitmix(uint64_t& s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double cell_uniform(uint64_t seed, int i, int j) {
    uint64_t s = seed ^ ((static_cast<uint64_t>(static_cast<uint32_t>(j)) << 32) |
                         static_cast<uint64_t>(static_cast<uint32_t>(i)));
    return static_cast<double>(splitmix(s) >> 11) * 0x1.0p-53;
}

V9<double> ic_kh(double x, double y) {  // test_cases.cpp:43-63
    const double gamma = 1.4;
    const double ux = 5.0 * (std::tanh(20.0 * (y + 0.5)) - (std::tanh(20.0 * (y - 0.5)) + 1.0));
    const double uy = 0.25 * std::sin(2.0 * kPi * x) *
                      (std::exp(-100.0 * (y + 0.5) * (y + 0.5)) - std::exp(-100.0 * (y - 0.5) * (y - 0.5)));
    return make_mhd(1.0, ux, uy, 0.0, 1.0, 0.0, 0.0, 50.0, gamma);
}

V9<double> ic_shock_bubble(double x, double y) {  // test_cases.cpp:76-106
    const double gamma = 1.4, rho_air = 1.29, rho_he = 0.214, p0 = 1.01325e5, Ms = 1.22;
    const double xc = 0.035, yc = 0.0445, Rb = 0.025, xs = 0.005;
    const double c1 = std::sqrt(gamma * p0 / rho_air);
    const double vs = Ms * c1;
    const double rho2 = rho_air * (gamma + 1.0) * Ms * Ms / ((gamma - 1.0) * Ms * Ms + 2.0);
    const double p2 = p0 * (2.0 * gamma * Ms * Ms - (gamma - 1.0)) / (gamma + 1.0);
    const double u2 = vs * (1.0 - rho_air / rho2);
    if (x < xs) return make_mhd(rho2, u2, 0.0, 0.0, 0.0, 0.0, 0.0, p2, gamma);
    const double r2 = (x - xc) * (x - xc) + (y - yc) * (y - yc);
    if (r2 < Rb * Rb) return make_mhd(rho_he, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, p0, gamma);
    return make_mhd(rho_air, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, p0, gamma);
}

V9<double> ic_brio_wu(double x) {  // test_cases.cpp:117-125
    if (x < 0.5) return make_mhd(1.0, 0.0, 0.0, 0.0, 0.75, 1.0, 0.0, 1.0, 2.0);
    return make_mhd(0.125, 0.0, 0.0, 0.0, 0.75, -1.0, 0.0, 0.1, 2.0);
}

V9<double> ic_orszag_tang(double x, double y) {  // test_cases.cpp:139-148
    const double gamma = 5.0 / 3.0;
    return make_mhd(gamma * gamma, -std::sin(y), std::sin(x), 0.0, -std::sin(y), std::sin(2.0 * x), 0.0, gamma,
                    gamma);
}

// Orszag-Tang on [0,1]^2 (rho = 25/(36 pi), p = 5/(12 pi), B = (-sin 2 pi y,
// sin 4 pi x) / sqrt(4 pi)), with an optional Gaussian density bump (p held, E
// recomputed).
V9<double> ic_ot_unit(double x, double y, const orc_ic_params* p, bool bump) {
    double rho = 25.0 / (36.0 * kPi);
    const double pr = 5.0 / (12.0 * kPi);
    const double b0 = 1.0 / std::sqrt(4.0 * kPi);
    if (bump) {
        const double r2 = (x - p->bump_cx) * (x - p->bump_cx) + (y - p->bump_cy) * (y - p->bump_cy);
        rho = rho + p->bump_amp * std::exp(-r2 / (2.0 * p->bump_sigma * p->bump_sigma));
    }
    const double sy = std::sin(2.0 * kPi * y);
    const double sx = std::sin(2.0 * kPi * x);
    return make_mhd(rho, -sy, sx, 0.0, -sy * b0, std::sin(4.0 * kPi * x) * b0, 0.0, pr, 5.0 / 3.0);
}

// MHD shock-cloud on [0,2]x[0,1]: post-shock slab x < 0.6, dense bubble r < 0.15
// at (0.8, 0.5) in pressure balance with seeded +-1 % density noise.
V9<double> ic_shock_cloud(double x, double y, int i, int j, uint64_t seed) {
    const double gamma = 5.0 / 3.0;
    if (x < 0.6) return make_mhd(3.86859, 11.2536, 0.0, 0.0, 0.0, 2.1826, -2.1826, 167.345, gamma);
    const double r2 = (x - 0.8) * (x - 0.8) + (y - 0.5) * (y - 0.5);
    if (r2 < 0.15 * 0.15) {
        const double rho = 10.0 * (1.0 + 0.01 * (2.0 * cell_uniform(seed, i, j) - 1.0));
        return make_mhd(rho, 0.0, 0.0, 0.0, 0.0, 0.56418, 0.56418, 1.0, gamma);
    }
    return make_mhd(1.0, 0.0, 0.0, 0.0, 0.0, 0.56418, 0.56418, 1.0, gamma);
}

}  // namespace

// ====================================================================================
// C ABI
// ====================================================================================
extern "C" {

void orc_cons_to_prim(const double* U, double gamma, double* V) {
    store9(cons_to_prim(load9<double>(U), gamma), V);
}
void orc_prim_to_cons(const double* V, double gamma, double* U) {
    store9(prim_to_cons(load9<double>(V), gamma), U);
}
void orc_flux(const double* U, int dir, double ch, double gamma, double* F) {
    store9(flux(load9<double>(U), axis_of(dir), ch, gamma), F);
}
double orc_fast_speed(const double* V, int dir, double gamma) {
    return fast_speed(load9<double>(V), axis_of(dir), gamma);
}
double orc_max_signal_speed(const double* V, int dir, double ch, double gamma) {
    return max_signal_speed(load9<double>(V), axis_of(dir), ch, gamma);
}
void orc_solver_flux(int kind, const double* UL, const double* UR, int dir, double ch, double gamma, double* F) {
    const V9<double> a = load9<double>(UL), b = load9<double>(UR);
    const Axis ax = axis_of(dir);
    V9<double> f = kind == 0 ? hll_flux(a, b, ax, ch, gamma)
                 : kind == 1 ? hllc_flux(a, b, ax, ch, gamma)
                 : kind == 2 ? hlld_flux(a, b, ax, ch, gamma)
                             : force_flux(a, b, ax, ch, gamma);
    store9(f, F);
}
void orc_riemann_flux(const double* UL, const double* UR, int dir, int solver, double ch, double gamma,
                      double* F) {
    store9(riemann(load9<double>(UL), load9<double>(UR), axis_of(dir), solver, ch, gamma), F);
}
void orc_apply_glm_flux(double* F, const double* UL, const double* UR, int dir, double ch) {
    V9<double> f = load9<double>(F);
    apply_glm(f, load9<double>(UL), load9<double>(UR), axis_of(dir), ch);
    store9(f, F);
}
double orc_limiter(double a, double b, int limiter) { return limit(a, b, limiter); }
int orc_is_physical(const double* V) { return is_physical(load9<double>(V)) ? 1 : 0; }
void orc_reconstruct(const double* Wm, const double* Wc, const double* Wp, double dt_over_d, int dir, double ch,
                     double gamma, int limiter, double* UL_star, double* UR_star) {
    V9<double> l, r;
    reconstruct(load9<double>(Wm), load9<double>(Wc), load9<double>(Wp), dt_over_d, axis_of(dir), ch, gamma,
                limiter, l, r);
    store9(l, UL_star);
    store9(r, UR_star);
}
uint64_t orc_splitmix64(uint64_t* state) { return splitmix(*state); }
double orc_cell_uniform(uint64_t seed, int i, int j) { return cell_uniform(seed, i, j); }
double orc_ic_sin(double a) { return mhd_ic::ic_sin(a); }

int orc_compute_dt(const double* U, const orc_geom* gm, double cfl, double gamma, double* ch_inout, double* dt) {
    const Geo g(*gm);
    return compute_dt(to_field(U, g), g, cfl, gamma, *ch_inout, *dt);
}

int orc_advance(const double* Uold, double* Utmp, double* Unew, const orc_geom* gm, double dt, double ch,
                double cr, double gamma, int solver, int limiter, const int* bc) {
    const Geo g(*gm);
    const Field<double> o = to_field(Uold, g);
    Field<double> t = to_field(Utmp, g), n = to_field(Unew, g);
    Workspace<double> ws;
    advance(o, t, n, g, dt, ch, cr, gamma, solver, limiter, bc, ws);
    from_field(t, Utmp);
    from_field(n, Unew);
    return 0;
}

void orc_apply_boundary(double* U, const orc_geom* gm, const int* bc) {
    const Geo g(*gm);
    Field<double> f = to_field(U, g);
    boundary(f, g, bc);
    from_field(f, U);
}

void orc_copy_ghost_cells(const double* src, double* dst, const orc_geom* gm) {
    const Geo g(*gm);
    const Field<double> s = to_field(src, g);
    Field<double> d = to_field(dst, g);
    copy_ghost(s, d, g);
    from_field(d, dst);
}

// scripts/cpu/main_cpu.cpp:173-222: Unew and Utmp start as copies of Uold.
int orc_run_loop(double* U, const orc_geom* gm, double cfl, double t_end, int max_steps, double gamma, double cr,
                 int solver, int limiter, const int* bc, double* ch_inout, double* dts, double* chs, int* nsteps,
                 double* t_out) {
    double s = 0.0;
    return orc_run_loop_timed(U, gm, cfl, t_end, max_steps, gamma, cr, solver, limiter, bc, ch_inout, dts, chs,
                              nsteps, t_out, &s);
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int orc_run_loop_timed(double* U, const orc_geom* gm, double cfl, double t_end, int max_steps, double gamma,
                       double cr, int solver, int limiter, const int* bc, double* ch_inout, double* dts,
                       double* chs, int* nsteps, double* t_out, double* seconds) {
    const Geo g(*gm);
    Field<double> a = to_field(U, g), b = a, tmp = a;
    Workspace<double> ws;
    const auto t0 = std::chrono::steady_clock::now();
    double t = 0.0, ch = *ch_inout;
    int step = 0, code = 0;
    while (t < t_end && step < max_steps) {
        double dt_raw = 0.0;
        if (compute_dt(a, g, cfl, gamma, ch, dt_raw) != 0) { code = 1; break; }
        const double dt = s_min(dt_raw, t_end - t);
        if (!std::isfinite(dt) || dt <= 0.0) { code = 2; break; }
        if (dt < 1.0e-8 * t_end) { code = 3; break; }
        advance(a, tmp, b, g, dt, ch, cr, gamma, solver, limiter, bc, ws);
        a.swap(b);
        t += dt;
        if (dts) dts[step] = dt;
        if (chs) chs[step] = ch;
        ++step;
    }
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    from_field(a, U);
    *ch_inout = ch;
    *nsteps = step;
    *t_out = t;
    return code;
}

int orc_case_config(int id, int n, double* c, int* bc) {
    if (n < 1) return -1;
    const double tw = 2.0 * kPi;
    double cfg[10];
    int b[4];
    switch (id) {
        case 0: { const double v[10] = {200, 400, 2, 0.0, 1.0, -1.0, 1.0, 0.3, 0.5, 1.4};
                  std::memcpy(cfg, v, sizeof v); b[0] = b[1] = b[2] = b[3] = 0; break; }
        case 1: {
            const double vs = 1.22 * std::sqrt(1.4 * 1.01325e5 / 1.29);
            const double t0 = (0.035 - 0.025 - 0.005) / vs, tau = 0.025 / vs;
            const double v[10] = {500, 197, 2, 0.0, 0.225, 0.0, 0.089, 0.4, t0 + 19.0 * tau, 1.4};
            std::memcpy(cfg, v, sizeof v); b[0] = b[1] = b[2] = b[3] = 1; break; }
        case 2: { const double v[10] = {800, 4, 2, 0.0, 1.0, 0.0, 4.0 / 800.0, 0.4, 0.1, 2.0};
                  std::memcpy(cfg, v, sizeof v); b[0] = b[1] = 1; b[2] = b[3] = 0; break; }
        case 3: { const double v[10] = {192, 192, 2, 0.0, tw, 0.0, tw, 0.4, tw, 5.0 / 3.0};
                  std::memcpy(cfg, v, sizeof v); b[0] = b[1] = b[2] = b[3] = 0; break; }
        case 4: { const double v[10] = {1024, 1024, 2, 0.0, 1.0, 0.0, 1.0, 0.8, 0.5, 5.0 / 3.0};
                  std::memcpy(cfg, v, sizeof v); b[0] = b[1] = b[2] = b[3] = 0; break; }
        case 5: { const double v[10] = {2048, 1024, 2, 0.0, 2.0, 0.0, 1.0, 0.8, 1.0, 5.0 / 3.0};
                  std::memcpy(cfg, v, sizeof v); b[0] = b[1] = b[2] = b[3] = 1; break; }
        case 6: { const double v[10] = {800, 4, 2, 0.0, 1.0, 0.0, 4.0 / 800.0, 0.8, 0.1, 2.0};
                  std::memcpy(cfg, v, sizeof v); b[0] = b[1] = b[2] = b[3] = 1; break; }
        default: return -1;
    }
    cfg[0] *= n;
    cfg[1] *= n;
    std::memcpy(c, cfg, sizeof cfg);
    std::memcpy(bc, b, sizeof b);
    return 0;
}

int orc_init_state(int id, const orc_geom* gm, const orc_ic_params* p, double* U) {
    const Geo g(*gm);
    for (int j = 0; j < g.tny; ++j)
        for (int i = 0; i < g.tnx; ++i) {
            // init.cpp:9-15 cell centres (ghosts included)
            const double x = gm->x_min + (static_cast<double>(i - g.ng) + 0.5) * g.dx;
            const double y = gm->y_min + (static_cast<double>(j - g.ng) + 0.5) * g.dy;
            V9<double> s;
            switch (id) {
                case 0: s = ic_kh(x, y); break;
                case 1: s = ic_shock_bubble(x, y); break;
                case 2: case 6: s = ic_brio_wu(x); break;
                case 3: s = ic_orszag_tang(x, y); break;
                case 4: s = ic_ot_unit(x, y, nullptr, false); break;
                case 5: s = ic_shock_cloud(x, y, i, j, p ? p->seed : 0); break;
                case 7: {  // the batch family is this build's own case: one shared definition
                    const mhd_ic::State f = mhd_ic::ot_family(x, y, *p);
                    for (int k = 0; k < 9; ++k) s[k] = f.q[k];
                    break;
                }
                case 8: s = ic_ot_unit(x, y, p, true); break;
                default: return -1;
            }
            store9(s, U + 9 * g.at(i, j));
        }
    return 0;
}

static V9<Dual> dual9(const double* v, const double* d) {
    V9<Dual> x;
    for (int k = 0; k < 9; ++k) x[k] = Dual(v[k], d[k]);
    return x;
}
static void split9(const V9<Dual>& x, double* v, double* d) {
    for (int k = 0; k < 9; ++k) {
        v[k] = x[k].v;
        d[k] = x[k].d;
    }
}

void orc_tan_cons_to_prim(const double* U, const double* dU, double gamma, double* V, double* dV) {
    split9(cons_to_prim(dual9(U, dU), gamma), V, dV);
}
void orc_tan_prim_to_cons(const double* V, const double* dV, double gamma, double* U, double* dU) {
    split9(prim_to_cons(dual9(V, dV), gamma), U, dU);
}
void orc_tan_flux(const double* U, const double* dU, int dir, double ch, double gamma, double* F, double* dF) {
    split9(flux(dual9(U, dU), axis_of(dir), Dual(ch), gamma), F, dF);
}
void orc_tan_fast_speed(const double* V, const double* dV, int dir, double gamma, double* c, double* dc) {
    const Dual r = fast_speed(dual9(V, dV), axis_of(dir), gamma);
    *c = r.v;
    *dc = r.d;
}
void orc_tan_limiter(double a, double da, double b, double db, int limiter, double* v, double* dv) {
    const Dual r = limit(Dual(a, da), Dual(b, db), limiter);
    *v = r.v;
    *dv = r.d;
}
void orc_tan_riemann(const double* UL, const double* dUL, const double* UR, const double* dUR, int dir, int solver,
                     double ch, double gamma, double* F, double* dF) {
    split9(riemann(dual9(UL, dUL), dual9(UR, dUR), axis_of(dir), solver, Dual(ch), gamma), F, dF);
}
void orc_tan_reconstruct(const double* Wm, const double* dWm, const double* Wc, const double* dWc, const double* Wp,
                         const double* dWp, double dt_over_d, int dir, double ch, double gamma, int limiter,
                         double* L, double* dL, double* R, double* dR) {
    V9<Dual> l, r;
    reconstruct(dual9(Wm, dWm), dual9(Wc, dWc), dual9(Wp, dWp), Dual(dt_over_d), axis_of(dir), Dual(ch), gamma,
                limiter, l, r);
    split9(l, L, dL);
    split9(r, R, dR);
}

// Frozen-branch central differences: for each probe q = (row, column, field), the
// copies U + eps e_q and U - eps e_q advance through nsteps steps with the given
// dt / c_h, every branch taken as on the unperturbed trajectory.  Writes the final
// density of the unperturbed run (rho0, tny x tnx) and of the 2 nprobe copies
// (rho_pm, [q][+/-][tny][tnx]).  nprobe <= 32.
int orc_frozen_fd_run(const double* U, const int* probes, int nprobe, double eps, const orc_geom* gm, int nsteps,
                      const double* dts, const double* chs, double cr, double gamma, int solver, int limiter,
                      const int* bc, double* rho0, double* rho_pm) {
    constexpr int N = 64;
    if (nprobe < 0 || 2 * nprobe > N) return 1;
    const Geo g(*gm);
    using S = Shadow<N>;
    Field<S> a(g.cells());
    for (std::size_t c = 0; c < g.cells(); ++c)
        for (int k = 0; k < 9; ++k) a[c][k] = S(U[9 * c + k]);
    for (int q = 0; q < nprobe; ++q) {
        const std::size_t c = g.at(probes[3 * q + 1], probes[3 * q]);
        const int f = probes[3 * q + 2];
        a[c][f].p[2 * q] = U[9 * c + f] + eps;
        a[c][f].p[2 * q + 1] = U[9 * c + f] - eps;
    }
    Field<S> b = a, tmp = a;
    Workspace<S> ws;
    for (int s = 0; s < nsteps; ++s) {
        advance(a, tmp, b, g, S(dts[s]), S(chs[s]), cr, gamma, solver, limiter, bc, ws);
        a.swap(b);
    }
    for (std::size_t c = 0; c < g.cells(); ++c) {
        rho0[c] = a[c][RHO].v;
        for (int k = 0; k < 2 * nprobe; ++k) rho_pm[k * g.cells() + c] = a[c][RHO].p[k];
    }
    return 0;
}

int orc_tangent_run(double* U, double* dU, const orc_geom* gm, int nsteps, const double* dts, const double* chs,
                    double cr, double gamma, int solver, int limiter, const int* bc) {
    const Geo g(*gm);
    Field<Dual> a(g.cells());
    for (std::size_t c = 0; c < g.cells(); ++c)
        for (int k = 0; k < 9; ++k) a[c][k] = Dual(U[9 * c + k], dU[9 * c + k]);
    Field<Dual> b = a, tmp = a;
    Workspace<Dual> ws;
    for (int s = 0; s < nsteps; ++s) {
        advance(a, tmp, b, g, Dual(dts[s]), Dual(chs[s]), cr, gamma, solver, limiter, bc, ws);
        a.swap(b);
    }
    for (std::size_t c = 0; c < g.cells(); ++c)
        for (int k = 0; k < 9; ++k) {
            U[9 * c + k] = a[c][k].v;
            dU[9 * c + k] = a[c][k].d;
        }
    return 0;
}

}  // extern "C"
