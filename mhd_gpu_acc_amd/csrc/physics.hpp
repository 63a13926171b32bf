// Point physics of the 2-D Dedner GLM-MHD system for the gfx950 kernels.
//
// States are held in a rotated frame: for a sweep along x the "normal"
// components are (rho u, Bx) and the tangential ones (rho v, By); along y
// they are exchanged.  Every expression of the reference (head/physics.hpp,
// head/riemann.hpp, src/cpu/solver_cpu.cpp) is invariant under that exchange
// up to commutation of a single + or *, which IEEE arithmetic preserves
// exactly, so one code path serves both sweeps bit-for-bit.
//
// Evaluation order follows the reference literally (C++ left-to-right
// association; `v/s` on a 9-vector is `(1/s)*v`, types.hpp:104-106;
// std::min/std::max NaN rules).  Built with -ffp-contract=off: no fused
// multiply-adds are formed, so every product and sum rounds as on the CPU.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

// Measurement knob (never the default): built with -DMHD_FP_CONTRACT_FAST=1 the forward
// kernels let the compiler fuse multiplies and adds, to time what bitwise parity with
// the reference's non-contracting flags (Makefile:8,18) costs.  Results then differ
// from the reference's in the last bits.
#if defined(MHD_FP_CONTRACT_FAST) && MHD_FP_CONTRACT_FAST
#define MHD_FP_CONTRACT _Pragma("clang fp contract(fast)")
#else
#define MHD_FP_CONTRACT _Pragma("clang fp contract(off)")
#endif
MHD_FP_CONTRACT

namespace mhd {

// Host and device: the host build backs the self-check entry points (hostcheck.cpp).
#define MHD_DI __host__ __device__ __forceinline__

// Conserved state, rotated: (rho, m_n, m_t, m_z, B_n, B_t, B_z, E, psi).
struct Cons {
    double r, mn, mt, mz, bn, bt, bz, e, psi;
};
// Primitive state, rotated: (rho, u_n, u_t, u_z, B_n, B_t, B_z, p, psi), plus
// |B|^2 as cons_to_prim forms it (physics.hpp:256): flux_x/flux_y and
// fast_speed evaluate the same expression on the same components, so reusing it
// is bitwise neutral.  Only to_prim fills b2.
struct Prim {
    double r, un, ut, uz, bn, bt, bz, p, psi;
    double b2;
};

// Per-launch physics constants.
struct Phys {
    double gamma;   // adiabatic index
    double gm1;     // gamma - 1 (physics.hpp:258, 268)
    double ch;      // GLM cleaning speed of this step
    double ch2;     // ch * ch (physics.hpp:307)
    double half_ch; // 0.5 * ch (riemann.hpp:204, 441, 558)
    double half_over_ch;  // 0.5 / ch (riemann.hpp:557)
};

MHD_DI bool finite(double x) { return __builtin_isfinite(x); }MHD_DI double std_min(double a, double b) { return (b < a) ? b : a; }  // std::min
MHD_DI double std_max(double a, double b) { return (a < b) ? b : a; }  // std::max

// ---------------------------------------------------------------------------------
// Quotients by a shared denominator
// ---------------------------------------------------------------------------------
// The device lowers a / b (FP64) to: div_scale of b and of a, rcp of the scaled b, two
// Newton-Raphson corrections of that reciprocal, the quotient, its residual (FMA),
// div_fmas and div_fixup.  div_scale leaves both operands unchanged, and div_fmas is
// then a plain FMA, whenever b is normal with a normal reciprocal, a is zero or at
// least 2^-969 in magnitude, the quotient is normal and a's biased exponent exceeds
// b's by less than 768.  div_fixup returns that quotient for finite non-zero operands
// and the IEEE result for a zero, infinite or NaN operand whatever quotient it is given.
// So for operands whose frexp exponents lie in [-DIV_EXP, DIV_EXP] (|x| in
// [2^-381, 2^380): every condition above holds with room), or that are zero, infinite
// or NaN (frexp exponent 0), the sequence without the two div_scale steps returns the
// bits of a / b.  The refined reciprocal depends on b alone, so divisions by one
// denominator share it: each further quotient is a multiply, two FMAs and the fixup,
// instead of the whole sequence around a quarter-rate rcp and its VCC-carried scale.
//
// FastDiv forms quotients that way and keeps the smallest and largest frexp exponent
// of every operand it divided; ExactDiv divides.  A caller evaluates with FastDiv and
// keeps the result when no lane of the wave saw an exponent outside the range (one
// ballot), else evaluates again with ExactDiv: either way every quotient is a / b.
constexpr int DIV_EXP = 380;

struct Rcp {  // a denominator and its refined reciprocal (FastDiv) or 0 (ExactDiv)
    double b, r;
};

// The reciprocal the device's division refines from rcp(b) (0 on the host).
MHD_DI double refined_rcp(double b) {
#ifdef __HIP_DEVICE_COMPILE__
    const double r0 = __builtin_amdgcn_rcp(b);
    const double e0 = __builtin_fma(-b, r0, 1.0);
    const double r1 = __builtin_fma(r0, e0, r0);
    const double e1 = __builtin_fma(-b, r1, 1.0);
    return __builtin_fma(r1, e1, r1);
#else
    (void)b;
    return 0.0;
#endif
}

struct ExactDiv {
    MHD_DI Rcp rcp(double b) { return Rcp{b, 0.0}; }
    MHD_DI double q(double a, const Rcp& d) { return a / d.b; }
};

struct FastDiv {
    int lo = 0, hi = 0;  // frexp exponents seen (0 for zero, infinite and NaN operands)
    MHD_DI void note(double x) {
#ifdef __HIP_DEVICE_COMPILE__
        const int e = __builtin_amdgcn_frexp_exp(x);
        lo = e < lo ? e : lo;
        hi = e > hi ? e : hi;
#else
        (void)x;
#endif
    }
    MHD_DI Rcp rcp(double b) {
        note(b);
        return Rcp{b, refined_rcp(b)};
    }
    MHD_DI double q(double a, const Rcp& d) {
        note(a);
#ifdef __HIP_DEVICE_COMPILE__
        const double q0 = a * d.r;
        const double e = __builtin_fma(-d.b, q0, a);
        return __builtin_amdgcn_div_fixup(__builtin_fma(e, d.r, q0), d.b, a);
#else
        return a / d.b;
#endif
    }
    // Every active lane's operands were in range (wave-uniform).
    MHD_DI bool wave_ok() const {
#ifdef __HIP_DEVICE_COMPILE__
        return __builtin_amdgcn_ballot_w64(lo < -DIV_EXP || hi > DIV_EXP) == 0;
#else
        return true;
#endif
    }
};

// FastDiv's sequence for operands the caller has already found in range.
struct CoreDiv {
    MHD_DI Rcp rcp(double b) { return Rcp{b, refined_rcp(b)}; }
    MHD_DI double q(double a, const Rcp& d) {
#ifdef __HIP_DEVICE_COMPILE__
        const double q0 = a * d.r;
        const double e = __builtin_fma(-d.b, q0, a);
        return __builtin_amdgcn_div_fixup(__builtin_fma(e, d.r, q0), d.b, a);
#else
        return a / d.b;
#endif
    }
};

// Whether every active lane's values have frexp exponents in [-DIV_EXP, DIV_EXP] (or are
// zero, infinite or NaN), so that CoreDiv may divide by and into them (wave-uniform;
// false on the host).
MHD_DI bool wave_in_range(double x0, double x1, double x2 = 0.0, double x3 = 0.0, double x4 = 0.0) {
#ifdef __HIP_DEVICE_COMPILE__
    FastDiv g;
    g.note(x0); g.note(x1); g.note(x2); g.note(x3); g.note(x4);
    return g.wave_ok();
#else
    (void)x0; (void)x1; (void)x2; (void)x3; (void)x4;
    return false;
#endif
}

// physics.hpp:69-78
template <class D> MHD_DI Prim to_prim(const Cons& u, const Phys& ph, D& dv) {
    const Rcp r = dv.rcp(u.r);
    const double inv = dv.q(1.0, r);
    Prim w;
    w.r = u.r;
    w.un = u.mn * inv;
    w.ut = u.mt * inv;
    w.uz = u.mz * inv;
    w.bn = u.bn;
    w.bt = u.bt;
    w.bz = u.bz;
    const double b2 = u.bn * u.bn + u.bt * u.bt + u.bz * u.bz;
    const double ke = 0.5 * u.r * (w.un * w.un + w.ut * w.ut + w.uz * w.uz);
    w.p = ph.gm1 * (u.e - ke - 0.5 * b2);
    w.psi = u.psi;
    w.b2 = b2;
    return w;
}
MHD_DI Prim to_prim(const Cons& u, const Phys& ph) {
    ExactDiv ex;
    return to_prim(u, ph, ex);
}

// physics.hpp:83-89.  rg: the reciprocal of gamma - 1 (rcp(ph.gm1) of the same policy).
template <class D> MHD_DI Cons to_cons(const Prim& w, const Phys& ph, D& dv, const Rcp& rg) {
    const double b2 = w.bn * w.bn + w.bt * w.bt + w.bz * w.bz;
    const double ke = 0.5 * w.r * (w.un * w.un + w.ut * w.ut + w.uz * w.uz);
    Cons u;
    u.r = w.r;
    u.mn = w.r * w.un;
    u.mt = w.r * w.ut;
    u.mz = w.r * w.uz;
    u.bn = w.bn;
    u.bt = w.bt;
    u.bz = w.bz;
    u.e = dv.q(w.p, rg) + ke + 0.5 * b2;
    u.psi = w.psi;
    return u;
}
MHD_DI Cons to_cons(const Prim& w, const Phys& ph) {
    ExactDiv ex;
    return to_cons(w, ph, ex, ex.rcp(ph.gm1));
}

// physics.hpp:105-164.  `w` must be to_prim(u): the reference recomputes the same
// expressions inside flux_x/flux_y, so passing them in is bitwise neutral.
MHD_DI Cons flux(const Cons& u, const Prim& w, const Phys& ph) {
    const double pt = w.p + 0.5 * w.b2;
    const double bu = u.bn * w.un + u.bt * w.ut + u.bz * w.uz;
    Cons f;
    f.r = u.mn;
    f.mn = u.mn * w.un + pt - u.bn * u.bn;
    f.mt = u.mn * w.ut - u.bn * u.bt;
    f.mz = u.mn * w.uz - u.bn * u.bz;
    f.bn = u.psi;
    f.bt = u.bt * w.un - u.bn * w.ut;
    f.bz = u.bz * w.un - u.bn * w.uz;
    f.e = (u.e + pt) * w.un - u.bn * bu;
    f.psi = ph.ch2 * u.bn;
    return f;
}

// Square root with the bits of sqrt().  On the device sqrt is lowered to rsq
// followed by two Newton-Raphson corrections, wrapped in a range scaling (taken for
// x below 2^-767) and a fixup that passes 0 and +inf through; for x in
// [2^-767, DBL_MAX] the wrapping is the identity.  A wave whose active lanes all
// hold such an x runs the core sequence alone, operation for operation as the
// lowering orders it (10 instructions instead of 18); any other wave calls sqrt.
MHD_DI double sqrt_cr(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    const bool core = x >= 0x1p-767 && x <= 0x1.fffffffffffffp+1023;
    if (__builtin_amdgcn_ballot_w64(!core) == 0) {
        double h = __builtin_amdgcn_rsq(x);
        double g = x * h;
        h = h * 0.5;
        const double e = __builtin_fma(-h, g, 0.5);
        g = __builtin_fma(g, e, g);
        h = __builtin_fma(h, e, h);
        double d = __builtin_fma(-g, g, x);
        g = __builtin_fma(d, h, g);
        d = __builtin_fma(-g, g, x);
        return __builtin_fma(d, h, g);
    }
#endif
    return sqrt(x);
}

// physics.hpp:175-190 (fast magnetosonic speed along the normal); the three
// divisions by rho share one reciprocal under FastDiv.
template <class D> MHD_DI double fast_speed(const Prim& w, const Phys& ph, D& dv) {
    const Rcp r = dv.rcp(w.r);
    const double a2 = dv.q(ph.gamma * w.p, r);
    const double b2 = dv.q(w.b2, r);
    const double bn2 = dv.q(w.bn * w.bn, r);
    const double disc = (a2 + b2) * (a2 + b2) - 4.0 * a2 * bn2;
    return sqrt_cr(0.5 * (a2 + b2 + sqrt_cr(disc > 0.0 ? disc : 0.0)));
}
MHD_DI double fast_speed(const Prim& w, const Phys& ph) {
    ExactDiv ex;
    return fast_speed(w, ph, ex);
}

// riemann.hpp:54-57
MHD_DI bool prim_physical(const Prim& w) {
    return finite(w.r) && finite(w.p) && w.r > 0.0 && w.p > 0.0;
}

// solver_cpu.cpp:49-56 (floors 1e-12, every component finite)
MHD_DI bool is_physical(const Prim& w) {
    return finite(w.r) && w.r > 1.0e-12 && finite(w.p) && w.p > 1.0e-12 && finite(w.un) && finite(w.ut) &&
           finite(w.uz) && finite(w.bn) && finite(w.bt) && finite(w.bz) && finite(w.psi);
}

MHD_DI Cons half_sum(const Cons& a, const Cons& b) {
    return Cons{0.5 * (a.r + b.r), 0.5 * (a.mn + b.mn), 0.5 * (a.mt + b.mt), 0.5 * (a.mz + b.mz),
                0.5 * (a.bn + b.bn), 0.5 * (a.bt + b.bt), 0.5 * (a.bz + b.bz), 0.5 * (a.e + b.e),
                0.5 * (a.psi + b.psi)};
}

// Star state from rotated components (riemann.hpp:219-230, 417-437).
MHD_DI Cons star(double rs, double uns, double uts, double uws, double bns, double bts, double bws, double es,
                 double psis) {
    return Cons{rs, rs * uns, rs * uts, rs * uws, bns, bts, bws, es, psis};
}

// f + s * (a - b), componentwise
MHD_DI Cons fan(const Cons& f, double s, const Cons& a, const Cons& b) {
    return Cons{f.r + s * (a.r - b.r),    f.mn + s * (a.mn - b.mn), f.mt + s * (a.mt - b.mt),
                f.mz + s * (a.mz - b.mz), f.bn + s * (a.bn - b.bn), f.bt + s * (a.bt - b.bt),
                f.bz + s * (a.bz - b.bz), f.e + s * (a.e - b.e),    f.psi + s * (a.psi - b.psi)};
}

// ---------------------------------------------------------------------------------
// HLL (riemann.hpp:70-107).  FL/FR and the prims are passed in.
// ---------------------------------------------------------------------------------
template <class D>
MHD_DI Cons hll(const Cons& ul, const Prim& vl, const Cons& fl, const Cons& ur, const Prim& vr, const Cons& fr,
                const Phys& ph, D& dv) {
    if (!prim_physical(vl) || !prim_physical(vr)) return half_sum(fl, fr);
    const double cfl = fast_speed(vl, ph, dv);
    const double cfr = fast_speed(vr, ph, dv);
    const double sl = fmin(fmin(vl.un - cfl, vr.un - cfr), -ph.ch);
    const double sr = fmax(fmax(vl.un + cfl, vr.un + cfr), ph.ch);
    if (sl >= 0.0) return fl;
    if (sr <= 0.0) return fr;
    const double den = sr - sl;
    if (fabs(den) < 1.0e-14) return half_sum(fl, fr);
    const double inv = 1.0 / den;
    const double ss = sl * sr;
    return Cons{inv * (sr * fl.r - sl * fr.r + ss * (ur.r - ul.r)),
                inv * (sr * fl.mn - sl * fr.mn + ss * (ur.mn - ul.mn)),
                inv * (sr * fl.mt - sl * fr.mt + ss * (ur.mt - ul.mt)),
                inv * (sr * fl.mz - sl * fr.mz + ss * (ur.mz - ul.mz)),
                inv * (sr * fl.bn - sl * fr.bn + ss * (ur.bn - ul.bn)),
                inv * (sr * fl.bt - sl * fr.bt + ss * (ur.bt - ul.bt)),
                inv * (sr * fl.bz - sl * fr.bz + ss * (ur.bz - ul.bz)),
                inv * (sr * fl.e - sl * fr.e + ss * (ur.e - ul.e)),
                inv * (sr * fl.psi - sl * fr.psi + ss * (ur.psi - ul.psi))};
}
MHD_DI Cons hll(const Cons& ul, const Prim& vl, const Cons& fl, const Cons& ur, const Prim& vr, const Cons& fr,
                const Phys& ph) {
    ExactDiv ex;
    return hll(ul, vl, fl, ur, vr, fr, ph, ex);
}

// ---------------------------------------------------------------------------------
// HLLC (riemann.hpp:122-241)
// ---------------------------------------------------------------------------------
// Returns false when the reference falls back to HLL (riemann.hpp:131-133, 188,
// 194-197); otherwise f is the flux.  Only the physical flux of the side the
// interface lies on is formed (FL for S_M >= 0 or S_L >= 0, FR otherwise), which
// is the one the reference's region selection returns or builds on.
template <class D>
MHD_DI bool hllc(const Cons& ul, const Prim& vl, const Cons& ur, const Prim& vr, const Phys& ph, Cons& f, D& dv) {
    if (!prim_physical(vl) || !prim_physical(vr)) return false;
    const double cfl = fast_speed(vl, ph, dv);
    const double cfr = fast_speed(vr, ph, dv);
    const double sl = fmin(vl.un - cfl, vr.un - cfr);
    const double sr = fmax(vl.un + cfl, vr.un + cfr);
    if (sl >= 0.0) { f = flux(ul, vl, ph); return true; }
    if (sr <= 0.0) { f = flux(ur, vr, ph); return true; }
    const double bl2 = vl.b2;  // B_n^2 + B_t^2 + B_w^2 (riemann.hpp:180-181)
    const double br2 = vr.b2;
    const double ptl = vl.p + 0.5 * bl2;
    const double ptr = vr.p + 0.5 * br2;
    const double num = (sr - vr.un) * vr.r * vr.un - (sl - vl.un) * vl.r * vl.un + ptl - ptr;
    const double den = (sr - vr.un) * vr.r - (sl - vl.un) * vl.r;
    if (fabs(den) < 1.0e-14) return false;
    const double sm = num / den;
    const double rls = vl.r * (sl - vl.un) / (sl - sm);
    const double rrs = vr.r * (sr - vr.un) / (sr - sm);
    if (rls <= 0.0 || rrs <= 0.0 || !finite(rls) || !finite(rrs)) return false;
    const double psim = 0.5 * (vl.psi + vr.psi) - ph.half_ch * (vr.bn - vl.bn);
    const bool left = sm >= 0.0;
    const Cons& u = left ? ul : ur;
    const Prim& v = left ? vl : vr;
    const double s = left ? sl : sr;
    const double pt = left ? ptl : ptr;
    const double pts = pt + v.r * (s - v.un) * (sm - v.un);
    const double bu = v.bn * v.un + v.bt * v.ut + v.bz * v.uz;
    const double bus = v.bn * sm + v.bt * v.ut + v.bz * v.uz;
    const double es = ((s - v.un) * u.e - pt * v.un + pts * sm + v.bn * (bus - bu)) / (s - sm);
    f = fan(flux(u, v, ph), s, star(left ? rls : rrs, sm, v.ut, v.uz, v.bn, v.bt, v.bz, es, psim), u);
    return true;
}

// ---------------------------------------------------------------------------------
// HLLD (riemann.hpp:253-480), Miyoshi-Kusano with B_n allowed to jump (GLM).
// ---------------------------------------------------------------------------------
struct SideStar {
    double ut, uw, bt, bw, e;  // single-star tangential velocity/field and energy
    double bu;                 // B.u of the single-star state
};

// L* / R* states (riemann.hpp:341-373).  s = S_L or S_R.  The four divisions by
// the denominator d share one reciprocal under FastDiv.
template <class D>
MHD_DI SideStar side_star(const Cons& u, const Prim& v, double s, double sm, double ptl, double ptstar, D& dv) {
    SideStar o;
    const double d = v.r * (s - v.un) * (s - sm) - v.bn * v.bn;
    if (fabs(d) < 1.0e-14) {
        o.ut = v.ut; o.uw = v.uz; o.bt = v.bt; o.bw = v.bz;
    } else {
        const Rcp rd = dv.rcp(d);
        o.ut = v.ut - dv.q(v.bn * v.bt * (sm - v.un), rd);
        o.uw = v.uz - dv.q(v.bn * v.bz * (sm - v.un), rd);
        const double k = v.r * (s - v.un) * (s - v.un) - v.bn * v.bn;
        o.bt = dv.q(v.bt * k, rd);
        o.bw = dv.q(v.bz * k, rd);
    }
    const double bu = v.bn * v.un + v.bt * v.ut + v.bz * v.uz;
    o.bu = v.bn * sm + o.bt * o.ut + o.bw * o.uw;
    o.e = (u.e * (s - v.un) - ptl * v.un + ptstar * sm + v.bn * (o.bu - bu)) / (s - sm);
    return o;
}
MHD_DI SideStar side_star(const Cons& u, const Prim& v, double s, double sm, double ptl, double ptstar) {
    ExactDiv ex;
    return side_star(u, v, s, sm, ptl, ptstar, ex);
}

// Returns false when the reference falls back to HLL (riemann.hpp:262-264,
// 301-303, 316-318, 325-328); otherwise f is the flux.  As in hllc, only the
// physical flux of the interface's side is formed.
template <class D>
MHD_DI bool hlld(const Cons& ul, const Prim& vl, const Cons& ur, const Prim& vr, const Phys& ph, Cons& f, D& dv) {
    if (!prim_physical(vl) || !prim_physical(vr)) return false;
    const double cfl = fast_speed(vl, ph, dv);
    const double cfr = fast_speed(vr, ph, dv);
    const double sl = fmin(vl.un - cfl, vr.un - cfr);
    const double sr = fmax(vl.un + cfl, vr.un + cfr);
    if (fabs(sr - sl) < 1.0e-14) return false;
    const double bl2 = vl.b2;  // B_n^2 + B_t^2 + B_w^2 (riemann.hpp:306-307)
    const double br2 = vr.b2;
    const double ptl = vl.p + 0.5 * bl2;
    const double ptr = vr.p + 0.5 * br2;
    const double num = (sr - vr.un) * vr.r * vr.un - (sl - vl.un) * vl.r * vl.un + ptl - ptr;
    const double den = (sr - vr.un) * vr.r - (sl - vl.un) * vl.r;
    if (fabs(den) < 1.0e-14) return false;
    const double sm = num / den;
    const double rls = vl.r * (sl - vl.un) / (sl - sm);
    const double rrs = vr.r * (sr - vr.un) / (sr - sm);
    if (rls <= 0.0 || rrs <= 0.0 || !finite(rls) || !finite(rrs)) return false;
    // The outermost regions are decided before any star state is needed
    // (riemann.hpp:448-449); the values they return do not depend on it.
    if (sl >= 0.0) { f = flux(ul, vl, ph); return true; }
    if (sr <= 0.0) { f = flux(ur, vr, ph); return true; }
    const double ptstar = ptl + vl.r * (sl - vl.un) * (sm - vl.un);
    const double sql = sqrt_cr(rls);
    const double sqr = sqrt_cr(rrs);
    const double sls = sm - fabs(vl.bn) / sql;
    const double srs = sm + fabs(vr.bn) / sqr;
    const double psim = 0.5 * (vl.psi + vr.psi) - ph.half_ch * (vr.bn - vl.bn);
    const SideStar L = side_star(ul, vl, sl, sm, ptl, ptstar, dv);
    const SideStar R = side_star(ur, vr, sr, sm, ptr, ptstar, dv);
    const bool left = sm >= 0.0;
    const bool outer = left ? (sls >= 0.0) : (srs <= 0.0);
    const Cons& us = left ? ul : ur;
    const Cons fs = flux(us, left ? vl : vr, ph);
    const SideStar& S = left ? L : R;
    const double s = left ? sl : sr;
    const double rs = left ? rls : rrs;
    const double bn = left ? vl.bn : vr.bn;
    const Cons u1 = star(rs, sm, S.ut, S.uw, bn, S.bt, S.bw, S.e, psim);
    if (outer) {  // L* or R* region
        f = fan(fs, s, u1, us);
        return true;
    }
    // Double-star states (riemann.hpp:376-413); sign of B_n from the left state.
    const double sgn = (vl.bn >= 0.0) ? 1.0 : -1.0;
    double ut2, uw2, bt2, bw2, e2;
    if (fabs(sql + sqr) < 1.0e-14) {
        ut2 = S.ut; uw2 = S.uw; bt2 = S.bt; bw2 = S.bw; e2 = S.e;
    } else {
        const double inv = 1.0 / (sql + sqr);
        ut2 = (sql * L.ut + sqr * R.ut + sgn * (R.bt - L.bt)) * inv;
        uw2 = (sql * L.uw + sqr * R.uw + sgn * (R.bw - L.bw)) * inv;
        bt2 = (sql * R.bt + sqr * L.bt + sgn * sql * sqr * (R.ut - L.ut)) * inv;
        bw2 = (sql * R.bw + sqr * L.bw + sgn * sql * sqr * (R.uw - L.uw)) * inv;
        const double bu2 = bn * sm + bt2 * ut2 + bw2 * uw2;
        e2 = left ? (S.e - sql * sgn * (bu2 - S.bu)) : (S.e + sqr * sgn * (bu2 - S.bu));
    }
    const double s2 = left ? sls : srs;
    const Cons u2 = star(rs, sm, ut2, uw2, bn, bt2, bw2, e2, psim);
    f = fan(fan(fs, s, u1, us), s2, u2, u1);  // L** or R** region
    return true;
}

// ---------------------------------------------------------------------------------
// FORCE (riemann.hpp:486-525)
// ---------------------------------------------------------------------------------
template <class D>
MHD_DI Cons force(const Cons& ul, const Prim& vl, const Cons& fl, const Cons& ur, const Prim& vr, const Cons& fr,
                  const Phys& ph, D& dv) {
    if (!prim_physical(vl) || !prim_physical(vr)) return half_sum(fl, fr);
    const double cfl = fast_speed(vl, ph, dv);
    const double cfr = fast_speed(vr, ph, dv);
    const double a = fmax(fmax(fabs(vl.un) + cfl, fabs(vr.un) + cfr), ph.ch);
    if (a < 1.0e-14) return half_sum(fl, fr);
    const double ha = 0.5 * a;
    const double hia = 0.5 * (1.0 / a);
    const Cons flf{0.5 * (fl.r + fr.r) - ha * (ur.r - ul.r),       0.5 * (fl.mn + fr.mn) - ha * (ur.mn - ul.mn),
                   0.5 * (fl.mt + fr.mt) - ha * (ur.mt - ul.mt),   0.5 * (fl.mz + fr.mz) - ha * (ur.mz - ul.mz),
                   0.5 * (fl.bn + fr.bn) - ha * (ur.bn - ul.bn),   0.5 * (fl.bt + fr.bt) - ha * (ur.bt - ul.bt),
                   0.5 * (fl.bz + fr.bz) - ha * (ur.bz - ul.bz),   0.5 * (fl.e + fr.e) - ha * (ur.e - ul.e),
                   0.5 * (fl.psi + fr.psi) - ha * (ur.psi - ul.psi)};
    const Cons uri{0.5 * (ul.r + ur.r) - hia * (fr.r - fl.r),       0.5 * (ul.mn + ur.mn) - hia * (fr.mn - fl.mn),
                   0.5 * (ul.mt + ur.mt) - hia * (fr.mt - fl.mt),   0.5 * (ul.mz + ur.mz) - hia * (fr.mz - fl.mz),
                   0.5 * (ul.bn + ur.bn) - hia * (fr.bn - fl.bn),   0.5 * (ul.bt + ur.bt) - hia * (fr.bt - fl.bt),
                   0.5 * (ul.bz + ur.bz) - hia * (fr.bz - fl.bz),   0.5 * (ul.e + ur.e) - hia * (fr.e - fl.e),
                   0.5 * (ul.psi + ur.psi) - hia * (fr.psi - fl.psi)};
    const Prim vri = to_prim(uri, ph);
    if (!prim_physical(vri)) return flf;
    return half_sum(flf, flux(uri, vri, ph));
}
MHD_DI Cons force(const Cons& ul, const Prim& vl, const Cons& fl, const Cons& ur, const Prim& vr, const Cons& fr,
                  const Phys& ph) {
    ExactDiv ex;
    return force(ul, vl, fl, ur, vr, fr, ph, ex);
}

// riemann.hpp:544-563: exact upwind flux of the (B_n, psi) subsystem.
MHD_DI void glm_overwrite(Cons& f, const Cons& ul, const Cons& ur, const Phys& ph) {
    if (ph.ch <= 0.0) return;
    const double bstar = 0.5 * (ul.bn + ur.bn) - ph.half_over_ch * (ur.psi - ul.psi);
    const double pstar = 0.5 * (ul.psi + ur.psi) - ph.half_ch * (ur.bn - ul.bn);
    f.bn = pstar;
    f.psi = ph.ch2 * bstar;
}

// riemann.hpp:568-592.  XDIR selects which rotated slot is the reference's
// F.Bx in the finiteness test (:583-586): B_n on x sweeps, B_t on y sweeps.
template <int SOLVER, bool XDIR, class D>
MHD_DI Cons riemann_with(const Cons& ul, const Prim& vl, const Cons& ur, const Prim& vr, const Phys& ph, D& dv) {
    if (SOLVER == 0 || SOLVER == 3) {
        const Cons fl = flux(ul, vl, ph);
        const Cons fr = flux(ur, vr, ph);
        if (SOLVER == 0) return hll(ul, vl, fl, ur, vr, fr, ph, dv);
        Cons f = force(ul, vl, fl, ur, vr, fr, ph, dv);
        const double fbx = XDIR ? f.bn : f.bt;
        if (!finite(f.r) || !finite(f.e) || !finite(fbx) || !finite(f.psi)) return hll(ul, vl, fl, ur, vr, fr, ph, dv);
        glm_overwrite(f, ul, ur, ph);
        return f;
    }
    Cons f;
    const bool ok = (SOLVER == 1) ? hllc(ul, vl, ur, vr, ph, f, dv) : hlld(ul, vl, ur, vr, ph, f, dv);
    // The solvers' own HLL fallback is an ordinary result of hllc_flux/hlld_flux:
    // it passes the finiteness test and gets the GLM overwrite like any other.
    if (!ok) f = hll(ul, vl, flux(ul, vl, ph), ur, vr, flux(ur, vr, ph), ph, dv);
    const double fbx = XDIR ? f.bn : f.bt;
    if (!finite(f.r) || !finite(f.e) || !finite(fbx) || !finite(f.psi))
        return ok ? hll(ul, vl, flux(ul, vl, ph), ur, vr, flux(ur, vr, ph), ph, dv) : f;
    glm_overwrite(f, ul, ur, ph);
    return f;
}

// On the device the face is evaluated with shared-reciprocal quotients first and again
// with plain divisions only when some lane of the wave divided an operand outside
// FastDiv's range; the flux is riemann_with<ExactDiv>'s bit for bit either way.
template <int SOLVER, bool XDIR>
MHD_DI Cons riemann(const Cons& ul, const Prim& vl, const Cons& ur, const Prim& vr, const Phys& ph) {
#ifdef __HIP_DEVICE_COMPILE__
    FastDiv fd;
    const Cons f = riemann_with<SOLVER, XDIR>(ul, vl, ur, vr, ph, fd);
    if (fd.wave_ok()) return f;
#endif
    ExactDiv ex;
    return riemann_with<SOLVER, XDIR>(ul, vl, ur, vr, ph, ex);
}

// ---------------------------------------------------------------------------------
// MUSCL-Hancock reconstruction (solver_cpu.cpp:76-119)
// ---------------------------------------------------------------------------------
// minmod (solver_cpu.cpp:26-29); van Leer 2ab/(a+b) on same-sign slopes.
template <int LIMITER> MHD_DI double limit(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    if (LIMITER == 0) return (a > 0.0) ? std_min(a, b) : std_max(a, b);
    return 2.0 * a * b / (a + b);
}

// Interface states of one cell with their primitives (the Riemann solver needs
// both, and the reference's guard already computes the primitive of the
// candidate state, solver_cpu.cpp:65-70).
struct Face {
    Cons u;
    Prim w;
};

// What the adjoint of reconstruct() needs of its forward values (adjoint.hpp):
// the limited slopes, the conserved interface states before the half step and the
// outcome of the four physical-state guards.  reconstruct() fills it when given one.
struct RecordNone {};
struct ReconRecord {
    double s[9];           // slopes, field order of Prim
    Cons ul, ur;           // to_cons of the guarded interface primitives
    bool pl, pr, pcl, pcr; // guards: W_L, W_R, then the half-stepped states
};

// dv.fast: whether to try CoreDiv where the operands allow it (never on the host).
struct ReconDiv {
    bool fast;
};
template <int LIMITER, class Record>
MHD_DI void reconstruct_with(const Prim& wm, const Prim& wc, const Prim& wp, double hdtd, const Phys& ph, Face& lo,
                             Face& hi, Record* rec, const ReconDiv& dv) {
    constexpr bool REC = !std::is_same<Record, RecordNone>::value;
    Prim wl, wr;
#define MHD_SLOPE(k, f)                                       \
    {                                                         \
        const double s = limit<LIMITER>(wc.f - wm.f, wp.f - wc.f); \
        if constexpr (REC) rec->s[k] = s;                     \
        wl.f = wc.f - 0.5 * s;                                \
        wr.f = wc.f + 0.5 * s;                                \
    }
    MHD_SLOPE(0, r) MHD_SLOPE(1, un) MHD_SLOPE(2, ut) MHD_SLOPE(3, uz) MHD_SLOPE(4, bn) MHD_SLOPE(5, bt)
    MHD_SLOPE(6, bz) MHD_SLOPE(7, p) MHD_SLOPE(8, psi)
#undef MHD_SLOPE
    if constexpr (REC) {
        rec->pl = is_physical(wl);
        rec->pr = is_physical(wr);
        if (!rec->pl) wl = wc;
        if (!rec->pr) wr = wc;
    } else {
        if (!is_physical(wl)) wl = wc;
        if (!is_physical(wr)) wr = wc;
    }
    Cons ul, ur;
    Prim vl, vr;
    // to_cons divides p by gamma - 1 and to_prim divides 1 by rho (to_cons keeps rho)
    if (dv.fast && wave_in_range(ph.gm1, wl.p, wr.p, wl.r, wr.r)) {
        CoreDiv cd;
        const Rcp rg = cd.rcp(ph.gm1);
        ul = to_cons(wl, ph, cd, rg);
        ur = to_cons(wr, ph, cd, rg);
        vl = to_prim(ul, ph, cd);
        vr = to_prim(ur, ph, cd);
    } else {
        ExactDiv ex;
        const Rcp rg = ex.rcp(ph.gm1);
        ul = to_cons(wl, ph, ex, rg);
        ur = to_cons(wr, ph, ex, rg);
        vl = to_prim(ul, ph, ex);
        vr = to_prim(ur, ph, ex);
    }
    const Cons fl = flux(ul, vl, ph);
    const Cons fr = flux(ur, vr, ph);
    Cons cl, cr;
#define MHD_HALF(f)                          \
    {                                        \
        const double h = hdtd * (fr.f - fl.f); \
        cl.f = ul.f - h;                     \
        cr.f = ur.f - h;                     \
    }
    MHD_HALF(r) MHD_HALF(mn) MHD_HALF(mt) MHD_HALF(mz) MHD_HALF(bn) MHD_HALF(bt) MHD_HALF(bz) MHD_HALF(e)
    MHD_HALF(psi)
#undef MHD_HALF
    Prim pcl, pcr;
    if (dv.fast && wave_in_range(cl.r, cr.r)) {
        CoreDiv cd;
        pcl = to_prim(cl, ph, cd);
        pcr = to_prim(cr, ph, cd);
    } else {
        ExactDiv ex;
        pcl = to_prim(cl, ph, ex);
        pcr = to_prim(cr, ph, ex);
    }
    if constexpr (REC) {
        rec->pcl = is_physical(pcl);
        rec->pcr = is_physical(pcr);
        rec->ul = ul;
        rec->ur = ur;
        if (rec->pcl) { lo.u = cl; lo.w = pcl; } else { lo.u = ul; lo.w = vl; }
        if (rec->pcr) { hi.u = cr; hi.w = pcr; } else { hi.u = ur; hi.w = vr; }
    } else {
        if (is_physical(pcl)) { lo.u = cl; lo.w = pcl; } else { lo.u = ul; lo.w = vl; }
        if (is_physical(pcr)) { hi.u = cr; hi.w = pcr; } else { hi.u = ur; hi.w = vr; }
    }
}

// The reconstruction's divisions by gamma - 1 and by rho: on the device, each group
// whose operands every lane of the wave holds in FastDiv's range (one ballot, before
// dividing) uses CoreDiv's sequence, any other group divides (the limiter keeps its
// division); lo and hi are bit for bit the same either way.
template <int LIMITER, class Record = RecordNone>
MHD_DI void reconstruct(const Prim& wm, const Prim& wc, const Prim& wp, double hdtd, const Phys& ph, Face& lo,
                        Face& hi, Record* rec = nullptr) {
#ifdef __HIP_DEVICE_COMPILE__
    const ReconDiv dv{true};
#else
    const ReconDiv dv{false};
#endif
    reconstruct_with<LIMITER>(wm, wc, wp, hdtd, ph, lo, hi, rec, dv);
}

// Conservative update with the non-finite revert (solver_cpu.cpp:386-392).
MHD_DI Cons update(const Cons& u, double dtd, const Cons& fp, const Cons& fm) {
    Cons n{u.r - dtd * (fp.r - fm.r),    u.mn - dtd * (fp.mn - fm.mn), u.mt - dtd * (fp.mt - fm.mt),
           u.mz - dtd * (fp.mz - fm.mz), u.bn - dtd * (fp.bn - fm.bn), u.bt - dtd * (fp.bt - fm.bt),
           u.bz - dtd * (fp.bz - fm.bz), u.e - dtd * (fp.e - fm.e),    u.psi - dtd * (fp.psi - fm.psi)};
    if (!finite(n.r) || !finite(n.e)) return u;
    return n;
}

// The exact candidate of a valid cell (physics.hpp:175-210 for both directions).
MHD_DI double cfl_candidate_exact(const Prim& w, const Phys& ph) {
    const double a2 = ph.gamma * w.p / w.r;
    const double b2 = w.b2 / w.r;
    const double bx2 = w.bn * w.bn / w.r;
    const double by2 = w.bt * w.bt / w.r;
    const double s = (a2 + b2) * (a2 + b2);
    const double dx = s - 4.0 * a2 * bx2;
    const double dy = s - 4.0 * a2 * by2;
    const double cx = sqrt_cr(0.5 * (a2 + b2 + sqrt_cr(dx > 0.0 ? dx : 0.0)));
    const double cy = sqrt_cr(0.5 * (a2 + b2 + sqrt_cr(dy > 0.0 ? dy : 0.0)));
    const double sx = fmax(fabs(w.un) + cx, ph.ch);
    const double sy = fmax(fabs(w.ut) + cy, ph.ch);
    if (finite(sx) && finite(sy)) return std_max(sx, sy);
    return 0.0;
}

// Sufficient test, without division or square root, that the fast speed along
// one direction cannot lift the reference's signal speed above c_h:
// fl(|u_n| + c_f) <= c_h, so max_signal_speed (physics.hpp:197-210), an fmax with
// c_h, returns c_h itself.  In exact arithmetic, with Z = c_h - |u_n| > 0,
//   c_f^2 = (S + sqrt(S^2 - 4 a^2 b_n^2)) / 2 <= Z^2,   S = a^2 + b^2,
// holds iff 2Z^2 >= S and Z^2 (S - Z^2) <= a^2 b_n^2; multiplied by rho^2, with
// Y = Z^2 rho and T = gamma p + |B|^2:  2Y >= T and Y (T - Y) <= gamma p B_n^2
// (Y >= T alone suffices: it bounds c_f by sqrt(S)).
// Margins.  The rounded c_f exceeds the exact one by at most a relative 2.1e-8
// (the discriminant cancels: an absolute error of 15 ulp of S^2 under a square
// root), so Z is taken from c_h (1 - 1e-7).  (Where that cancellation happens,
// a^2 ~ b_n^2 ~ b^2, the second condition is near equality and its own margin
// already demands Z^2 >= c_f^2 (1 + ~2e-6).)  Each rounded product and sum of the
// test is within a few ulp of its exact value; relative margins of 1e-12 absorb
// that, always towards "not proven".  The caller keeps gamma p in [1e-100, 1e100],
// T and rho <= 1e100, and y is held <= 1e100, so that every product a passing test
// relies on is normal and finite (a y that underflows is far below T and fails),
// except B_n^2, whose underflow error (<= 2.5e-224 after the product) the
// absolute margin covers.  Outside those bounds the test fails and the speed is
// evaluated exactly.
MHD_DI bool speed_below_ch(double un, double bn, double gp, double t, double rho, double ch) {
    constexpr double M = 1.0e-12;
    const double z = ch * (1.0 - 1.0e-7) - fabs(un);
    if (!(z > 0.0)) return false;
    const double y = z * z * rho;
    if (!(y <= 1.0e100)) return false;
    if (y >= t * (1.0 + M)) return true;
    if (!(2.0 * y >= t * (1.0 + M))) return false;
    const double lhs = y * ((t - y) + M * (t + y)) * (1.0 + M);
    const double rhs = gp * (bn * bn) * (1.0 - M) - 1.0e-220;
    return lhs <= rhs;
}

// Both directions of a valid cell within the magnitude bounds speed_below_ch needs.
MHD_DI bool speeds_below_ch(const Prim& w, const Phys& ph) {
    const double gp = ph.gamma * w.p;
    const double t = gp + w.b2;
    return gp >= 1.0e-100 && t <= 1.0e100 && w.r <= 1.0e100 && speed_below_ch(w.un, w.bn, gp, t, w.r, ph.ch) &&
           speed_below_ch(w.ut, w.bt, gp, t, w.r, ph.ch);
}

// Candidate of compute_dt for one cell (solver_cpu.cpp:291-300): the larger of the
// x and y signal speeds, or 0 when the cell is skipped.  `u` is in the x frame.
// a^2 and b^2 are the same expressions for both directions (physics.hpp:176-177,
// 185-186) and are formed once.  Cells whose speeds are provably at most c_h
// (speeds_below_ch) return c_h, the value the exact evaluation returns for them,
// without the divisions and square roots; c_h ratchets, so that is most cells.
MHD_DI double cfl_candidate(const Cons& u, const Phys& ph) {
    const Prim w = to_prim(u, ph);
    if (!finite(w.r) || !finite(w.p) || w.r <= 0.0 || w.p <= 0.0) return 0.0;
    if (speeds_below_ch(w, ph)) return ph.ch;
    return cfl_candidate_exact(w, ph);
}

}  // namespace mhd
