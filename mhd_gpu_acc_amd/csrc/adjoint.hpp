// Hand-written discrete adjoint of the point physics in physics.hpp.
//
// Every function here takes the primal inputs of its forward counterpart and the
// adjoint (bar) of the forward's output, recomputes the forward's intermediates,
// and ACCUMULATES (+=) the adjoints of the inputs.  Branches are those the
// forward takes on the same inputs: solver regions and fallbacks, limiter zero
// and selection rules, the physical-state guards, fmin/fmax selections.  dt and
// c_h are constants.  Conventions at non-differentiable points, shared with the
// oracle's tangent-linear model:
//   fmin(a, b), fmax(a, b): the derivative follows a when the result equals a
//       and a is not NaN, else b;
//   std::min/std::max: the derivative follows the selected argument;
//   |x|: derivative sign(x), 0 at x = 0;
//   sqrt(x): derivative 1/(2 sqrt x) for x > 0, 0 at x = 0;
//   a limited slope whose two one-sided differences are both below rounding level
//   (|a| <= 2^-44 (|wc| + |wm|) and |b| <= 2^-44 (|wp| + |wc|)) is differentiated as
//   the zero slope: their signs are noise (a uniform field recovered from conserved
//   variables), the limiter's derivative would depend only on the noise ratio, and
//   no finite difference can observe it.  Values are never changed.
//
// States are in the rotated frame of physics.hpp.  A Prim adjoint carries a
// tenth component, b2 (|B|^2 as to_prim forms it), because flux, fast_speed and
// the Riemann solvers read that intermediate.
#pragma once

#include "physics.hpp"

#pragma clang fp contract(off)

namespace mhd {

MHD_DI Cons cons_zero() { return Cons{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0}; }
MHD_DI Prim prim_zero() { return Prim{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0}; }

MHD_DI void acc(Cons& a, const Cons& b) {
    a.r += b.r; a.mn += b.mn; a.mt += b.mt; a.mz += b.mz; a.bn += b.bn; a.bt += b.bt; a.bz += b.bz;
    a.e += b.e; a.psi += b.psi;
}
MHD_DI void acc_scaled(Cons& a, double s, const Cons& b) {
    a.r += s * b.r; a.mn += s * b.mn; a.mt += s * b.mt; a.mz += s * b.mz; a.bn += s * b.bn;
    a.bt += s * b.bt; a.bz += s * b.bz; a.e += s * b.e; a.psi += s * b.psi;
}
MHD_DI double dot(const Cons& a, const Cons& b) {
    return a.r * b.r + a.mn * b.mn + a.mt * b.mt + a.mz * b.mz + a.bn * b.bn + a.bt * b.bt + a.bz * b.bz +
           a.e * b.e + a.psi * b.psi;
}
// a - b, componentwise
MHD_DI Cons diff(const Cons& a, const Cons& b) {
    return Cons{a.r - b.r,   a.mn - b.mn, a.mt - b.mt, a.mz - b.mz, a.bn - b.bn,
                a.bt - b.bt, a.bz - b.bz, a.e - b.e,   a.psi - b.psi};
}

MHD_DI double d_sqrt(double r, double rb) { return r > 0.0 ? rb / (2.0 * r) : 0.0; }
MHD_DI double sign0(double x) { return x < 0.0 ? -1.0 : (x > 0.0 ? 1.0 : 0.0); }
// Whether fmin(a, b) / fmax(a, b) passes the derivative to a.
MHD_DI bool picks_a(double m, double a) { return m == a && !__builtin_isnan(a); }
MHD_DI bool slope_noise(double wm, double wc, double wp) {
    constexpr double tau = 0x1p-44;
    return fabs(wc - wm) <= tau * (fabs(wc) + fabs(wm)) && fabs(wp - wc) <= tau * (fabs(wp) + fabs(wc));
}

// w = to_prim(u)
MHD_DI void to_prim_adj(const Cons& u, const Prim& w, const Prim& wb, const Phys& ph, Cons& ub) {
    const double inv = 1.0 / u.r;
    const double eb = ph.gm1 * wb.p;          // p = gm1 * (e - ke - 0.5 b2)
    const double keb = -eb;
    const double b2b = wb.b2 - 0.5 * eb;
    const double unb = wb.un + u.r * w.un * keb;  // ke = 0.5 r |v|^2
    const double utb = wb.ut + u.r * w.ut * keb;
    const double uzb = wb.uz + u.r * w.uz * keb;
    const double v2 = w.un * w.un + w.ut * w.ut + w.uz * w.uz;
    ub.r += wb.r + 0.5 * v2 * keb - inv * (w.un * unb + w.ut * utb + w.uz * uzb);
    ub.mn += inv * unb;
    ub.mt += inv * utb;
    ub.mz += inv * uzb;
    ub.bn += wb.bn + 2.0 * u.bn * b2b;
    ub.bt += wb.bt + 2.0 * u.bt * b2b;
    ub.bz += wb.bz + 2.0 * u.bz * b2b;
    ub.e += eb;
    ub.psi += wb.psi;
}

// u = to_cons(w) (b2 recomputed from the field components; w.b2 is not read)
MHD_DI void to_cons_adj(const Prim& w, const Cons& ub, const Phys& ph, Prim& wb) {
    const double eb = ub.e;
    const double v2 = w.un * w.un + w.ut * w.ut + w.uz * w.uz;
    wb.r += ub.r + w.un * ub.mn + w.ut * ub.mt + w.uz * ub.mz + 0.5 * v2 * eb;
    wb.un += w.r * ub.mn + w.r * w.un * eb;
    wb.ut += w.r * ub.mt + w.r * w.ut * eb;
    wb.uz += w.r * ub.mz + w.r * w.uz * eb;
    wb.bn += ub.bn + w.bn * eb;
    wb.bt += ub.bt + w.bt * eb;
    wb.bz += ub.bz + w.bz * eb;
    wb.p += eb / ph.gm1;
    wb.psi += ub.psi;
}

// f = flux(u, w)
MHD_DI void flux_adj(const Cons& u, const Prim& w, const Cons& fb, const Phys& ph, Cons& ub, Prim& wb) {
    const double pt = w.p + 0.5 * w.b2;
    const double bu = u.bn * w.un + u.bt * w.ut + u.bz * w.uz;
    const double ptb = fb.mn + fb.e * w.un;
    const double bub = -u.bn * fb.e;
    ub.mn += fb.r + w.un * fb.mn + w.ut * fb.mt + w.uz * fb.mz;
    wb.un += u.mn * fb.mn + u.bt * fb.bt + u.bz * fb.bz + (u.e + pt) * fb.e + u.bn * bub;
    wb.ut += u.mn * fb.mt - u.bn * fb.bt + u.bt * bub;
    wb.uz += u.mn * fb.mz - u.bn * fb.bz + u.bz * bub;
    ub.bn += -2.0 * u.bn * fb.mn - u.bt * fb.mt - u.bz * fb.mz - w.ut * fb.bt - w.uz * fb.bz - bu * fb.e +
             w.un * bub + ph.ch2 * fb.psi;
    ub.bt += -u.bn * fb.mt + w.un * fb.bt + w.ut * bub;
    ub.bz += -u.bn * fb.mz + w.un * fb.bz + w.uz * bub;
    ub.e += w.un * fb.e;
    ub.psi += fb.bn;
    wb.p += ptb;
    wb.b2 += 0.5 * ptb;
}

// Intermediates of fast_speed(w) (physics.hpp), c bitwise equal to it.
struct FastSpeed {
    double a2, b2, bn2, s, disc, q, c;
};
MHD_DI FastSpeed fast_speed_fwd(const Prim& w, const Phys& ph) {
    FastSpeed f;
    f.a2 = ph.gamma * w.p / w.r;
    f.b2 = w.b2 / w.r;
    f.bn2 = w.bn * w.bn / w.r;
    f.s = f.a2 + f.b2;
    f.disc = f.s * f.s - 4.0 * f.a2 * f.bn2;
    f.q = sqrt_cr(f.disc > 0.0 ? f.disc : 0.0);
    f.c = sqrt_cr(0.5 * (f.s + f.q));
    return f;
}

// c = fast_speed(w), given its intermediates
MHD_DI void fast_speed_adj(const Prim& w, const FastSpeed& f, double cb, const Phys& ph, Prim& wb) {
    const double hb = d_sqrt(f.c, cb);
    double sb = 0.5 * hb;
    const double qb = 0.5 * hb;
    const double discb = f.disc > 0.0 ? d_sqrt(f.q, qb) : 0.0;
    sb += 2.0 * f.s * discb;
    const double a2b = sb - 4.0 * f.bn2 * discb;
    const double b2nb = sb;
    const double bn2b = -4.0 * f.a2 * discb;
    const double inv = 1.0 / w.r;
    wb.p += ph.gamma * a2b * inv;
    wb.b2 += b2nb * inv;
    wb.bn += 2.0 * w.bn * bn2b * inv;
    wb.r -= (f.a2 * a2b + f.b2 * b2nb + f.bn2 * bn2b) * inv;
}
MHD_DI void fast_speed_adj(const Prim& w, double cb, const Phys& ph, Prim& wb) {
    fast_speed_adj(w, fast_speed_fwd(w, ph), cb, ph, wb);
}

// Wave-speed bars of S_L = fmin(vl.un - cfl, vr.un - cfr) and S_R = fmax(vl.un + cfl,
// vr.un + cfr) (optionally against -c_h / +c_h, as HLL bounds them) into the states.
MHD_DI void wave_speeds_adj(const Prim& vl, const Prim& vr, const FastSpeed& fl, const FastSpeed& fr,
                            const Phys& ph, bool with_ch, double slb, double srb, Prim& vlb, Prim& vrb) {
    const double A = vl.un - fl.c, B = vr.un - fr.c, C = vl.un + fl.c, D = vr.un + fr.c;
    const double il = fmin(A, B), ir = fmax(C, D);
    if (with_ch) {
        if (!picks_a(fmin(il, -ph.ch), il)) slb = 0.0;
        if (!picks_a(fmax(ir, ph.ch), ir)) srb = 0.0;
    }
    double cflb = 0.0, cfrb = 0.0;
    if (picks_a(il, A)) { vlb.un += slb; cflb -= slb; } else { vrb.un += slb; cfrb -= slb; }
    if (picks_a(ir, C)) { vlb.un += srb; cflb += srb; } else { vrb.un += srb; cfrb += srb; }
    fast_speed_adj(vl, fl, cflb, ph, vlb);
    fast_speed_adj(vr, fr, cfrb, ph, vrb);
}

// f = hll(ul, vl, fl, ur, vr, fr)
MHD_DI void hll_adj(const Cons& ul, const Prim& vl, const Cons& fl, const Cons& ur, const Prim& vr, const Cons& fr,
                    const Phys& ph, const Cons& fb, Cons& ulb, Prim& vlb, Cons& flb, Cons& urb, Prim& vrb,
                    Cons& frb) {
    if (!prim_physical(vl) || !prim_physical(vr)) {
        acc_scaled(flb, 0.5, fb);
        acc_scaled(frb, 0.5, fb);
        return;
    }
    const FastSpeed tl = fast_speed_fwd(vl, ph);
    const FastSpeed tr = fast_speed_fwd(vr, ph);
    const double cfl = tl.c, cfr = tr.c;
    const double sl = fmin(fmin(vl.un - cfl, vr.un - cfr), -ph.ch);
    const double sr = fmax(fmax(vl.un + cfl, vr.un + cfr), ph.ch);
    if (sl >= 0.0) { acc(flb, fb); return; }
    if (sr <= 0.0) { acc(frb, fb); return; }
    const double den = sr - sl;
    if (fabs(den) < 1.0e-14) {
        acc_scaled(flb, 0.5, fb);
        acc_scaled(frb, 0.5, fb);
        return;
    }
    const double inv = 1.0 / den;
    const double ss = sl * sr;
    const Cons du = diff(ur, ul);
    // F_k = inv * T_k, T_k = sr fl_k - sl fr_k + ss du_k
    double invb = 0.0, srb = 0.0, slb = 0.0, ssb = 0.0;
#define MHD_HLL_ADJ(f)                                    \
    {                                                     \
        const double T = sr * fl.f - sl * fr.f + ss * du.f; \
        invb += fb.f * T;                                 \
        const double Tb = inv * fb.f;                     \
        flb.f += sr * Tb;                                 \
        frb.f -= sl * Tb;                                 \
        urb.f += ss * Tb;                                 \
        ulb.f -= ss * Tb;                                 \
        srb += fl.f * Tb;                                 \
        slb -= fr.f * Tb;                                 \
        ssb += du.f * Tb;                                 \
    }
    MHD_HLL_ADJ(r) MHD_HLL_ADJ(mn) MHD_HLL_ADJ(mt) MHD_HLL_ADJ(mz) MHD_HLL_ADJ(bn) MHD_HLL_ADJ(bt) MHD_HLL_ADJ(bz)
    MHD_HLL_ADJ(e) MHD_HLL_ADJ(psi)
#undef MHD_HLL_ADJ
    slb += sr * ssb;
    srb += sl * ssb;
    const double denb = -inv * inv * invb;
    srb += denb;
    slb -= denb;
    wave_speeds_adj(vl, vr, tl, tr, ph, true, slb, srb, vlb, vrb);
}

// f.bn, f.psi overwritten by glm_overwrite(f, ul, ur): consumes those bars.
MHD_DI void glm_overwrite_adj(Cons& fb, const Phys& ph, Cons& ulb, Cons& urb) {
    if (ph.ch <= 0.0) return;
    const double pb = fb.bn;
    const double bb = ph.ch2 * fb.psi;
    fb.bn = 0.0;
    fb.psi = 0.0;
    ulb.psi += 0.5 * pb + ph.half_over_ch * bb;
    urb.psi += 0.5 * pb - ph.half_over_ch * bb;
    ulb.bn += ph.half_ch * pb + 0.5 * bb;
    urb.bn += -ph.half_ch * pb + 0.5 * bb;
}

// Bars of a SideStar
struct SideStarBar {
    double ut, uw, bt, bw, e, bu;
};
MHD_DI SideStarBar side_bar_zero() { return SideStarBar{0.0, 0.0, 0.0, 0.0, 0.0, 0.0}; }

// o = side_star(u, v, s, sm, pt, ptstar)
MHD_DI void side_star_adj(const Cons& u, const Prim& v, double s, double sm, double pt, double ptstar,
                          const SideStar& o, const SideStarBar& ob, Cons& ub, Prim& vb, double& sb, double& smb,
                          double& ptb, double& ptstarb) {
    const double h = s - v.un;
    const double j = s - sm;
    const double d = v.r * h * j - v.bn * v.bn;
    const double bu = v.bn * v.un + v.bt * v.ut + v.bz * v.uz;
    // e = Num / j  (the reverse uses 1/j and 1/d: its values need not be correctly rounded)
    const double ij = 1.0 / j;
    const double numb = ob.e * ij;
    double jb = -o.e * numb;
    double hb = u.e * numb;
    ub.e += h * numb;
    ptb -= v.un * numb;
    vb.un -= pt * numb;
    ptstarb += sm * numb;
    smb += ptstar * numb;
    vb.bn += (o.bu - bu) * numb;
    const double obub = ob.bu + v.bn * numb;
    const double bub = -v.bn * numb;
    // o.bu = v.bn sm + bt ut + bw uw
    vb.bn += sm * obub;
    smb += v.bn * obub;
    const double btb = ob.bt + o.ut * obub;
    const double utb = ob.ut + o.bt * obub;
    const double bwb = ob.bw + o.uw * obub;
    const double uwb = ob.uw + o.bw * obub;
    // bu = v.bn v.un + v.bt v.ut + v.bz v.uz
    vb.bn += v.un * bub;
    vb.un += v.bn * bub;
    vb.bt += v.ut * bub;
    vb.ut += v.bt * bub;
    vb.bz += v.uz * bub;
    vb.uz += v.bz * bub;
    double db = 0.0;
    if (fabs(d) < 1.0e-14) {
        vb.ut += utb;
        vb.uz += uwb;
        vb.bt += btb;
        vb.bz += bwb;
    } else {
        const double g = sm - v.un;
        const double k = v.r * h * h - v.bn * v.bn;
        const double id = 1.0 / d;
        double gb = 0.0, kb = 0.0;
        // ut = v.ut - P/d, P = v.bn v.bt g
        {
            const double P = v.bn * v.bt * g;
            const double Pb = -utb * id;
            db -= P * id * Pb;
            vb.ut += utb;
            vb.bn += v.bt * g * Pb;
            vb.bt += v.bn * g * Pb;
            gb += v.bn * v.bt * Pb;
        }
        // uw = v.uz - Q/d, Q = v.bn v.bz g
        {
            const double Q = v.bn * v.bz * g;
            const double Qb = -uwb * id;
            db -= Q * id * Qb;
            vb.uz += uwb;
            vb.bn += v.bz * g * Qb;
            vb.bz += v.bn * g * Qb;
            gb += v.bn * v.bz * Qb;
        }
        // bt = v.bt k / d, bw = v.bz k / d
        {
            const double nb = btb * id;
            db -= o.bt * nb;
            vb.bt += k * nb;
            kb += v.bt * nb;
            const double mb = bwb * id;
            db -= o.bw * mb;
            vb.bz += k * mb;
            kb += v.bz * mb;
        }
        // k = v.r h h - v.bn^2
        vb.r += h * h * kb;
        hb += 2.0 * v.r * h * kb;
        vb.bn -= 2.0 * v.bn * kb;
        // g = sm - v.un
        smb += gb;
        vb.un -= gb;
    }
    // d = v.r h j - v.bn^2
    vb.r += h * j * db;
    hb += v.r * j * db;
    jb += v.r * h * db;
    vb.bn -= 2.0 * v.bn * db;
    sb += jb;
    smb -= jb;
    sb += hb;
    vb.un -= hb;
}

// star(rs, sm, ut, uw, bn, bt, bw, e, psi) = {rs, rs sm, rs ut, rs uw, bn, bt, bw, e, psi}
struct StarBar {
    double rs, sm, ut, uw, bn, bt, bw, e, psi;
};
MHD_DI void star_adj(double rs, double sm, double ut, double uw, const Cons& b, StarBar& o) {
    o.rs += b.r + sm * b.mn + ut * b.mt + uw * b.mz;
    o.sm += rs * b.mn;
    o.ut += rs * b.mt;
    o.uw += rs * b.mz;
    o.bn += b.bn;
    o.bt += b.bt;
    o.bw += b.bz;
    o.e += b.e;
    o.psi += b.psi;
}

// Adjoint of the HLLD flux as riemann() uses it, given the bar of riemann()'s result.
// Returns 0 when hlld() falls back to HLL (nothing accumulated), 2 when the HLLD
// flux fails riemann()'s finiteness test (riemann() then returns the HLL flux
// without the GLM overwrite; nothing accumulated), 1 otherwise (the GLM overwrite
// and the HLLD flux are reversed).
template <bool XDIR>
MHD_DI int hlld_adj(const Cons& ul, const Prim& vl, const Cons& ur, const Prim& vr, const Phys& ph, Cons fb,
                    Cons& ulb, Prim& vlb, Cons& urb, Prim& vrb) {
    if (!prim_physical(vl) || !prim_physical(vr)) return 0;
    const FastSpeed tl = fast_speed_fwd(vl, ph);
    const FastSpeed tr = fast_speed_fwd(vr, ph);
    const double cfl = tl.c, cfr = tr.c;
    const double sl = fmin(vl.un - cfl, vr.un - cfr);
    const double sr = fmax(vl.un + cfl, vr.un + cfr);
    if (fabs(sr - sl) < 1.0e-14) return 0;
    const double ptl = vl.p + 0.5 * vl.b2;
    const double ptr = vr.p + 0.5 * vr.b2;
    const double a1 = sl - vl.un;
    const double b1 = sr - vr.un;
    const double num = b1 * vr.r * vr.un - a1 * vl.r * vl.un + ptl - ptr;
    const double den = b1 * vr.r - a1 * vl.r;
    if (fabs(den) < 1.0e-14) return 0;
    const double sm = num / den;
    const double yl = sl - sm, yr = sr - sm;
    const double rls = vl.r * a1 / yl;
    const double rrs = vr.r * b1 / yr;
    if (rls <= 0.0 || rrs <= 0.0 || !finite(rls) || !finite(rrs)) return 0;

    double slb = 0.0, srb = 0.0, smb = 0.0, ptlb = 0.0, ptrb = 0.0, rlsb = 0.0, rrsb = 0.0;
    if (sl >= 0.0 || sr <= 0.0) {
        const bool l = sl >= 0.0;
        const Cons f = l ? flux(ul, vl, ph) : flux(ur, vr, ph);
        const double fbx = XDIR ? f.bn : f.bt;
        if (!finite(f.r) || !finite(f.e) || !finite(fbx) || !finite(f.psi)) return 2;
        glm_overwrite_adj(fb, ph, ulb, urb);
        if (l) flux_adj(ul, vl, fb, ph, ulb, vlb);
        else flux_adj(ur, vr, fb, ph, urb, vrb);
        return 1;
    }
    const double a2 = sm - vl.un;
    const double ptstar = ptl + vl.r * a1 * a2;
    const double sql = sqrt_cr(rls);
    const double sqr = sqrt_cr(rrs);
    const double ql = fabs(vl.bn) / sql;
    const double qr = fabs(vr.bn) / sqr;
    const double sls = sm - ql;
    const double srs = sm + qr;
    const double psim = 0.5 * (vl.psi + vr.psi) - ph.half_ch * (vr.bn - vl.bn);
    const SideStar L = side_star(ul, vl, sl, sm, ptl, ptstar);
    const SideStar R = side_star(ur, vr, sr, sm, ptr, ptstar);
    const bool left = sm >= 0.0;
    const bool outer = left ? (sls >= 0.0) : (srs <= 0.0);
    const Cons& us = left ? ul : ur;
    const Prim& vs = left ? vl : vr;
    const SideStar& S = left ? L : R;
    const double s = left ? sl : sr;
    const double rs = left ? rls : rrs;
    const double bn = left ? vl.bn : vr.bn;
    const Cons fs = flux(us, vs, ph);
    const Cons u1 = star(rs, sm, S.ut, S.uw, bn, S.bt, S.bw, S.e, psim);
    const Cons f1 = fan(fs, s, u1, us);

    // Double-star states, and the flux riemann() tests for finiteness.
    const double sgn = (vl.bn >= 0.0) ? 1.0 : -1.0;
    const bool degen = fabs(sql + sqr) < 1.0e-14;
    double inv = 0.0, ut2 = S.ut, uw2 = S.uw, bt2 = S.bt, bw2 = S.bw, e2 = S.e, bu2 = 0.0;
    double N1 = 0.0, N2 = 0.0, N3 = 0.0, N4 = 0.0, M = 0.0;
    Cons u2 = u1, f = f1;
    if (!outer) {
        if (!degen) {
            inv = 1.0 / (sql + sqr);
            M = sgn * sql * sqr;
            N1 = sql * L.ut + sqr * R.ut + sgn * (R.bt - L.bt);
            N2 = sql * L.uw + sqr * R.uw + sgn * (R.bw - L.bw);
            N3 = sql * R.bt + sqr * L.bt + M * (R.ut - L.ut);
            N4 = sql * R.bw + sqr * L.bw + M * (R.uw - L.uw);
            ut2 = N1 * inv;
            uw2 = N2 * inv;
            bt2 = N3 * inv;
            bw2 = N4 * inv;
            bu2 = bn * sm + bt2 * ut2 + bw2 * uw2;
            e2 = left ? (S.e - sql * sgn * (bu2 - S.bu)) : (S.e + sqr * sgn * (bu2 - S.bu));
        }
        u2 = star(rs, sm, ut2, uw2, bn, bt2, bw2, e2, psim);
        f = fan(f1, left ? sls : srs, u2, u1);
    }
    {
        const double fbx = XDIR ? f.bn : f.bt;
        if (!finite(f.r) || !finite(f.e) || !finite(fbx) || !finite(f.psi)) return 2;
    }
    glm_overwrite_adj(fb, ph, ulb, urb);

    // ---- reverse ----
    StarBar sb1{0, 0, 0, 0, 0, 0, 0, 0, 0};   // bars of u1's arguments
    SideStarBar Lb = side_bar_zero(), Rb = side_bar_zero();
    double sqlb = 0.0, sqrb = 0.0, slsb = 0.0, srsb = 0.0;
    Cons u1b = cons_zero();
    if (!outer) {
        // f = f1 + s2 (u2 - u1)
        const double s2 = left ? sls : srs;
        const double s2b = dot(fb, diff(u2, u1));
        Cons u2b = cons_zero();
        acc_scaled(u2b, s2, fb);
        acc_scaled(u1b, -s2, fb);
        if (left) slsb += s2b; else srsb += s2b;
        StarBar sb2{0, 0, 0, 0, 0, 0, 0, 0, 0};
        star_adj(rs, sm, ut2, uw2, u2b, sb2);
        rlsb += left ? sb2.rs : 0.0;
        rrsb += left ? 0.0 : sb2.rs;
        smb += sb2.sm;
        sb1.bn += sb2.bn;
        sb1.psi += sb2.psi;
        SideStarBar& Sb = left ? Lb : Rb;
        if (degen) {
            Sb.ut += sb2.ut; Sb.uw += sb2.uw; Sb.bt += sb2.bt; Sb.bw += sb2.bw; Sb.e += sb2.e;
        } else {
            double ut2b = sb2.ut, uw2b = sb2.uw, bt2b = sb2.bt, bw2b = sb2.bw;
            const double e2b = sb2.e;
            const double dbu = bu2 - S.bu;
            Sb.e += e2b;
            double bu2b;
            if (left) {
                sqlb -= sgn * dbu * e2b;
                bu2b = -(sql * sgn) * e2b;
            } else {
                sqrb += sgn * dbu * e2b;
                bu2b = (sqr * sgn) * e2b;
            }
            Sb.bu -= bu2b;
            // bu2 = bn sm + bt2 ut2 + bw2 uw2
            sb1.bn += sm * bu2b;
            smb += bn * bu2b;
            bt2b += ut2 * bu2b;
            ut2b += bt2 * bu2b;
            bw2b += uw2 * bu2b;
            uw2b += bw2 * bu2b;
            const double N1b = ut2b * inv, N2b = uw2b * inv, N3b = bt2b * inv, N4b = bw2b * inv;
            const double invb = N1 * ut2b + N2 * uw2b + N3 * bt2b + N4 * bw2b;
            // N1 = sql L.ut + sqr R.ut + sgn (R.bt - L.bt)
            sqlb += L.ut * N1b; Lb.ut += sql * N1b; sqrb += R.ut * N1b; Rb.ut += sqr * N1b;
            Rb.bt += sgn * N1b; Lb.bt -= sgn * N1b;
            // N2 = sql L.uw + sqr R.uw + sgn (R.bw - L.bw)
            sqlb += L.uw * N2b; Lb.uw += sql * N2b; sqrb += R.uw * N2b; Rb.uw += sqr * N2b;
            Rb.bw += sgn * N2b; Lb.bw -= sgn * N2b;
            // N3 = sql R.bt + sqr L.bt + M (R.ut - L.ut)
            double Mb = (R.ut - L.ut) * N3b + (R.uw - L.uw) * N4b;
            sqlb += R.bt * N3b; Rb.bt += sql * N3b; sqrb += L.bt * N3b; Lb.bt += sqr * N3b;
            Rb.ut += M * N3b; Lb.ut -= M * N3b;
            // N4 = sql R.bw + sqr L.bw + M (R.uw - L.uw)
            sqlb += R.bw * N4b; Rb.bw += sql * N4b; sqrb += L.bw * N4b; Lb.bw += sqr * N4b;
            Rb.uw += M * N4b; Lb.uw -= M * N4b;
            // M = sgn sql sqr
            sqlb += sgn * sqr * Mb;
            sqrb += sgn * sql * Mb;
            // inv = 1 / (sql + sqr)
            const double tb = -inv * inv * invb;
            sqlb += tb;
            sqrb += tb;
        }
    }
    // f1 = fs + s (u1 - us)
    const double sbar = dot(fb, diff(u1, us));
    acc_scaled(u1b, s, fb);
    Cons usb = cons_zero();
    acc_scaled(usb, -s, fb);
    star_adj(rs, sm, S.ut, S.uw, u1b, sb1);
    {
        SideStarBar& Sb = left ? Lb : Rb;
        Sb.ut += sb1.ut; Sb.uw += sb1.uw; Sb.bt += sb1.bt; Sb.bw += sb1.bw; Sb.e += sb1.e;
    }
    smb += sb1.sm;
    if (left) {
        flux_adj(ul, vl, fb, ph, ulb, vlb);
        acc(ulb, usb);
        slb += sbar;
        rlsb += sb1.rs;
        vlb.bn += sb1.bn;
    } else {
        flux_adj(ur, vr, fb, ph, urb, vrb);
        acc(urb, usb);
        srb += sbar;
        rrsb += sb1.rs;
        vrb.bn += sb1.bn;
    }
    // side stars (in a single-star region only the interface's side carries a bar)
    double ptstarb = 0.0;
    if (!outer || left) side_star_adj(ul, vl, sl, sm, ptl, ptstar, L, Lb, ulb, vlb, slb, smb, ptlb, ptstarb);
    if (!outer || !left) side_star_adj(ur, vr, sr, sm, ptr, ptstar, R, Rb, urb, vrb, srb, smb, ptrb, ptstarb);
    // psim = 0.5 (vl.psi + vr.psi) - half_ch (vr.bn - vl.bn)
    vlb.psi += 0.5 * sb1.psi;
    vrb.psi += 0.5 * sb1.psi;
    vrb.bn -= ph.half_ch * sb1.psi;
    vlb.bn += ph.half_ch * sb1.psi;
    // sls = sm - |vl.bn| / sql, srs = sm + |vr.bn| / sqr
    smb += slsb + srsb;
    const double isql = 1.0 / sql, isqr = 1.0 / sqr;  // rls, rrs > 0 here
    {
        const double qlb = -slsb * isql;
        vlb.bn += sign0(vl.bn) * qlb;
        sqlb -= ql * qlb;
        const double qrb = srsb * isqr;
        vrb.bn += sign0(vr.bn) * qrb;
        sqrb -= qr * qrb;
    }
    rlsb += 0.5 * sqlb * isql;
    rrsb += 0.5 * sqrb * isqr;
    // ptstar = ptl + vl.r a1 a2
    double a1b = 0.0, b1b = 0.0;
    ptlb += ptstarb;
    vlb.r += a1 * a2 * ptstarb;
    a1b += vl.r * a2 * ptstarb;
    {
        const double a2b = vl.r * a1 * ptstarb;
        smb += a2b;
        vlb.un -= a2b;
    }
    // rls = vl.r a1 / yl, rrs = vr.r b1 / yr
    {
        const double xlb = rlsb / yl;
        const double ylb = -rls * xlb;
        vlb.r += a1 * xlb;
        a1b += vl.r * xlb;
        slb += ylb;
        smb -= ylb;
        const double xrb = rrsb / yr;
        const double yrb = -rrs * xrb;
        vrb.r += b1 * xrb;
        b1b += vr.r * xrb;
        srb += yrb;
        smb -= yrb;
    }
    // sm = num / den
    const double numb = smb / den;
    const double denb = -sm * numb;
    // num = b1 vr.r vr.un - a1 vl.r vl.un + ptl - ptr
    b1b += vr.r * vr.un * numb;
    vrb.r += b1 * vr.un * numb;
    vrb.un += b1 * vr.r * numb;
    a1b -= vl.r * vl.un * numb;
    vlb.r -= a1 * vl.un * numb;
    vlb.un -= a1 * vl.r * numb;
    ptlb += numb;
    ptrb -= numb;
    // den = b1 vr.r - a1 vl.r
    b1b += vr.r * denb;
    vrb.r += b1 * denb;
    a1b -= vl.r * denb;
    vlb.r -= a1 * denb;
    // a1 = sl - vl.un, b1 = sr - vr.un
    slb += a1b;
    vlb.un -= a1b;
    srb += b1b;
    vrb.un -= b1b;
    // pt = p + 0.5 b2
    vlb.p += ptlb;
    vlb.b2 += 0.5 * ptlb;
    vrb.p += ptrb;
    vrb.b2 += 0.5 * ptrb;
    wave_speeds_adj(vl, vr, tl, tr, ph, false, slb, srb, vlb, vrb);
    return 1;
}

// f = force(ul, vl, fl, ur, vr, fr): half the Lax-Friedrichs flux with speed a plus
// half the flux of the Richtmyer state uri (physics.hpp force).
MHD_DI void force_adj(const Cons& ul, const Prim& vl, const Cons& fl, const Cons& ur, const Prim& vr, const Cons& fr,
                      const Phys& ph, const Cons& fb, Cons& ulb, Prim& vlb, Cons& flb, Cons& urb, Prim& vrb,
                      Cons& frb) {
    if (!prim_physical(vl) || !prim_physical(vr)) {
        acc_scaled(flb, 0.5, fb);
        acc_scaled(frb, 0.5, fb);
        return;
    }
    const FastSpeed tl = fast_speed_fwd(vl, ph);
    const FastSpeed tr = fast_speed_fwd(vr, ph);
    const double A = fabs(vl.un) + tl.c, B = fabs(vr.un) + tr.c;
    const double m = fmax(A, B);
    const double a = fmax(m, ph.ch);
    if (a < 1.0e-14) {
        acc_scaled(flb, 0.5, fb);
        acc_scaled(frb, 0.5, fb);
        return;
    }
    const double ha = 0.5 * a;
    const double hia = 0.5 * (1.0 / a);
    const Cons uri{0.5 * (ul.r + ur.r) - hia * (fr.r - fl.r),       0.5 * (ul.mn + ur.mn) - hia * (fr.mn - fl.mn),
                   0.5 * (ul.mt + ur.mt) - hia * (fr.mt - fl.mt),   0.5 * (ul.mz + ur.mz) - hia * (fr.mz - fl.mz),
                   0.5 * (ul.bn + ur.bn) - hia * (fr.bn - fl.bn),   0.5 * (ul.bt + ur.bt) - hia * (fr.bt - fl.bt),
                   0.5 * (ul.bz + ur.bz) - hia * (fr.bz - fl.bz),   0.5 * (ul.e + ur.e) - hia * (fr.e - fl.e),
                   0.5 * (ul.psi + ur.psi) - hia * (fr.psi - fl.psi)};
    const Prim vri = to_prim(uri, ph);
    Cons flfb = fb, urib = cons_zero();
    if (prim_physical(vri)) {
        // f = 0.5 (flf + flux(uri, vri))
        flfb = cons_zero();
        acc_scaled(flfb, 0.5, fb);
        Cons fxb = cons_zero();
        acc_scaled(fxb, 0.5, fb);
        Prim vrib = prim_zero();
        flux_adj(uri, vri, fxb, ph, urib, vrib);
        to_prim_adj(uri, vri, vrib, ph, urib);
    }
    double hab = 0.0, hiab = 0.0;
#define MHD_FORCE_ADJ(f)                               \
    {                                                  \
        flb.f += 0.5 * flfb.f;                         \
        frb.f += 0.5 * flfb.f;                         \
        urb.f -= ha * flfb.f;                          \
        ulb.f += ha * flfb.f;                          \
        hab -= (ur.f - ul.f) * flfb.f;                 \
        ulb.f += 0.5 * urib.f;                         \
        urb.f += 0.5 * urib.f;                         \
        frb.f -= hia * urib.f;                         \
        flb.f += hia * urib.f;                         \
        hiab -= (fr.f - fl.f) * urib.f;                \
    }
    MHD_FORCE_ADJ(r) MHD_FORCE_ADJ(mn) MHD_FORCE_ADJ(mt) MHD_FORCE_ADJ(mz) MHD_FORCE_ADJ(bn) MHD_FORCE_ADJ(bt)
    MHD_FORCE_ADJ(bz) MHD_FORCE_ADJ(e) MHD_FORCE_ADJ(psi)
#undef MHD_FORCE_ADJ
    // ha = 0.5 a, hia = 0.5 (1/a); a = fmax(fmax(A, B), c_h)
    const double ab = 0.5 * hab - 0.5 * hiab / (a * a);
    if (!picks_a(a, m)) return;  // a = c_h, a constant
    double cflb = 0.0, cfrb = 0.0;
    if (picks_a(m, A)) {
        vlb.un += sign0(vl.un) * ab;
        cflb = ab;
    } else {
        vrb.un += sign0(vr.un) * ab;
        cfrb = ab;
    }
    fast_speed_adj(vl, tl, cflb, ph, vlb);
    fast_speed_adj(vr, tr, cfrb, ph, vrb);
}

// Adjoint of the HLLC flux as riemann() uses it (physics.hpp hllc), given the bar of
// riemann()'s result.  Return codes as hlld_adj: 0 when hllc() falls back to HLL, 2
// when its flux fails riemann()'s finiteness test (nothing accumulated in both), 1
// otherwise (the GLM overwrite and the HLLC flux are reversed).
template <bool XDIR>
MHD_DI int hllc_adj(const Cons& ul, const Prim& vl, const Cons& ur, const Prim& vr, const Phys& ph, Cons fb,
                    Cons& ulb, Prim& vlb, Cons& urb, Prim& vrb) {
    if (!prim_physical(vl) || !prim_physical(vr)) return 0;
    const FastSpeed tl = fast_speed_fwd(vl, ph);
    const FastSpeed tr = fast_speed_fwd(vr, ph);
    const double sl = fmin(vl.un - tl.c, vr.un - tr.c);
    const double sr = fmax(vl.un + tl.c, vr.un + tr.c);
    if (sl >= 0.0 || sr <= 0.0) {
        const bool l = sl >= 0.0;
        const Cons f = l ? flux(ul, vl, ph) : flux(ur, vr, ph);
        const double fbx = XDIR ? f.bn : f.bt;
        if (!finite(f.r) || !finite(f.e) || !finite(fbx) || !finite(f.psi)) return 2;
        glm_overwrite_adj(fb, ph, ulb, urb);
        if (l) flux_adj(ul, vl, fb, ph, ulb, vlb);
        else flux_adj(ur, vr, fb, ph, urb, vrb);
        return 1;
    }
    const double ptl = vl.p + 0.5 * vl.b2;
    const double ptr = vr.p + 0.5 * vr.b2;
    const double a1 = sl - vl.un;
    const double b1 = sr - vr.un;
    const double num = b1 * vr.r * vr.un - a1 * vl.r * vl.un + ptl - ptr;
    const double den = b1 * vr.r - a1 * vl.r;
    if (fabs(den) < 1.0e-14) return 0;
    const double sm = num / den;
    const double yl = sl - sm, yr = sr - sm;
    const double rls = vl.r * a1 / yl;
    const double rrs = vr.r * b1 / yr;
    if (rls <= 0.0 || rrs <= 0.0 || !finite(rls) || !finite(rrs)) return 0;
    const double psim = 0.5 * (vl.psi + vr.psi) - ph.half_ch * (vr.bn - vl.bn);
    const bool left = sm >= 0.0;
    const Cons& u = left ? ul : ur;
    const Prim& v = left ? vl : vr;
    const double s = left ? sl : sr;
    const double pt = left ? ptl : ptr;
    const double rs = left ? rls : rrs;
    const double c1 = s - v.un, c2 = sm - v.un;
    const double pts = pt + v.r * c1 * c2;
    const double bu = v.bn * v.un + v.bt * v.ut + v.bz * v.uz;
    const double bus = v.bn * sm + v.bt * v.ut + v.bz * v.uz;
    const double ys = s - sm;
    const double Ne = c1 * u.e - pt * v.un + pts * sm + v.bn * (bus - bu);
    const double es = Ne / ys;
    const Cons us = star(rs, sm, v.ut, v.uz, v.bn, v.bt, v.bz, es, psim);
    const Cons f = fan(flux(u, v, ph), s, us, u);
    {
        const double fbx = XDIR ? f.bn : f.bt;
        if (!finite(f.r) || !finite(f.e) || !finite(fbx) || !finite(f.psi)) return 2;
    }
    glm_overwrite_adj(fb, ph, ulb, urb);

    // ---- reverse ----  f = flux(u, v) + s (us - u)
    Cons& ub = left ? ulb : urb;
    Prim& vb = left ? vlb : vrb;
    double sb = dot(fb, diff(us, u));
    Cons usb = cons_zero();
    acc_scaled(usb, s, fb);
    acc_scaled(ub, -s, fb);
    flux_adj(u, v, fb, ph, ub, vb);
    StarBar st{0, 0, 0, 0, 0, 0, 0, 0, 0};
    star_adj(rs, sm, v.ut, v.uz, usb, st);
    vb.ut += st.ut;
    vb.uz += st.uw;
    vb.bn += st.bn;
    vb.bt += st.bt;
    vb.bz += st.bw;
    double smb = st.sm, ptb = 0.0;
    // es = Ne / ys, ys = s - sm
    const double Neb = st.e / ys;
    const double ysb = -es * Neb;
    sb += ysb;
    smb -= ysb;
    // Ne = c1 u.e - pt v.un + pts sm + v.bn (bus - bu)
    double c1b = u.e * Neb, c2b = 0.0;
    ub.e += c1 * Neb;
    ptb -= v.un * Neb;
    vb.un -= pt * Neb;
    const double ptsb = sm * Neb;
    smb += pts * Neb;
    vb.bn += (bus - bu) * Neb;
    const double busb = v.bn * Neb, bub = -v.bn * Neb;
    // bus = v.bn sm + v.bt v.ut + v.bz v.uz;  bu = v.bn v.un + v.bt v.ut + v.bz v.uz
    vb.bn += sm * busb + v.un * bub;
    smb += v.bn * busb;
    vb.un += v.bn * bub;
    vb.bt += v.ut * (busb + bub);
    vb.ut += v.bt * (busb + bub);
    vb.bz += v.uz * (busb + bub);
    vb.uz += v.bz * (busb + bub);
    // pts = pt + v.r c1 c2
    ptb += ptsb;
    vb.r += c1 * c2 * ptsb;
    c1b += v.r * c2 * ptsb;
    c2b += v.r * c1 * ptsb;
    // c1 = s - v.un, c2 = sm - v.un
    sb += c1b;
    vb.un -= c1b + c2b;
    smb += c2b;
    // psim = 0.5 (vl.psi + vr.psi) - half_ch (vr.bn - vl.bn)
    vlb.psi += 0.5 * st.psi;
    vrb.psi += 0.5 * st.psi;
    vrb.bn -= ph.half_ch * st.psi;
    vlb.bn += ph.half_ch * st.psi;
    double slb = left ? sb : 0.0, srb = left ? 0.0 : sb;
    double ptlb = left ? ptb : 0.0, ptrb = left ? 0.0 : ptb;
    const double rlsb = left ? st.rs : 0.0, rrsb = left ? 0.0 : st.rs;
    // rls = vl.r a1 / yl, rrs = vr.r b1 / yr
    double a1b = 0.0, b1b = 0.0;
    {
        const double xlb = rlsb / yl;
        const double ylb = -rls * xlb;
        vlb.r += a1 * xlb;
        a1b += vl.r * xlb;
        slb += ylb;
        smb -= ylb;
        const double xrb = rrsb / yr;
        const double yrb = -rrs * xrb;
        vrb.r += b1 * xrb;
        b1b += vr.r * xrb;
        srb += yrb;
        smb -= yrb;
    }
    // sm = num / den
    const double numb = smb / den;
    const double denb = -sm * numb;
    b1b += vr.r * vr.un * numb;
    vrb.r += b1 * vr.un * numb;
    vrb.un += b1 * vr.r * numb;
    a1b -= vl.r * vl.un * numb;
    vlb.r -= a1 * vl.un * numb;
    vlb.un -= a1 * vl.r * numb;
    ptlb += numb;
    ptrb -= numb;
    b1b += vr.r * denb;
    vrb.r += b1 * denb;
    a1b -= vl.r * denb;
    vlb.r -= a1 * denb;
    slb += a1b;
    vlb.un -= a1b;
    srb += b1b;
    vrb.un -= b1b;
    vlb.p += ptlb;
    vlb.b2 += 0.5 * ptlb;
    vrb.p += ptrb;
    vrb.b2 += 0.5 * ptrb;
    wave_speeds_adj(vl, vr, tl, tr, ph, false, slb, srb, vlb, vrb);
    return 1;
}

// Adjoint of riemann<SOLVER, XDIR>() for every solver: 0 HLL, 1 HLLC, 2 HLLD, 3 FORCE.
// Bars of the two interface states are accumulated as (cons, prim) pairs.
template <int SOLVER, bool XDIR>
MHD_DI void riemann_adj(const Cons& ul, const Prim& vl, const Cons& ur, const Prim& vr, const Phys& ph,
                        const Cons& fb, Cons& ulb, Prim& vlb, Cons& urb, Prim& vrb) {
    static_assert(SOLVER >= 0 && SOLVER <= 3, "adjoint: unknown solver");
    if (SOLVER == 3) {
        const Cons fl = flux(ul, vl, ph);
        const Cons fr = flux(ur, vr, ph);
        Cons flb = cons_zero(), frb = cons_zero();
        const Cons f = force(ul, vl, fl, ur, vr, fr, ph);
        const double fbx = XDIR ? f.bn : f.bt;
        if (!finite(f.r) || !finite(f.e) || !finite(fbx) || !finite(f.psi)) {
            hll_adj(ul, vl, fl, ur, vr, fr, ph, fb, ulb, vlb, flb, urb, vrb, frb);  // riemann()'s HLL fallback
        } else {
            Cons g = fb;
            glm_overwrite_adj(g, ph, ulb, urb);
            force_adj(ul, vl, fl, ur, vr, fr, ph, g, ulb, vlb, flb, urb, vrb, frb);
        }
        flux_adj(ul, vl, flb, ph, ulb, vlb);
        flux_adj(ur, vr, frb, ph, urb, vrb);
        return;
    }
    if (SOLVER == 1 || SOLVER == 2) {
        const int code = SOLVER == 1 ? hllc_adj<XDIR>(ul, vl, ur, vr, ph, fb, ulb, vlb, urb, vrb)
                                     : hlld_adj<XDIR>(ul, vl, ur, vr, ph, fb, ulb, vlb, urb, vrb);
        if (code == 1) return;
        const Cons fl = flux(ul, vl, ph);
        const Cons fr = flux(ur, vr, ph);
        Cons flb = cons_zero(), frb = cons_zero();
        Cons g = fb;
        if (code == 0) {
            // HLL fallback of hllc()/hlld(): finiteness test, then the GLM overwrite
            const Cons f = hll(ul, vl, fl, ur, vr, fr, ph);
            const double fbx = XDIR ? f.bn : f.bt;
            if (finite(f.r) && finite(f.e) && finite(fbx) && finite(f.psi)) glm_overwrite_adj(g, ph, ulb, urb);
        }
        hll_adj(ul, vl, fl, ur, vr, fr, ph, g, ulb, vlb, flb, urb, vrb, frb);
        flux_adj(ul, vl, flb, ph, ulb, vlb);
        flux_adj(ur, vr, frb, ph, urb, vrb);
        return;
    }
    const Cons fl = flux(ul, vl, ph);
    const Cons fr = flux(ur, vr, ph);
    Cons flb = cons_zero(), frb = cons_zero();
    hll_adj(ul, vl, fl, ur, vr, fr, ph, fb, ulb, vlb, flb, urb, vrb, frb);
    flux_adj(ul, vl, flb, ph, ulb, vlb);
    flux_adj(ur, vr, frb, ph, urb, vrb);
}

// s = limit<LIMITER>(a, b)
template <int LIMITER> MHD_DI void limit_adj(double a, double b, double sb, double& ab, double& bb) {
    if (a * b <= 0.0) return;
    if (LIMITER == 0) {
        if (a > 0.0) { if (b < a) bb += sb; else ab += sb; }
        else { if (a < b) bb += sb; else ab += sb; }
        return;
    }
    const double iD = 1.0 / (a + b);
    const double Nb = sb * iD;
    const double Db = -(2.0 * a * b) * iD * Nb;
    ab += 2.0 * b * Nb + Db;
    bb += 2.0 * a * Nb + Db;
}

// Adjoint of reconstruct<LIMITER>(wm, wc, wp) given the bars of the two interface
// states as conserved vectors (the bar of a face's primitive folded in by
// to_prim_adj), from reconstruct()'s record of its forward values.  Accumulates into
// the nine primitive slots of wmb, wcb, wpb.
template <int LIMITER>
MHD_DI void reconstruct_adj_rec(const Prim& wm, const Prim& wc, const Prim& wp, double hdtd, const Phys& ph,
                                const ReconRecord& rec, const Cons& lob, const Cons& hib, Prim& wmb, Prim& wcb,
                                Prim& wpb) {
    const bool pl = rec.pl, pr = rec.pr, pcl = rec.pcl, pcr = rec.pcr;
    // the guarded interface primitives, bit for bit as reconstruct() formed them
    Prim wl, wr;
#define MHD_SIDES(k, f)                                \
    {                                                  \
        wl.f = pl ? wc.f - 0.5 * rec.s[k] : wc.f;      \
        wr.f = pr ? wc.f + 0.5 * rec.s[k] : wc.f;      \
    }
    MHD_SIDES(0, r) MHD_SIDES(1, un) MHD_SIDES(2, ut) MHD_SIDES(3, uz) MHD_SIDES(4, bn) MHD_SIDES(5, bt)
    MHD_SIDES(6, bz) MHD_SIDES(7, p) MHD_SIDES(8, psi)
#undef MHD_SIDES
    wl.b2 = 0.0;
    wr.b2 = 0.0;
    const Cons& ul = rec.ul;
    const Cons& ur = rec.ur;
    const Prim vl = to_prim(ul, ph);
    const Prim vr = to_prim(ur, ph);
    // lo = pcl ? cl : ul, hi = pcr ? cr : ur; cl = ul - h, cr = ur - h
    Cons ulb = cons_zero(), urb = cons_zero(), hb = cons_zero();
    if (pcl) { acc(ulb, lob); acc_scaled(hb, -1.0, lob); } else acc(ulb, lob);
    if (pcr) { acc(urb, hib); acc_scaled(hb, -1.0, hib); } else acc(urb, hib);
    // h = hdtd (fr - fl)
    Cons frb = cons_zero(), flb = cons_zero();
    acc_scaled(frb, hdtd, hb);
    acc_scaled(flb, -hdtd, hb);
    Prim vlb = prim_zero(), vrb = prim_zero();
    flux_adj(ul, vl, flb, ph, ulb, vlb);
    flux_adj(ur, vr, frb, ph, urb, vrb);
    to_prim_adj(ul, vl, vlb, ph, ulb);
    to_prim_adj(ur, vr, vrb, ph, urb);
    Prim wlb = prim_zero(), wrb = prim_zero();
    to_cons_adj(wl, ulb, ph, wlb);
    to_cons_adj(wr, urb, ph, wrb);
#define MHD_SLOPE_ADJ(f)                                         \
    {                                                            \
        double sb = 0.0;                                         \
        wcb.f += wlb.f;                                          \
        if (pl) sb -= 0.5 * wlb.f;                               \
        wcb.f += wrb.f;                                          \
        if (pr) sb += 0.5 * wrb.f;                               \
        double ab = 0.0, bb = 0.0;                               \
        if (!slope_noise(wm.f, wc.f, wp.f))                      \
            limit_adj<LIMITER>(wc.f - wm.f, wp.f - wc.f, sb, ab, bb); \
        wcb.f += ab - bb;                                        \
        wmb.f -= ab;                                             \
        wpb.f += bb;                                             \
    }
    MHD_SLOPE_ADJ(r) MHD_SLOPE_ADJ(un) MHD_SLOPE_ADJ(ut) MHD_SLOPE_ADJ(uz) MHD_SLOPE_ADJ(bn) MHD_SLOPE_ADJ(bt)
    MHD_SLOPE_ADJ(bz) MHD_SLOPE_ADJ(p) MHD_SLOPE_ADJ(psi)
#undef MHD_SLOPE_ADJ
}

// The same, recomputing the forward values.
template <int LIMITER>
MHD_DI void reconstruct_adj(const Prim& wm, const Prim& wc, const Prim& wp, double hdtd, const Phys& ph,
                            const Cons& lob, const Cons& hib, Prim& wmb, Prim& wcb, Prim& wpb) {
    Face lo, hi;
    ReconRecord rec;
    reconstruct<LIMITER>(wm, wc, wp, hdtd, ph, lo, hi, &rec);
    reconstruct_adj_rec<LIMITER>(wm, wc, wp, hdtd, ph, rec, lob, hib, wmb, wcb, wpb);
}

}  // namespace mhd
