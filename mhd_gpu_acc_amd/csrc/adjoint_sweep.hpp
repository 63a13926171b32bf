// Sweep-level pieces of the discrete adjoint, shared by the device kernels
// (adjoint.hip) and the host build of the same arithmetic (hostcheck.cpp).
//
// The forward step (solver_cpu.cpp:327-440) is, per sweep,
//   U1[c] = U[c] - dtd (F[c+1/2] - F[c-1/2])      for interior cells c,
//   F[c+1/2] = riemann(hi[c], lo[c+1]),  (lo[c], hi[c]) = reconstruct(W[c-1], W[c], W[c+1]),
//   W = to_prim(U),
// followed by the ghost fill of U1 from its own interior (and, after the y
// sweep, psi damping of the interior).  Given g[c], the adjoint of U1 at interior
// cells with the ghost-cell adjoints folded onto their source cells, the adjoint
// of U along one line of the sweep is
//   Fb[c+1/2] = dtd (g[c+1] - g[c])
//   (hib[c], lob[c+1]) = riemann^T(Fb[c+1/2])
//   (Wmb[c], Wcb[c], Wpb[c]) = reconstruct^T(lob[c], hib[c])
//   Ub[c] = to_prim^T((Wpb[c-1] + Wcb[c]) + Wmb[c+1]) + [c interior] g[c].
// Every quantity is a function of its position alone, so a tiled device sweep and
// a plain host loop produce the same bits.
#pragma once

#include "adjoint.hpp"
#include "sweeps.hpp"

#pragma clang fp contract(off)

namespace mhd {

// Roles of the cell at index a along a sweep line with n interior cells; `line`
// says whether the line itself is an interior row (x) or column (y).
struct AdjFlags {
    bool interior;  // updated by the sweep
    bool face;      // the face (a | a+1) enters an interior update
    bool rec;       // the cell's interface states enter such a face
};
MHD_DI AdjFlags adj_flags(int a, int n, bool line) {
    AdjFlags f;
    f.interior = line && a >= NG && a < NG + n;
    f.face = line && a >= NG - 1 && a < NG + n;
    f.rec = line && a >= NG - 1 && a < NG + n + 1;
    return f;
}

// Adjoint of the face flux (a | a+1): hi = upper interface state of cell a, lo the
// lower one of cell a+1; gown/gnext the output adjoints of the two cells.
template <int SOLVER, bool XDIR>
MHD_DI void adj_face(const Face& hi, const Face& lo, const Cons& gown, const Cons& gnext, double dtd,
                     const Phys& ph, Cons& hib, Cons& lob) {
    const Cons fb{dtd * (gnext.r - gown.r),   dtd * (gnext.mn - gown.mn), dtd * (gnext.mt - gown.mt),
                  dtd * (gnext.mz - gown.mz), dtd * (gnext.bn - gown.bn), dtd * (gnext.bt - gown.bt),
                  dtd * (gnext.bz - gown.bz), dtd * (gnext.e - gown.e),   dtd * (gnext.psi - gown.psi)};
    hib = cons_zero();
    lob = cons_zero();
    Prim hwb = prim_zero(), lwb = prim_zero();
    riemann_adj<SOLVER, XDIR>(hi.u, hi.w, lo.u, lo.w, ph, fb, hib, hwb, lob, lwb);
    to_prim_adj(hi.u, hi.w, hwb, ph, hib);
    to_prim_adj(lo.u, lo.w, lwb, ph, lob);
}

// Output adjoint of cell a: pprev = Wpb[a-1], cown = Wcb[a], mnext = Wmb[a+1].
MHD_DI Cons adj_cell(const Cons& u, const Prim& w, const Prim& pprev, const Prim& cown, const Prim& mnext,
                     bool interior, const Cons& g, const Phys& ph) {
    Prim lw;
    lw.r = (pprev.r + cown.r) + mnext.r;
    lw.un = (pprev.un + cown.un) + mnext.un;
    lw.ut = (pprev.ut + cown.ut) + mnext.ut;
    lw.uz = (pprev.uz + cown.uz) + mnext.uz;
    lw.bn = (pprev.bn + cown.bn) + mnext.bn;
    lw.bt = (pprev.bt + cown.bt) + mnext.bt;
    lw.bz = (pprev.bz + cown.bz) + mnext.bz;
    lw.p = (pprev.p + cown.p) + mnext.p;
    lw.psi = (pprev.psi + cown.psi) + mnext.psi;
    lw.b2 = 0.0;
    Cons out = cons_zero();
    to_prim_adj(u, w, lw, ph, out);
    if (interior) acc(out, g);
    return out;
}

// Adjoint of the interior cell (i, j) of a grid after its ghost fill: the cell's own
// adjoint (psi scaled by the damping factor that followed the fill, 1 when none) plus
// the adjoints of every ghost cell ghost_source() maps onto it, with the reflections'
// sign flips.  (Ghost cells copied from the previous grid through a Dirichlet side are
// transposed by dirichlet_gather.)
// lam(i, j) returns the adjoint at a cell in the rotated frame of the sweep.
template <bool XDIR, class Lam>
MHD_DI Cons gather_adj(const Lam& lam, const Geom& g, const int* bc, int i, int j, double psi_scale) {
    Cons v = lam(i, j);
    v.psi = v.psi * psi_scale;
    const int li = i - NG, lj = j - NG;
    if (li >= NG && li < g.nx - NG && lj >= NG && lj < g.ny - NG) return v;
    // Candidates (gx(a), gy(b)), a, b in 0..4: ghost columns x {j} and ghost rows,
    // {i} x ghost rows, and corners, where gx = {i, 0, 1, nx+2, nx+3} and gy likewise.
    // ghost_source maps the two axes independently, so a candidate's source is (i, j)
    // itself exactly when its column maps to i and its row to j, neither through a
    // Dirichlet side, with the flips of the two maps.  Ten axis maps give the set of
    // matching candidates as a 25-bit mask (bit 5a + b); its bits are taken in
    // increasing order, the order of a loop over a then b.
    unsigned mx = 0, my = 0, fx = 0, fy = 0;
#pragma unroll
    for (int a = 0; a < 5; ++a) {
        const int cx = a == 0 ? i : (a < 3 ? a - 1 : g.nx + a - 1);
        const int cy = a == 0 ? j : (a < 3 ? a - 1 : g.ny + a - 1);
        int s;
        bool f;
        if (map_axis(cx, g.nx, bc[0], bc[1], s, f) && s == i) mx |= 1u << a;
        if (f) fx |= 1u << a;
        if (map_axis(cy, g.ny, bc[2], bc[3], s, f) && s == j) my |= 1u << a;
        if (f) fy |= 1u << a;
    }
    unsigned m = 0;
#pragma unroll
    for (int a = 0; a < 5; ++a)
        if (mx & (1u << a)) m |= my << (5 * a);
    m &= ~1u;  // (i, j) itself
    while (m) {
        const int k = __builtin_ctz(m);
        m &= m - 1;
        const int a = k / 5, b = k % 5;
        const int ci = a == 0 ? i : (a < 3 ? a - 1 : g.nx + a - 1);
        const int cj = b == 0 ? j : (b < 3 ? b - 1 : g.ny + b - 1);
        Cons w = lam(ci, cj);
        // reference-frame flips: x negates rho*u and Bx, y negates rho*v and By
        const bool flx = (fx >> a) & 1u, fly = (fy >> b) & 1u;
        const bool fn = XDIR ? flx : fly;  // normal slots
        const bool ft = XDIR ? fly : flx;  // tangential slots
        if (fn) { w.mn = -w.mn; w.bn = -w.bn; }
        if (ft) { w.mt = -w.mt; w.bt = -w.bt; }
        acc(v, w);
    }
    return v;
}

// Adjoint of the ghost cells that a ghost fill copies from the previous grid through a
// Dirichlet side (ghost_source(...).prev: copy_ghost_cells followed by a side the fill
// leaves alone, head/cpu/boundary_cpu.hpp:12-32, solver_cpu.cpp:396-398,434-436).  Such a
// source is always a ghost cell of the previous grid.  The adjoint of the previous grid
// at its ghost cell (ti, tj) is v plus the adjoint of every ghost cell of the filled
// grid whose value was copied from there, with the reflections' sign flips; lam(i, j)
// returns the filled grid's adjoint in the reference frame (x frame).  The candidates
// are those of gather_adj, each once, in a fixed order.
template <class Lam>
MHD_DI Cons dirichlet_gather(const Lam& lam, const Geom& g, const int* bc, int ti, int tj, Cons v) {
    const int gx[5] = {ti, 0, 1, NG + g.nx, NG + g.nx + 1};
    const int gy[5] = {tj, 0, 1, NG + g.ny, NG + g.ny + 1};
    for (int a = 0; a < 5; ++a) {
        if (a > 0 && gx[a] == ti) continue;
        for (int b = 0; b < 5; ++b) {
            if (b > 0 && gy[b] == tj) continue;
            const Src s = ghost_source(gx[a], gy[b], g, bc);
            if (!s.prev || s.i != ti || s.j != tj) continue;
            Cons w = lam(gx[a], gy[b]);
            if (s.flip_x) { w.mn = -w.mn; w.bn = -w.bn; }
            if (s.flip_y) { w.mt = -w.mt; w.bt = -w.bt; }
            acc(v, w);
        }
    }
    return v;
}

MHD_DI bool any_dirichlet(const int* bc) { return bc[0] == 3 || bc[1] == 3 || bc[2] == 3 || bc[3] == 3; }

}  // namespace mhd
