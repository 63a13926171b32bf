// Host build of the adjoint arithmetic (adjoint.hpp, adjoint_sweep.hpp) for the
// CPU test suite: the point adjoints and one reverse step over host grids, with
// the same functions, flags and summation order as the kernels of adjoint.hip, so
// that the device reverse step can be compared with it bit for bit and this one,
// on machines without a GPU, with the oracle's tangent-linear model by
// dot-product tests.  Not part of the product library (libmhd_hostcheck.so).
//
// Arrays are in the reference field order (rho, rho u, rho v, rho w, Bx, By, Bz,
// E, psi); grids are AoS row-major with ghost width 2, as head/cpu/grid_cpu.hpp.
#include <cstring>
#include <vector>

#include "adjoint_sweep.hpp"

using namespace mhd;

namespace {

Cons rot(const double* a, bool xdir) {
    return xdir ? Cons{a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8]}
                : Cons{a[0], a[2], a[1], a[3], a[5], a[4], a[6], a[7], a[8]};
}
void unrot(const Cons& c, bool xdir, double* a) {
    a[0] = c.r; a[3] = c.mz; a[6] = c.bz; a[7] = c.e; a[8] = c.psi;
    if (xdir) { a[1] = c.mn; a[2] = c.mt; a[4] = c.bn; a[5] = c.bt; }
    else { a[2] = c.mn; a[1] = c.mt; a[5] = c.bn; a[4] = c.bt; }
}
// Primitive arrays (rho, u, v, w, Bx, By, Bz, p, psi) in the rotated frame.
Prim rotp(const double* a, bool xdir) {
    const Cons c = rot(a, xdir);
    return Prim{c.r, c.mn, c.mt, c.mz, c.bn, c.bt, c.bz, c.e, c.psi, 0.0};
}
void unrotp(const Prim& p, bool xdir, double* a) {
    unrot(Cons{p.r, p.un, p.ut, p.uz, p.bn, p.bt, p.bz, p.p, p.psi}, xdir, a);
}
Phys phys(double gamma, double ch) { return Phys{gamma, gamma - 1.0, ch, ch * ch, 0.5 * ch, 0.5 / ch}; }

struct HostLam {
    const double* L;
    int tnx;
    bool xdir;
    Cons operator()(int i, int j) const { return rot(L + 9 * ((size_t)j * tnx + i), xdir); }
};

// One sweep's reverse over host grids: Lin = adjoint of the sweep output (full
// grid), Lout receives the adjoint of the sweep input on the cells the device pass
// writes (x: every column of the interior rows; y: every row of the interior columns).
template <int SOLVER, int LIMITER, bool XDIR>
void reverse_sweep(const double* U, const double* Lin, double* Lout, const Geom& g, const Phys& ph, double dtd,
                   double hdtd, double psi_scale, const int* bc) {
    const int nlines = XDIR ? g.ny : g.nx;
    const int len = XDIR ? g.tnx : g.tny;
    const int n = XDIR ? g.nx : g.ny;
    const HostLam lin{Lin, g.tnx, XDIR};
#pragma omp parallel if ((long long)nlines * len > 65536)
    {
    std::vector<Cons> u(len), gv(len), hib(len), lob(len);
    std::vector<Prim> w(len), mb(len), cb(len), pb(len);
    std::vector<Face> lo(len), hi(len);
    std::vector<AdjFlags> fl(len);
#pragma omp for schedule(static)
    for (int l = 0; l < nlines; ++l) {
        auto cell = [&](int a, int& i, int& j) {
            i = XDIR ? a : NG + l;
            j = XDIR ? NG + l : a;
        };
        for (int a = 0; a < len; ++a) {
            int i, j;
            cell(a, i, j);
            u[a] = rot(U + 9 * ((size_t)j * g.tnx + i), XDIR);
            w[a] = to_prim(u[a], ph);
            fl[a] = adj_flags(a, n, true);
        }
        for (int a = 0; a < len; ++a) {
            if (fl[a].rec) reconstruct<LIMITER>(w[a - 1], w[a], w[a + 1], hdtd, ph, lo[a], hi[a]);
            else { lo[a].u = u[a]; lo[a].w = w[a]; hi[a] = lo[a]; }
            int i, j;
            cell(a, i, j);
            gv[a] = fl[a].interior ? gather_adj<XDIR>(lin, g, bc, i, j, psi_scale) : cons_zero();
        }
        for (int a = 0; a < len; ++a) { hib[a] = cons_zero(); lob[a] = cons_zero(); }
        for (int a = 0; a < len; ++a)
            if (fl[a].face) adj_face<SOLVER, XDIR>(hi[a], lo[a + 1], gv[a], gv[a + 1], dtd, ph, hib[a], lob[a + 1]);
        for (int a = 0; a < len; ++a) {
            mb[a] = prim_zero(); cb[a] = prim_zero(); pb[a] = prim_zero();
            if (fl[a].rec)
                reconstruct_adj<LIMITER>(w[a - 1], w[a], w[a + 1], hdtd, ph, lob[a], hib[a], mb[a], cb[a], pb[a]);
        }
        for (int a = 0; a < len; ++a) {
            const Prim z = prim_zero();
            const Cons out = adj_cell(u[a], w[a], a > 0 ? pb[a - 1] : z, cb[a], a < len - 1 ? mb[a + 1] : z,
                                      fl[a].interior, gv[a], ph);
            int i, j;
            cell(a, i, j);
            unrot(out, XDIR, Lout + 9 * ((size_t)j * g.tnx + i));
        }
    }
    }
}

// The Dirichlet transpose over every ghost cell of the previous grid's adjoint LP from
// the filled grid's adjoint LG (dirichlet_gather), as adj_dirichlet_kernel: the cells
// the preceding reverse pass wrote keep their value and gain the copies' adjoints
// (keep_all: every ghost cell; otherwise the ghost rows of the interior columns), the
// others are set to the copies' adjoints alone.
void dirichlet_pass(const double* LG, double* LP, const Geom& g, const int* bc, bool keep_all) {
    const HostLam lg{LG, g.tnx, true};
    for (int t = 0; t < ghost_count(g); ++t) {
        int ti, tj;
        ghost_cell(g, t, ti, tj);
        double* p = LP + 9 * ((size_t)tj * g.tnx + ti);
        const bool keep = keep_all || (ti >= NG && ti < NG + g.nx);
        const Cons v = dirichlet_gather(lg, g, bc, ti, tj, keep ? rot(p, true) : cons_zero());
        unrot(v, true, p);
    }
}

template <int S, int L>
void reverse_step(const double* Uold, const double* Utmp, double* lam, const Geom& g, double dx, double dy, double dt,
                  double ch, double cr, double gamma, const int* bc) {
    const Phys ph = phys(gamma, ch);
    const double dtdx = dt / dx, dtdy = dt / dy;
    const double damp = std::exp(-dt * ch / cr);
    const bool damping = !(damp >= 1.0);
    std::vector<double> B(9 * (size_t)g.tnx * g.tny, 0.0);
    reverse_sweep<S, L, false>(Utmp, lam, B.data(), g, ph, dtdy, 0.5 * dtdy, damping ? damp : 1.0, bc);
    if (any_dirichlet(bc)) dirichlet_pass(lam, B.data(), g, bc, false);
    reverse_sweep<S, L, true>(Uold, B.data(), lam, g, ph, dtdx, 0.5 * dtdx, 1.0, bc);
    for (int j = 0; j < g.tny; ++j) {
        if (j >= NG && j < NG + g.ny) continue;
        std::memset(lam + 9 * (size_t)j * g.tnx, 0, 9 * (size_t)g.tnx * sizeof(double));
    }
    if (any_dirichlet(bc)) dirichlet_pass(B.data(), lam, g, bc, true);
}

}  // namespace

extern "C" {

// F = riemann(UL, UR): bars of UL and UR (conserved, to_prim folded in) from Fb.
void hc_riemann_adj(const double* UL, const double* UR, int dir, int solver, double ch, double gamma,
                    const double* Fb, double* ULb, double* URb) {
    const bool x = dir == 0;
    const Phys ph = phys(gamma, ch);
    const Cons ul = rot(UL, x), ur = rot(UR, x);
    const Face l{ul, to_prim(ul, ph)}, r{ur, to_prim(ur, ph)};
    // adj_face takes output adjoints g; with gown = 0, gnext = Fb / 1 and dtd = 1
    // the face bar is Fb itself.
    const Cons fb = rot(Fb, x), z = cons_zero();
    Cons lb, rb;
    switch (solver) {
#define MHD_HC_FACE(S)                                        \
    case S:                                                   \
        if (x) adj_face<S, true>(l, r, z, fb, 1.0, ph, lb, rb); \
        else adj_face<S, false>(l, r, z, fb, 1.0, ph, lb, rb);  \
        break;
        MHD_HC_FACE(0) MHD_HC_FACE(1) MHD_HC_FACE(2) MHD_HC_FACE(3)
#undef MHD_HC_FACE
    }
    unrot(lb, x, ULb);
    unrot(rb, x, URb);
}

// (UL*, UR*) = reconstruct(Wm, Wc, Wp) with half-step factor 0.5 dtd.
void hc_reconstruct_adj(const double* Wm, const double* Wc, const double* Wp, double dtd, int dir, double ch,
                        double gamma, int limiter, const double* Lb, const double* Rb, double* Wmb, double* Wcb,
                        double* Wpb) {
    const bool x = dir == 0;
    const Phys ph = phys(gamma, ch);
    Prim mb = prim_zero(), cb = prim_zero(), pb = prim_zero();
    if (limiter == 0)
        reconstruct_adj<0>(rotp(Wm, x), rotp(Wc, x), rotp(Wp, x), 0.5 * dtd, ph, rot(Lb, x), rot(Rb, x), mb, cb, pb);
    else
        reconstruct_adj<1>(rotp(Wm, x), rotp(Wc, x), rotp(Wp, x), 0.5 * dtd, ph, rot(Lb, x), rot(Rb, x), mb, cb, pb);
    unrotp(mb, x, Wmb);
    unrotp(cb, x, Wcb);
    unrotp(pb, x, Wpb);
}

// F = flux(U) (primitives from U): bar of U.
void hc_flux_adj(const double* U, int dir, double ch, double gamma, const double* Fb, double* Ub) {
    const bool x = dir == 0;
    const Phys ph = phys(gamma, ch);
    const Cons u = rot(U, x);
    const Prim w = to_prim(u, ph);
    Cons ub = cons_zero();
    Prim wb = prim_zero();
    flux_adj(u, w, rot(Fb, x), ph, ub, wb);
    to_prim_adj(u, w, wb, ph, ub);
    unrot(ub, x, Ub);
}

// V = cons_to_prim(U): bar of U.
void hc_cons_to_prim_adj(const double* U, double gamma, const double* Vb, double* Ub) {
    const Phys ph = phys(gamma, 1.0);
    const Cons u = rot(U, true);
    Cons ub = cons_zero();
    to_prim_adj(u, to_prim(u, ph), rotp(Vb, true), ph, ub);
    unrot(ub, true, Ub);
}

// U = prim_to_cons(V): bar of V.
void hc_prim_to_cons_adj(const double* V, double gamma, const double* Ub, double* Vb) {
    const Phys ph = phys(gamma, 1.0);
    Prim vb = prim_zero();
    to_cons_adj(rotp(V, true), rot(Ub, true), ph, vb);
    unrotp(vb, true, Vb);
}

// c = fast_speed(V) along dir (|B|^2 from V): bar of V.
void hc_fast_speed_adj(const double* V, int dir, double gamma, double cb, double* Vb) {
    const bool x = dir == 0;
    const Phys ph = phys(gamma, 1.0);
    Prim v = rotp(V, x);
    v.b2 = v.bn * v.bn + v.bt * v.bt + v.bz * v.bz;
    Prim vb = prim_zero();
    fast_speed_adj(v, cb, ph, vb);
    vb.bn += 2.0 * v.bn * vb.b2;
    vb.bt += 2.0 * v.bt * vb.b2;
    vb.bz += 2.0 * v.bz * vb.b2;
    unrotp(vb, x, Vb);
}

void hc_limiter_adj(double a, double b, int limiter, double sb, double* ab, double* bb) {
    *ab = 0.0;
    *bb = 0.0;
    if (limiter == 0) limit_adj<0>(a, b, sb, *ab, *bb);
    else limit_adj<1>(a, b, sb, *ab, *bb);
}

// One reverse step: lam = dL/d(step output) in, dL/dUold out (AoS, full grid).
// Returns 0, or 1 for an unsupported solver / limiter / boundary.
int hc_adjoint_step(const double* Uold, const double* Utmp, int nx, int ny, double x0, double x1, double y0,
                    double y1, double dt, double ch, double cr, double gamma, int solver, int limiter, const int* bc,
                    double* lam) {
    for (int k = 0; k < 4; ++k)
        if (bc[k] < 0 || bc[k] > 3) return 1;
    const Geom g{nx, ny, nx + 2 * NG, ny + 2 * NG, 0};
    const double dx = (x1 - x0) / static_cast<double>(nx), dy = (y1 - y0) / static_cast<double>(ny);
    switch (solver * 2 + limiter) {
        case 0: reverse_step<0, 0>(Uold, Utmp, lam, g, dx, dy, dt, ch, cr, gamma, bc); return 0;
        case 1: reverse_step<0, 1>(Uold, Utmp, lam, g, dx, dy, dt, ch, cr, gamma, bc); return 0;
        case 2: reverse_step<1, 0>(Uold, Utmp, lam, g, dx, dy, dt, ch, cr, gamma, bc); return 0;
        case 3: reverse_step<1, 1>(Uold, Utmp, lam, g, dx, dy, dt, ch, cr, gamma, bc); return 0;
        case 4: reverse_step<2, 0>(Uold, Utmp, lam, g, dx, dy, dt, ch, cr, gamma, bc); return 0;
        case 5: reverse_step<2, 1>(Uold, Utmp, lam, g, dx, dy, dt, ch, cr, gamma, bc); return 0;
        case 6: reverse_step<3, 0>(Uold, Utmp, lam, g, dx, dy, dt, ch, cr, gamma, bc); return 0;
        case 7: reverse_step<3, 1>(Uold, Utmp, lam, g, dx, dy, dt, ch, cr, gamma, bc); return 0;
    }
    return 1;
}

// compute_dt candidates of n cells (conserved, reference order, 9 per cell): the
// kernels' cfl_candidate, which returns c_h for cells its speed test clears, and
// the exact evaluation of every cell.  hit[i] = 1 where the test cleared cell i.
void hc_cfl_candidates(const double* U, int n, double gamma, double ch, double* fast, double* exact, int* hit) {
    const Phys ph = phys(gamma, ch);
    for (int i = 0; i < n; ++i) {
        const Cons u = rot(U + 9 * (size_t)i, true);
        fast[i] = cfl_candidate(u, ph);
        const Prim w = to_prim(u, ph);
        const bool valid = finite(w.r) && finite(w.p) && w.r > 0.0 && w.p > 0.0;
        exact[i] = valid ? cfl_candidate_exact(w, ph) : 0.0;
        hit[i] = valid && speeds_below_ch(w, ph);
    }
}

// The ghost rows the x sweep writes itself (row_push): out = {n, src[4], dst[4], flip}.
void hc_row_push(int ny, int bc_lo, int bc_hi, int* out) {
    const RowPush rp = row_push(ny, bc_lo, bc_hi);
    out[0] = rp.n;
    for (int e = 0; e < 4; ++e) {
        out[1 + e] = rp.src[e];
        out[5 + e] = rp.dst[e];
    }
    out[9] = (int)rp.flip;
}

// The closed-form source of ghost cell (gi, gj) (ghost_source): out = {i, j, prev, flip_x, flip_y}.
void hc_ghost_source(int gi, int gj, int nx, int ny, const int* bc, int* out) {
    const Geom g{nx, ny, nx + 2 * NG, ny + 2 * NG, (long long)(nx + 2 * NG) * (ny + 2 * NG)};
    const Src s = ghost_source(gi, gj, g, bc);
    out[0] = s.i;
    out[1] = s.j;
    out[2] = s.prev;
    out[3] = s.flip_x;
    out[4] = s.flip_y;
}

}  // extern "C"
