// Discrete adjoint of one step on gfx950: the reverse y sweep, then the reverse x
// sweep, each one launch.  Memory layout as the forward (kernels.hip): nine SoA
// planes with ghost width 2, pitch nx+4, plane stride ps.
//
// A workgroup of WG lanes holds a tile of NL lines x LT cells along the sweep
// (x: one flattened run of interior rows, NL = 1; y: NL columns of LT rows, so
// that NL consecutive lanes read 8 NL contiguous bytes).  Each lane
// owns one cell and computes, with four LDS exchanges between neighbours along the
// line: the cell's primitive, its interface states, the adjoint of the face above
// it, the adjoint of its reconstruction, and finally its own output adjoint.  The
// output of a cell needs the cells within 3 along the line, so a tile writes its
// LT-6 central cells and tiles overlap by 6.  Nothing is accumulated across lanes
// with atomics: every output is a gather, so results are deterministic and equal,
// bit for bit, to the host build of the same arithmetic (hostcheck.cpp).
//
// Adjoint buffers: A holds the adjoint of the step's output (read by the y pass)
// and receives the adjoint of its input (written by the x pass); B holds the
// adjoint of the intermediate U*.  The x pass writes every column of the interior
// rows and zeroes the ghost rows of A; the y pass writes every row of the interior
// columns of B, whose ghost columns stay zero (B is cleared when allocated).
#include <hip/hip_runtime.h>

#include "adjoint_sweep.hpp"
#include "launch.hpp"

#pragma clang fp contract(off)

namespace mhd {

struct Bc4 {
    int s[4];
};

// The adjoint planes of a grid seen from a sweep (rotated frame).
template <bool XDIR, bool BUF> struct LamPlanes {
    std::conditional_t<BUF, BufPlanes<XDIR>, Planes<XDIR, const double>> p;
    int tnx;
    MHD_DI LamPlanes(const double* L, long long ps, int tnx_) : p(L, ps), tnx(tnx_) {}
    MHD_DI Cons operator()(int i, int j) const { return p.load(j * tnx + i); }
};

// Knob (compile time): MHD_ADJ_WAVES_PER_EU asks the compiler for that many resident
// waves per SIMD (0: its own choice).
#ifndef MHD_ADJ_WAVES_PER_EU
#define MHD_ADJ_WAVES_PER_EU 0
#endif
#if MHD_ADJ_WAVES_PER_EU > 0
#define MHD_ADJ_ATTR __attribute__((amdgpu_waves_per_eu(MHD_ADJ_WAVES_PER_EU)))
#else
#define MHD_ADJ_ATTR
#endif

// A lane's record of its reconstruction (ReconRecord), kept from step 2 to step 4 so
// that the reconstruction adjoint does not form the reconstruction a second time.  Only
// the lane that wrote a slot reads it: no barrier involved.
template <int WG> struct RecLds {
    double s[9][WG];
    Lds9<WG> ul, ur;
    unsigned char flags[WG];
    MHD_DI void put(int t, const ReconRecord& r) {
#pragma unroll
        for (int k = 0; k < 9; ++k) s[k][t] = r.s[k];
        ul.put(t, r.ul);
        ur.put(t, r.ur);
        flags[t] = (unsigned char)((r.pl ? 1 : 0) | (r.pr ? 2 : 0) | (r.pcl ? 4 : 0) | (r.pcr ? 8 : 0));
    }
    MHD_DI ReconRecord get(int t) const {
        ReconRecord r;
#pragma unroll
        for (int k = 0; k < 9; ++k) r.s[k] = s[k][t];
        r.ul = ul.get(t);
        r.ur = ur.get(t);
        const unsigned f = flags[t];
        r.pl = f & 1;
        r.pr = f & 2;
        r.pcl = f & 4;
        r.pcr = f & 8;
        return r;
    }
};

template <int SOLVER, int LIMITER, bool XDIR, int NL, int WG, bool BUF>
__global__ __launch_bounds__(WG) MHD_ADJ_ATTR void adj_sweep_kernel(const double* __restrict__ U, const double* __restrict__ Lin,
                                                       double* __restrict__ Lout, Geom g, Phys ph, double dtd,
                                                       double hdtd, double psi_scale, Bc4 bc, int ntiles) {
    constexpr int LT = WG / NL, NOUT = LT - 6;
    __shared__ Lds9<WG> sW;      // cell primitives
    __shared__ Lds9<WG> sA;      // lower interface state (cons), then the bar of W[c-1]
    __shared__ Lds9<WG, 10> sB;  // its primitive, then the bar of W[c+1]
    __shared__ Lds9<WG> sG;      // output adjoint g, then the bar of the lower interface state
    __shared__ RecLds<WG> sR;    // each lane's own forward record of its reconstruction
    const int t = threadIdx.x;

    if (XDIR && (int)blockIdx.x >= ntiles) {
        // zero the ghost rows (rows 0, 1, ny+2, ny+3) of the input adjoint
        const int q = ((int)blockIdx.x - ntiles) * WG + t;
        if (q < 2 * NG * g.tnx) {
            const int r = q / g.tnx;
            const int row = r < NG ? r : g.ny + r;
            const long long k = (long long)row * g.tnx + q % g.tnx;
#pragma unroll
            for (int f = 0; f < 9; ++f) Lout[f * g.ps + k] = 0.0;
        }
        return;
    }

    const int line = t % NL, pos = t / NL;
    int i, j;           // cell (clamped into the grid for loads)
    int a, n;           // index along the sweep and the interior count along it
    bool inb, lineok;   // cell exists; its row (x) / column (y) is interior
    if (XDIR) {
        const long long p = (long long)NG * g.tnx + (long long)blockIdx.x * NOUT - 3 + pos;
        inb = p >= 0 && p < (long long)g.tnx * g.tny;
        const long long pc = inb ? p : 0;
        j = (int)(pc / g.tnx);
        i = (int)(pc % g.tnx);
        a = i;
        n = g.nx;
        lineok = inb && j >= NG && j < NG + g.ny;
    } else {
        const int ii = NG + (int)blockIdx.x * NL + line;
        const int jj = (int)blockIdx.y * NOUT - 3 + pos;
        inb = ii < NG + g.nx && jj >= 0 && jj < g.tny;
        i = ii < NG + g.nx ? ii : NG + g.nx - 1;
        j = jj < 0 ? 0 : (jj >= g.tny ? g.tny - 1 : jj);
        a = jj;
        n = g.ny;
        lineok = ii < NG + g.nx;
    }
    const AdjFlags fl = adj_flags(a, n, lineok);
    const int kc = j * g.tnx + i;
    const std::conditional_t<BUF, BufPlanes<XDIR>, Planes<XDIR, const double>> src(U, g.ps);
    const LamPlanes<XDIR, BUF> lin(Lin, g.ps, g.tnx);
    const int tm = pos > 0 ? t - NL : t;
    const int tp = pos < LT - 1 ? t + NL : t;

    // 1. primitives
    const Cons u = src.load(kc);
    const Prim w = to_prim(u, ph);
    sW.put(t, w);
    lds_barrier();

    // 2. interface states and the output adjoint of the cell
    Face lo, hi;
    if (fl.rec) {
        ReconRecord rec;
        reconstruct<LIMITER>(sW.getp(tm), w, sW.getp(tp), hdtd, ph, lo, hi, &rec);
        sR.put(t, rec);
    } else {
        lo.u = u; lo.w = w; hi = lo;
    }
    const Cons gown = fl.interior ? gather_adj<XDIR>(lin, g, bc.s, i, j, psi_scale) : cons_zero();
    sA.put(t, lo.u);
    sB.put(t, lo.w);
    sG.put(t, gown);
    lds_barrier();

    // 3. face (a | a+1)
    Cons hib = cons_zero(), lonb = cons_zero();
    if (fl.face) {
        const Face lon{sA.get(tp), sB.getp(tp)};
        adj_face<SOLVER, XDIR>(hi, lon, gown, sG.get(tp), dtd, ph, hib, lonb);
    }
    lds_barrier();
    sG.put(t, lonb);
    lds_barrier();

    // 4. reconstruction of the cell
    Prim wmb = prim_zero(), wcb = prim_zero(), wpb = prim_zero();
    if (fl.rec) {
        const Cons lob = pos > 0 ? sG.get(tm) : cons_zero();
        reconstruct_adj_rec<LIMITER>(sW.getp(tm), w, sW.getp(tp), hdtd, ph, sR.get(t), lob, hib, wmb, wcb, wpb);
    }
    sA.put(t, Cons{wmb.r, wmb.un, wmb.ut, wmb.uz, wmb.bn, wmb.bt, wmb.bz, wmb.p, wmb.psi});
    sB.put(t, wpb);
    lds_barrier();

    // 5. output
    const Prim zero = prim_zero();
    const Prim pprev = pos > 0 ? sB.getp(tm) : zero;
    const Cons mc = pos < LT - 1 ? sA.get(tp) : cons_zero();
    const Prim mnext{mc.r, mc.mn, mc.mt, mc.mz, mc.bn, mc.bt, mc.bz, mc.e, mc.psi, 0.0};
    const Cons out = adj_cell(u, w, pprev, wcb, mnext, fl.interior, gown, ph);
    if (inb && lineok && pos >= 3 && pos < LT - 3)
        std::conditional_t<BUF, BufPlanes<XDIR>, Planes<XDIR, double>>(Lout, g.ps).store(kc, out);
}

// Dirichlet transpose (dirichlet_gather) over every ghost cell of the previous grid's
// adjoint LP, from the filled grid's adjoint LG.  keep_all: every target keeps the value
// the preceding reverse pass wrote (after the x pass); otherwise only the ghost rows of
// the interior columns do (after the y pass, which writes no other ghost cell of U*'s
// adjoint) and the rest are set, so nothing of an earlier step survives there.
__global__ __launch_bounds__(256) void adj_dirichlet_kernel(const double* __restrict__ LG, double* __restrict__ LP,
                                                            Geom g, Bc4 bc, int keep_all) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ghost_count(g)) return;
    int ti, tj;
    ghost_cell(g, t, ti, tj);
    const LamPlanes<true, false> lg(LG, g.ps, g.tnx);
    const Planes<true, double> lp(LP, g.ps);
    const int k = tj * g.tnx + ti;
    const bool keep = keep_all || (ti >= NG && ti < NG + g.nx);
    lp.store(k, dirichlet_gather(lg, g, bc.s, ti, tj, keep ? lp.load(k) : cons_zero()));
}

// Tile shapes (compile-time knobs): workgroup width and, for the y pass, columns
// per tile.  At one wave per SIMD (the kernels hold ~410 registers) a CU runs
// whatever workgroups fit, and small workgroups keep a barrier from idling the
// whole CU: x one wave of 64 cells (58 written), y 4 columns x 32 rows (26 written).
#ifndef MHD_ADJ_XWG
#define MHD_ADJ_XWG 64
#endif
#ifndef MHD_ADJ_YWG
#define MHD_ADJ_YWG 128
#endif
#ifndef MHD_ADJ_YNL
#define MHD_ADJ_YNL 4
#endif

template <int S, int L, bool BUF>
static hipError_t launch_adjoint_sla(const AdjArgs& a) {
    const Geom g{a.nx, a.ny, a.nx + 2 * NG, a.ny + 2 * NG, a.ps};
    const Phys ph{a.gamma, a.gamma - 1.0, a.ch, a.ch * a.ch, 0.5 * a.ch, 0.5 / a.ch};
    const double dtdx = a.dt / a.dx, dtdy = a.dt / a.dy;
    const bool damping = !(a.damp >= 1.0);  // solver_cpu.cpp:265
    Bc4 bc{{a.bc[0], a.bc[1], a.bc[2], a.bc[3]}};
    constexpr int XWG = MHD_ADJ_XWG, YWG = MHD_ADJ_YWG, NLY = MHD_ADJ_YNL;
    constexpr int NOUTY = YWG / NLY - 6, NOUTX = XWG - 6;
    dim3 gy((unsigned)((g.nx + NLY - 1) / NLY), (unsigned)((g.tny + NOUTY - 1) / NOUTY));
    hipLaunchKernelGGL((adj_sweep_kernel<S, L, false, NLY, YWG, BUF>), gy, dim3(YWG), 0, a.stream, a.Utmp, a.lam,
                       a.lam_tmp, g, ph, dtdy, 0.5 * dtdy, damping ? a.damp : 1.0, bc, 0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const bool dirichlet = any_dirichlet(bc.s);
    const int nghost = ghost_count(g);
    if (dirichlet) {  // U_next's ghosts copied from U* -> U*'s adjoint
        hipLaunchKernelGGL(adj_dirichlet_kernel, dim3((nghost + 255) / 256), dim3(256), 0, a.stream, a.lam,
                           a.lam_tmp, g, bc, 0);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    const long long flat = (long long)g.tnx * g.ny;
    const int ntiles = (int)((flat + NOUTX - 1) / NOUTX);
    const int nzero = (2 * NG * g.tnx + XWG - 1) / XWG;
    hipLaunchKernelGGL((adj_sweep_kernel<S, L, true, 1, XWG, BUF>), dim3((unsigned)(ntiles + nzero)), dim3(XWG), 0,
                       a.stream, a.Uold, a.lam_tmp, a.lam, g, ph, dtdx, 0.5 * dtdx, 1.0, bc, ntiles);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (dirichlet) {  // U*'s ghosts copied from U -> U's adjoint
        hipLaunchKernelGGL(adj_dirichlet_kernel, dim3((nghost + 255) / 256), dim3(256), 0, a.stream, a.lam_tmp,
                           a.lam, g, bc, 1);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

// Buffer-resource addressing for grids of at most 4 GiB, as the forward sweeps.
template <int S, int L>
static hipError_t launch_adjoint_sl(const AdjArgs& a) {
    if (a.addr_buf && 9ull * (unsigned long long)a.ps * 8 <= BUF_MAX_BYTES) return launch_adjoint_sla<S, L, true>(a);
    return launch_adjoint_sla<S, L, false>(a);
}

hipError_t launch_adjoint(const AdjArgs& a) {
    switch (a.solver * 2 + a.limiter) {
        case 0: return launch_adjoint_sl<0, 0>(a);
        case 1: return launch_adjoint_sl<0, 1>(a);
        case 2: return launch_adjoint_sl<1, 0>(a);
        case 3: return launch_adjoint_sl<1, 1>(a);
        case 4: return launch_adjoint_sl<2, 0>(a);
        case 5: return launch_adjoint_sl<2, 1>(a);
        case 6: return launch_adjoint_sl<3, 0>(a);
        case 7: return launch_adjoint_sl<3, 1>(a);
    }
    return hipErrorInvalidValue;
}

}  // namespace mhd

namespace mhd {

__global__ void count_nonfinite_kernel(const double* __restrict__ U, long long n, long long ps,
                                       unsigned int* __restrict__ count) {
    unsigned int bad = 0;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
#pragma unroll
        for (int f = 0; f < 9; ++f) bad += finite(U[f * ps + k]) ? 0u : 1u;
    for (int off = 32; off > 0; off >>= 1) bad += __shfl_xor(bad, off);
    atomicAdd(count, (threadIdx.x & 63) == 0 ? bad : 0u);
}

hipError_t launch_count_nonfinite(const double* U, long long n, long long ps, unsigned int* count, hipStream_t s) {
    long long nb = (n + 255) / 256;
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(count_nonfinite_kernel, dim3((unsigned)nb), dim3(256), 0, s, U, n, ps, count);
    return hipGetLastError();
}

}  // namespace mhd
