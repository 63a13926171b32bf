// gfx950 kernels of the GLM-MHD MUSCL-Hancock step (one call of
// advance_second_order, src/cpu/solver_cpu.cpp:327-440, plus compute_dt :283-312).
//
// Memory: 9 SoA float64 planes per grid (reference field order), row pitch
// tnx = nx + 4, ghost width 2, plane stride `ps` doubles.
//
//   xsweep  : Uold -> Utmp interior.  A workgroup walks the flattened row-major
//             sequence of interior rows; the two ghost columns of each row are
//             the stencil halo of that row's edge cells, so one tiling serves any
//             nx.  Neighbour values go through LDS.
//   bc      : every ghost cell of a grid from its closed-form source (see
//             ghost_source), in one launch.
//   ysweep  : Utmp -> Unew interior.  One lane per column marches down a segment
//             of rows with a register pipeline (prims of two rows, the upper
//             interface state and the lower face flux carried).  Fuses the psi
//             damping (solver_cpu.cpp:263-275) and the next step's CFL candidate.
//   cfl     : stand-alone compute_dt reduction for a state not produced by ysweep.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "physics.hpp"

#pragma clang fp contract(off)

namespace mhd {

constexpr int NG = 2;

struct Geom {
    int nx, ny, tnx, tny;
    long long ps;  // plane stride (doubles)
};

// Field indices in the reference order.
enum : int { F_RHO = 0, F_MX = 1, F_MY = 2, F_MZ = 3, F_BX = 4, F_BY = 5, F_BZ = 6, F_E = 7, F_PSI = 8 };

template <bool XDIR>
MHD_DI Cons load_rot(const double* __restrict__ U, long long ps, long long k) {
    Cons u;
    u.r = U[k];
    u.mn = U[(XDIR ? F_MX : F_MY) * ps + k];
    u.mt = U[(XDIR ? F_MY : F_MX) * ps + k];
    u.mz = U[F_MZ * ps + k];
    u.bn = U[(XDIR ? F_BX : F_BY) * ps + k];
    u.bt = U[(XDIR ? F_BY : F_BX) * ps + k];
    u.bz = U[F_BZ * ps + k];
    u.e = U[F_E * ps + k];
    u.psi = U[F_PSI * ps + k];
    return u;
}

template <bool XDIR>
MHD_DI void store_rot(double* __restrict__ U, long long ps, long long k, const Cons& u) {
    U[k] = u.r;
    U[(XDIR ? F_MX : F_MY) * ps + k] = u.mn;
    U[(XDIR ? F_MY : F_MX) * ps + k] = u.mt;
    U[F_MZ * ps + k] = u.mz;
    U[(XDIR ? F_BX : F_BY) * ps + k] = u.bn;
    U[(XDIR ? F_BY : F_BX) * ps + k] = u.bt;
    U[F_BZ * ps + k] = u.bz;
    U[F_E * ps + k] = u.e;
    U[F_PSI * ps + k] = u.psi;
}

// LDS staging of a 9-double record, field-major so that lane-consecutive
// accesses hit consecutive banks.
template <int WG> struct Lds9 {
    double v[9][WG];
    MHD_DI void put(int t, const Cons& u) {
        v[0][t] = u.r; v[1][t] = u.mn; v[2][t] = u.mt; v[3][t] = u.mz; v[4][t] = u.bn;
        v[5][t] = u.bt; v[6][t] = u.bz; v[7][t] = u.e; v[8][t] = u.psi;
    }
    MHD_DI Cons get(int t) const {
        return Cons{v[0][t], v[1][t], v[2][t], v[3][t], v[4][t], v[5][t], v[6][t], v[7][t], v[8][t]};
    }
    MHD_DI void put(int t, const Prim& w) {
        v[0][t] = w.r; v[1][t] = w.un; v[2][t] = w.ut; v[3][t] = w.uz; v[4][t] = w.bn;
        v[5][t] = w.bt; v[6][t] = w.bz; v[7][t] = w.p; v[8][t] = w.psi;
    }
    MHD_DI Prim getp(int t) const {
        return Prim{v[0][t], v[1][t], v[2][t], v[3][t], v[4][t], v[5][t], v[6][t], v[7][t], v[8][t]};
    }
};

// ---------------------------------------------------------------------------------------
// x sweep
// ---------------------------------------------------------------------------------------
// Block b covers flattened cells [base + b*(WG-4), +WG) of the interior rows
// (base = jb*tnx - 2); lanes 2..WG-3 that land on an interior column write Utmp.
template <int SOLVER, int LIMITER, int WG>
__global__ __launch_bounds__(WG) void xsweep_kernel(const double* __restrict__ Uo, double* __restrict__ Ut, Geom g,
                                                    Phys ph, double dtd, double hdtd,
                                                    unsigned long long* __restrict__ acc) {
    __shared__ Lds9<WG> sw;   // primitives, later face fluxes
    __shared__ Lds9<WG> su;   // lower interface states
    __shared__ Lds9<WG> sv;   // their primitives
    const int t = threadIdx.x;
    if (acc && blockIdx.x == 0 && t == 0) *acc = 0ull;  // CFL accumulator for this step's ysweep
    const long long total = (long long)g.tnx * g.tny;
    const long long k0 = (long long)NG * g.tnx - 2 + (long long)blockIdx.x * (WG - 4);
    long long k = k0 + t;
    const long long kl = k < total ? k : total - 1;  // clamp: tail lanes read valid memory
    const Cons u = load_rot<true>(Uo, g.ps, kl);
    const Prim w = to_prim(u, ph);
    sw.put(t, w);
    __syncthreads();
    Face lo, hi;
    {
        const int tm = t > 0 ? t - 1 : t;
        const int tp = t < WG - 1 ? t + 1 : t;
        reconstruct<LIMITER>(sw.getp(tm), w, sw.getp(tp), hdtd, ph, lo, hi);
    }
    su.put(t, lo.u);
    sv.put(t, lo.w);
    __syncthreads();
    // Face between this cell and the next one.
    const int tp = t < WG - 1 ? t + 1 : t;
    const Cons f = riemann<SOLVER, true>(hi.u, hi.w, su.get(tp), sv.getp(tp), ph);
    sw.put(t, f);
    __syncthreads();
    if (t < 2 || t >= WG - 2 || k >= total) return;
    const int i = (int)(k % g.tnx);
    const int j = (int)(k / g.tnx);
    if (i < NG || i >= NG + g.nx || j >= NG + g.ny) return;
    const Cons n = update(u, dtd, f, sw.get(t - 1));
    store_rot<true>(Ut, g.ps, k, n);
}

// ---------------------------------------------------------------------------------------
// y sweep (column march)
// ---------------------------------------------------------------------------------------
// Undamped psi of the interior cells within NG of an edge, for the ghost fill
// that follows (the reference damps the interior after its BC, so ghosts carry
// undamped psi, solver_cpu.cpp:435-439).  Layout: [0, 2NG*nx) bottom/top strips
// (row r: r < NG -> jb + r, else je - 2NG + r), then [2NG*nx, +2NG*ny) left/right
// strips (column c: c < NG -> ib + c, else ie - 2NG + c).
MHD_DI void frame_put(double* __restrict__ fr, const Geom& g, int i, int j, double psi) {
    const int li = i - NG, lj = j - NG;
    if (lj < NG) fr[lj * g.nx + li] = psi;
    if (lj >= g.ny - NG) fr[(lj - (g.ny - 2 * NG)) * g.nx + li] = psi;
    double* side = fr + 2 * NG * g.nx;
    if (li < NG) side[li * g.ny + lj] = psi;
    if (li >= g.nx - NG) side[(li - (g.nx - 2 * NG)) * g.ny + lj] = psi;
}
MHD_DI double frame_get(const double* __restrict__ fr, const Geom& g, int i, int j) {
    const int li = i - NG, lj = j - NG;
    if (lj < NG) return fr[lj * g.nx + li];
    if (lj >= g.ny - NG) return fr[(lj - (g.ny - 2 * NG)) * g.nx + li];
    const double* side = fr + 2 * NG * g.nx;
    if (li < NG) return side[li * g.ny + lj];
    return side[(li - (g.nx - 2 * NG)) * g.ny + lj];
}

MHD_DI void wave_max_to(unsigned long long* acc, double m) {
    for (int off = 32; off > 0; off >>= 1) m = std_max(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(acc, (unsigned long long)__double_as_longlong(m));
}

template <int SOLVER, int LIMITER>
__global__ __launch_bounds__(256) void ysweep_kernel(const double* __restrict__ Ut, double* __restrict__ Un, Geom g,
                                                     Phys ph, double dtd, double hdtd, double damp, int seg,
                                                     double* __restrict__ frame,
                                                     unsigned long long* __restrict__ acc) {
    const int i = NG + blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i < NG + g.nx;
    const int ic = active ? i : NG;  // idle lanes shadow column NG (no stores)
    const int j0 = NG + blockIdx.y * seg;
    const int j1 = min(j0 + seg, NG + g.ny);
    const bool damping = !(damp >= 1.0);  // solver_cpu.cpp:265 skips only when factor >= 1
    double cand = 0.0;
    Prim w0 = to_prim(load_rot<false>(Ut, g.ps, (long long)(j0 - 2) * g.tnx + ic), ph);
    Prim w1 = to_prim(load_rot<false>(Ut, g.ps, (long long)(j0 - 1) * g.tnx + ic), ph);
    Face hprev;  // upper interface state of the previous cell
    Cons fprev;  // face flux below the cell being updated
    for (int r = j0; r <= j1 + 1; ++r) {
        const Prim w2 = to_prim(load_rot<false>(Ut, g.ps, (long long)r * g.tnx + ic), ph);
        Face lo, hi;
        reconstruct<LIMITER>(w0, w1, w2, hdtd, ph, lo, hi);  // cell r-1
        if (r >= j0 + 1) {
            const Cons f = riemann<SOLVER, false>(hprev.u, hprev.w, lo.u, lo.w, ph);  // face (r-2 | r-1)
            if (r >= j0 + 2) {
                const int j = r - 2;
                const long long k = (long long)j * g.tnx + ic;
                const Cons u = load_rot<false>(Ut, g.ps, k);
                Cons n = update(u, dtd, f, fprev);
                if (active) {
                    if (frame && (j < 2 * NG || j >= g.ny || i < 2 * NG || i >= g.nx)) frame_put(frame, g, i, j, n.psi);
                    if (damping) n.psi *= damp;
                    store_rot<false>(Un, g.ps, k, n);
                    if (acc) {
                        Cons x = n;  // back to the x frame for the CFL candidate
                        x.mn = n.mt; x.mt = n.mn; x.bn = n.bt; x.bt = n.bn;
                        cand = std_max(cand, cfl_candidate(x, ph));
                    }
                }
            }
            fprev = f;
        }
        hprev = hi;
        w0 = w1;
        w1 = w2;
    }
    if (acc) wave_max_to(acc, cand);
}

// ---------------------------------------------------------------------------------------
// Ghost cells (head/cpu/boundary_cpu.hpp:12-149 composed into a closed form)
// ---------------------------------------------------------------------------------------
// The reference fills a grid's ghosts by copy_ghost_cells(prev, dst), then the
// left/right pass over all rows, then the bottom/top pass over all columns.  The
// result at any ghost cell is therefore either an interior cell of dst with the
// reflections of the sides it was mapped through, or (through a Dirichlet side)
// the value the previous grid holds at a partially mapped position.
struct Src {
    int i, j;       // source cell
    bool prev;      // read the previous grid instead of dst
    bool flip_x;    // negate rho*u, Bx
    bool flip_y;    // negate rho*v, By
};

// Maps a ghost index along one axis: returns false for Dirichlet (index kept).
MHD_DI bool map_axis(int c, int n, int type_lo, int type_hi, int& s, bool& flip) {
    const int b = NG, e = NG + n;
    flip = false;
    if (c < b) {
        const int h = b - 1 - c;
        if (type_lo == 0) { s = e - 1 - h; return true; }
        if (type_lo == 1) { s = b; return true; }
        if (type_lo == 2) { s = b + h; flip = true; return true; }
        s = c;
        return false;
    }
    if (c >= e) {
        const int h = c - e;
        if (type_hi == 0) { s = b + h; return true; }
        if (type_hi == 1) { s = e - 1; return true; }
        if (type_hi == 2) { s = e - 1 - h; flip = true; return true; }
        s = c;
        return false;
    }
    s = c;
    return true;
}

MHD_DI Src ghost_source(int gi, int gj, const Geom& g, const int* bc) {
    Src s{gi, gj, false, false, false};
    int si, sj;
    bool fx, fy;
    const bool ghost_row = gj < NG || gj >= NG + g.ny;
    if (ghost_row) {
        const bool okj = map_axis(gj, g.ny, bc[2], bc[3], sj, fy);
        if (okj) {
            // bottom/top pass reads row sj (interior) after the left/right pass
            s.j = sj; s.flip_y = fy;
            const bool oki = map_axis(gi, g.nx, bc[0], bc[1], si, fx);
            if (oki) { s.i = si; s.flip_x = fx; }
            else { s.i = gi; s.prev = true; }  // Dirichlet ghost column at an interior row
        } else {
            // Dirichlet bottom/top: value after the left/right pass on the copied row
            s.j = gj;
            s.prev = true;
            const bool oki = map_axis(gi, g.nx, bc[0], bc[1], si, fx);
            if (oki) { s.i = si; s.flip_x = fx; }
        }
    } else {
        const bool oki = map_axis(gi, g.nx, bc[0], bc[1], si, fx);
        if (oki) { s.i = si; s.flip_x = fx; }
        else s.prev = true;
    }
    return s;
}

// One thread per ghost cell: 2NG columns x ny interior rows, then 2NG rows x tnx.
__global__ __launch_bounds__(256) void bc_kernel(double* __restrict__ U, const double* __restrict__ prev, Geom g,
                                                 int bc0, int bc1, int bc2, int bc3,
                                                 const double* __restrict__ frame) {
    const int bc[4] = {bc0, bc1, bc2, bc3};
    const int n_side = 2 * NG * g.ny;
    const int n_tb = 2 * NG * g.tnx;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_side + n_tb) return;
    int gi, gj;
    if (t < n_side) {
        const int c = t / g.ny;
        gj = NG + t % g.ny;
        gi = c < NG ? c : g.nx + c;  // 0,1 then ie, ie+1
    } else {
        const int q = t - n_side;
        const int r = q / g.tnx;
        gi = q % g.tnx;
        gj = r < NG ? r : g.ny + r;
    }
    const Src s = ghost_source(gi, gj, g, bc);
    const double* src = s.prev ? prev : U;
    const long long ks = (long long)s.j * g.tnx + s.i;
    const long long kd = (long long)gj * g.tnx + gi;
    double v[9];
#pragma unroll
    for (int f = 0; f < 9; ++f) v[f] = src[f * g.ps + ks];
    // psi of interior sources comes from the undamped frame when one is given
    if (frame && !s.prev) v[F_PSI] = frame_get(frame, g, s.i, s.j);
    if (s.flip_x) { v[F_MX] = -v[F_MX]; v[F_BX] = -v[F_BX]; }
    if (s.flip_y) { v[F_MY] = -v[F_MY]; v[F_BY] = -v[F_BY]; }
#pragma unroll
    for (int f = 0; f < 9; ++f) U[f * g.ps + kd] = v[f];
}

// ---------------------------------------------------------------------------------------
// Stand-alone CFL reduction (solver_cpu.cpp:283-312)
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void cfl_kernel(const double* __restrict__ U, Geom g, Phys ph,
                                                  unsigned long long* __restrict__ acc) {
    const long long n = (long long)g.nx * g.ny;
    double m = 0.0;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
        const int j = NG + (int)(q / g.nx);
        const int i = NG + (int)(q % g.nx);
        m = std_max(m, cfl_candidate(load_rot<true>(U, g.ps, (long long)j * g.tnx + i), ph));
    }
    wave_max_to(acc, m);
}

}  // namespace mhd

// ===========================================================================================
// Launch helpers used by the host layer (api.cpp)
// ===========================================================================================
#include "launch.hpp"

namespace mhd {

template <int S, int L>
static hipError_t launch_step_sl(const StepArgs& a) {
    const Geom g{a.nx, a.ny, a.nx + 2 * NG, a.ny + 2 * NG, a.ps};
    Phys ph{a.gamma, a.gamma - 1.0, a.ch, a.ch * a.ch, 0.5 * a.ch, 0.5 / a.ch};
    const double dtdx = a.dt / a.dx, dtdy = a.dt / a.dy;
    constexpr int WG = 256;
    const long long flat = (long long)g.tnx * g.ny + 4;
    const unsigned nbx = (unsigned)((flat + (WG - 4) - 1) / (WG - 4));
    hipError_t e;
    if (a.ev && (e = hipEventRecord(a.ev[0], a.stream)) != hipSuccess) return e;
    hipLaunchKernelGGL((xsweep_kernel<S, L, WG>), dim3(nbx), dim3(WG), 0, a.stream, a.Uold, a.Utmp, g, ph, dtdx,
                       0.5 * dtdx, a.acc);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (a.ev && (e = hipEventRecord(a.ev[1], a.stream)) != hipSuccess) return e;
    const int nbc = 2 * NG * g.ny + 2 * NG * g.tnx;
    hipLaunchKernelGGL(bc_kernel, dim3((nbc + 255) / 256), dim3(256), 0, a.stream, a.Utmp, a.Uold, g, a.bc[0],
                       a.bc[1], a.bc[2], a.bc[3], (const double*)nullptr);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (a.ev && (e = hipEventRecord(a.ev[2], a.stream)) != hipSuccess) return e;
    const int seg = a.seg > 0 ? a.seg : 64;
    dim3 gy((unsigned)((g.nx + 255) / 256), (unsigned)((g.ny + seg - 1) / seg));
    hipLaunchKernelGGL((ysweep_kernel<S, L>), gy, dim3(256), 0, a.stream, a.Utmp, a.Unew, g, ph, dtdy, 0.5 * dtdy,
                       a.damp, seg, a.frame, a.acc);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (a.ev && (e = hipEventRecord(a.ev[3], a.stream)) != hipSuccess) return e;
    hipLaunchKernelGGL(bc_kernel, dim3((nbc + 255) / 256), dim3(256), 0, a.stream, a.Unew, a.Utmp, g, a.bc[0],
                       a.bc[1], a.bc[2], a.bc[3], (const double*)a.frame);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (a.ev && (e = hipEventRecord(a.ev[4], a.stream)) != hipSuccess) return e;
    return hipSuccess;
}

hipError_t launch_step(const StepArgs& a) {
    switch (a.solver * 2 + a.limiter) {
        case 0: return launch_step_sl<0, 0>(a);
        case 1: return launch_step_sl<0, 1>(a);
        case 2: return launch_step_sl<1, 0>(a);
        case 3: return launch_step_sl<1, 1>(a);
        case 4: return launch_step_sl<2, 0>(a);
        case 5: return launch_step_sl<2, 1>(a);
        case 6: return launch_step_sl<3, 0>(a);
        case 7: return launch_step_sl<3, 1>(a);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_cfl(const double* U, int nx, int ny, long long ps, double gamma, double ch,
                      unsigned long long* acc, hipStream_t s) {
    const Geom g{nx, ny, nx + 2 * NG, ny + 2 * NG, ps};
    Phys ph{gamma, gamma - 1.0, ch, ch * ch, 0.5 * ch, 0.5 / ch};
    const long long n = (long long)nx * ny;
    long long nb = (n + 255) / 256;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(cfl_kernel, dim3((unsigned)nb), dim3(256), 0, s, U, g, ph, acc);
    return hipGetLastError();
}

hipError_t launch_bc(double* U, const double* prev, int nx, int ny, long long ps, const int* bc, hipStream_t s) {
    const Geom g{nx, ny, nx + 2 * NG, ny + 2 * NG, ps};
    const int nbc = 2 * NG * g.ny + 2 * NG * g.tnx;
    hipLaunchKernelGGL(bc_kernel, dim3((nbc + 255) / 256), dim3(256), 0, s, U, prev, g, bc[0], bc[1], bc[2], bc[3],
                       (const double*)nullptr);
    return hipGetLastError();
}

}  // namespace mhd
