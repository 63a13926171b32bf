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

// The nine field planes of a grid, rotated for a sweep direction.  The bases are
// uniform (scalar registers); lanes add a 32-bit cell offset.
template <bool XDIR, class T> struct Planes {
    T* r; T* mn; T* mt; T* mz; T* bn; T* bt; T* bz; T* e; T* psi;
    MHD_DI Planes(T* U, long long ps)
        : r(U), mn(U + (XDIR ? F_MX : F_MY) * ps), mt(U + (XDIR ? F_MY : F_MX) * ps), mz(U + F_MZ * ps),
          bn(U + (XDIR ? F_BX : F_BY) * ps), bt(U + (XDIR ? F_BY : F_BX) * ps), bz(U + F_BZ * ps),
          e(U + F_E * ps), psi(U + F_PSI * ps) {}
    MHD_DI Cons load(int k) const { return Cons{r[k], mn[k], mt[k], mz[k], bn[k], bt[k], bz[k], e[k], psi[k]}; }
    MHD_DI void store(int k, const Cons& u) const {
        r[k] = u.r; mn[k] = u.mn; mt[k] = u.mt; mz[k] = u.mz; bn[k] = u.bn; bt[k] = u.bt; bz[k] = u.bz;
        e[k] = u.e; psi[k] = u.psi;
    }
};

template <bool XDIR>
MHD_DI Cons load_rot(const double* __restrict__ U, long long ps, long long k) {
    return Planes<XDIR, const double>(U, ps).load((int)k);
}

// Workgroup barrier that orders LDS only.  __syncthreads() fences global memory
// too, which makes every lane wait for its in-flight global loads (the next
// segment's prefetch) before the barrier.
MHD_DI void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// LDS staging of a 9-double record, field-major so that lane-consecutive
// accesses hit consecutive banks.
template <int WG, int NF = 9> struct Lds9 {
    double v[NF][WG];
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
        if constexpr (NF > 9) v[9][t] = w.b2;
    }
    MHD_DI Prim getp(int t) const {
        double b2 = 0.0;
        if constexpr (NF > 9) b2 = v[9][t];
        return Prim{v[0][t], v[1][t], v[2][t], v[3][t], v[4][t], v[5][t], v[6][t], v[7][t], v[8][t], b2};
    }
};

// ---------------------------------------------------------------------------------------
// x sweep
// ---------------------------------------------------------------------------------------
// Segment s covers flattened cells [jb*tnx - 2 + s*(WG-4), +WG) of the interior
// rows; lanes 2..WG-3 that land on an interior column write Utmp.  A block walks
// segments blockIdx.x, +gridDim.x, ... and loads the next segment's cells before
// computing the current one, so global-load latency overlaps the arithmetic.
template <int SOLVER, int LIMITER, int WG>
__global__ __launch_bounds__(WG) void xsweep_kernel(const double* __restrict__ Uo, double* __restrict__ Ut, Geom g,
                                                    Phys ph, double dtd, double hdtd,
                                                    unsigned long long* __restrict__ acc, int nseg) {
    __shared__ Lds9<WG> sw;      // primitives, later face fluxes
    __shared__ Lds9<WG> su;      // lower interface states
    __shared__ Lds9<WG, 10> sv;  // their primitives
    const int t = threadIdx.x;
    if (acc && blockIdx.x == 0 && t == 0) *acc = 0ull;  // CFL accumulator for this step's ysweep
    const Planes<true, const double> src(Uo, g.ps);
    const Planes<true, double> dst(Ut, g.ps);
    const int last = g.tnx * g.tny - 1;
    const int kb = NG * g.tnx - 2 + t;
    int s = blockIdx.x;
    if (s >= nseg) return;
    int k = kb + s * (WG - 4);
    Cons un = src.load(min(k, last));  // clamp: tail lanes read valid memory
    for (;;) {
        const Cons u = un;  // the segment loaded one iteration ago
        const int sn = s + gridDim.x;
        const int kn = kb + sn * (WG - 4);
        if (sn < nseg) un = src.load(min(kn, last));
        const Prim w = to_prim(u, ph);
        sw.put(t, w);
        lds_barrier();
        Face lo, hi;
        {
            const int tm = t > 0 ? t - 1 : t;
            const int tp = t < WG - 1 ? t + 1 : t;
            reconstruct<LIMITER>(sw.getp(tm), w, sw.getp(tp), hdtd, ph, lo, hi);
        }
        su.put(t, lo.u);
        sv.put(t, lo.w);
        lds_barrier();
        const int tp = t < WG - 1 ? t + 1 : t;
        const Cons f = riemann<SOLVER, true>(hi.u, hi.w, su.get(tp), sv.getp(tp), ph);  // face (this | next)
        sw.put(t, f);
        lds_barrier();
        if (t >= 2 && t < WG - 2 && k <= last) {
            const int i = k % g.tnx;
            const int j = k / g.tnx;
            if (i >= NG && i < NG + g.nx && j < NG + g.ny) dst.store(k, update(u, dtd, f, sw.get(t - 1)));
        }
        if (sn >= nseg) break;
        lds_barrier();  // sw is rewritten by the next segment
        s = sn;
        k = kn;
    }
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

// Knobs (compile time): MHD_Y_WAVES_PER_EU asks the compiler for that many
// resident waves per SIMD; MHD_Y_CARRY_PRIM keeps the primitive of the upper
// interface state across rows instead of recomputing it.
#ifndef MHD_Y_WAVES_PER_EU
#define MHD_Y_WAVES_PER_EU 0
#endif
#ifndef MHD_Y_CARRY_PRIM
#define MHD_Y_CARRY_PRIM 1
#endif
#if MHD_Y_WAVES_PER_EU > 0
#define MHD_Y_ATTR __attribute__((amdgpu_waves_per_eu(MHD_Y_WAVES_PER_EU)))
#else
#define MHD_Y_ATTR
#endif

template <int SOLVER, int LIMITER>
__global__ __launch_bounds__(256) MHD_Y_ATTR void ysweep_kernel(const double* __restrict__ Ut,
                                                                double* __restrict__ Un, Geom g, Phys ph,
                                                                double dtd, double hdtd, double damp, int seg,
                                                                double* __restrict__ frame,
                                                                unsigned long long* __restrict__ acc) {
    const Planes<false, const double> src(Ut, g.ps);
    const Planes<false, double> dst(Un, g.ps);
    const int i = NG + blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i < NG + g.nx;
    const int ic = active ? i : NG;  // idle lanes shadow column NG (no stores)
    const int j0 = NG + blockIdx.y * seg;
    const int j1 = min(j0 + seg, NG + g.ny);
    const bool damping = !(damp >= 1.0);  // solver_cpu.cpp:265 skips only when factor >= 1
    double cand = 0.0;
    Prim w0 = to_prim(src.load((j0 - 2) * g.tnx + ic), ph);
    Prim w1 = to_prim(src.load((j0 - 1) * g.tnx + ic), ph);
    Cons hu;     // upper interface state of the previous cell
#if MHD_Y_CARRY_PRIM
    Prim hw;     // and its primitive
#endif
    Cons fprev;  // face flux below the cell being updated
    for (int r = j0; r <= j1 + 1; ++r) {
        const Prim w2 = to_prim(src.load(r * g.tnx + ic), ph);
        Face lo, hi;
        reconstruct<LIMITER>(w0, w1, w2, hdtd, ph, lo, hi);  // cell r-1
        if (r >= j0 + 1) {
#if MHD_Y_CARRY_PRIM
            const Cons f = riemann<SOLVER, false>(hu, hw, lo.u, lo.w, ph);  // face (r-2 | r-1)
#else
            const Cons f = riemann<SOLVER, false>(hu, to_prim(hu, ph), lo.u, lo.w, ph);
#endif
            if (r >= j0 + 2) {
                const int j = r - 2;
                const int k = j * g.tnx + ic;
                Cons n = update(src.load(k), dtd, f, fprev);
                if (active) {
                    if (frame && (j < 2 * NG || j >= g.ny || i < 2 * NG || i >= g.nx)) frame_put(frame, g, i, j, n.psi);
                    if (damping) n.psi *= damp;
                    dst.store(k, n);
                    if (acc) {
                        Cons x = n;  // back to the x frame for the CFL candidate
                        x.mn = n.mt; x.mt = n.mn; x.bn = n.bt; x.bt = n.bn;
                        cand = std_max(cand, cfl_candidate(x, ph));
                    }
                }
            }
            fprev = f;
        }
        hu = hi.u;
#if MHD_Y_CARRY_PRIM
        hw = hi.w;
#endif
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

#ifndef MHD_FUSED_CFL
#define MHD_FUSED_CFL 1
#endif

namespace mhd {

bool fused_cfl() { return MHD_FUSED_CFL != 0; }

template <int S, int L, int WG>
static void launch_x(const StepArgs& a, const Geom& g, const Phys& ph, double dtdx) {
    const long long flat = (long long)g.tnx * g.ny;
    const int nseg = (int)((flat + (WG - 4) - 1) / (WG - 4));
    // xgrid > 0: persistent grid of xgrid blocks per CU; 0: one block per segment
    const int nb = a.xgrid > 0 ? min(nseg, a.xgrid * a.cus) : nseg;
    hipLaunchKernelGGL((xsweep_kernel<S, L, WG>), dim3((unsigned)nb), dim3(WG), 0, a.stream, a.Uold, a.Utmp, g, ph,
                       dtdx, 0.5 * dtdx, a.acc, nseg);
}

template <int S, int L>
static hipError_t launch_step_sl(const StepArgs& a) {
    const Geom g{a.nx, a.ny, a.nx + 2 * NG, a.ny + 2 * NG, a.ps};
    Phys ph{a.gamma, a.gamma - 1.0, a.ch, a.ch * a.ch, 0.5 * a.ch, 0.5 / a.ch};
    const double dtdx = a.dt / a.dx, dtdy = a.dt / a.dy;
    hipError_t e;
    if (a.ev && (e = hipEventRecord(a.ev[0], a.stream)) != hipSuccess) return e;
    if (a.xwg == 64) launch_x<S, L, 64>(a, g, ph, dtdx);
    else if (a.xwg == 128) launch_x<S, L, 128>(a, g, ph, dtdx);
    else launch_x<S, L, 256>(a, g, ph, dtdx);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (a.ev && (e = hipEventRecord(a.ev[1], a.stream)) != hipSuccess) return e;
    const int nbc = 2 * NG * g.ny + 2 * NG * g.tnx;
    hipLaunchKernelGGL(bc_kernel, dim3((nbc + 255) / 256), dim3(256), 0, a.stream, a.Utmp, a.Uold, g, a.bc[0],
                       a.bc[1], a.bc[2], a.bc[3], (const double*)nullptr);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (a.ev && (e = hipEventRecord(a.ev[2], a.stream)) != hipSuccess) return e;
    const int seg = a.seg > 0 ? a.seg : 32;
    const int ywg = (a.ywg == 64 || a.ywg == 128) ? a.ywg : 256;
    dim3 gy((unsigned)((g.nx + ywg - 1) / ywg), (unsigned)((g.ny + seg - 1) / seg));
    hipLaunchKernelGGL((ysweep_kernel<S, L>), gy, dim3(ywg), 0, a.stream, a.Utmp, a.Unew, g, ph, dtdy, 0.5 * dtdy,
                       a.damp, seg, a.frame, MHD_FUSED_CFL ? a.acc : nullptr);
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
