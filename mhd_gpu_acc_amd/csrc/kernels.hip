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
//             interface state and the lower face flux carried; the raw rows
//             for the update in an LDS ring).  Fuses the psi
//             damping (solver_cpu.cpp:263-275) and the next step's CFL candidate.
//   cfl     : stand-alone compute_dt reduction for a state not produced by ysweep.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "sweeps.hpp"

MHD_FP_CONTRACT

namespace mhd {

// Knob (compile time): MHD_X_WAVES_PER_EU asks the compiler for that many resident waves
// per SIMD in the x sweep (0: its own choice, 2 for HLLD).
#ifndef MHD_X_WAVES_PER_EU
#define MHD_X_WAVES_PER_EU 0
#endif
#if MHD_X_WAVES_PER_EU > 0
#define MHD_X_ATTR __attribute__((amdgpu_waves_per_eu(MHD_X_WAVES_PER_EU)))
#else
#define MHD_X_ATTR
#endif

// Segments 0 .. npers-1 go to the first gp blocks, block b taking b, b+gp, ... (loading
// each next segment while it computes the current one); every further block takes one of
// the remaining segments.  Blocks are dispatched in order, so those one-segment blocks
// fill the chip while the last rounds of the looping blocks finish.
template <int SOLVER, int LIMITER, int WG, bool BUF>
__global__ __launch_bounds__(WG) MHD_X_ATTR void xsweep_kernel(const double* __restrict__ Uo, double* __restrict__ Ut, Geom g,
                                                    Phys ph, double dtd, double hdtd,
                                                    unsigned long long* __restrict__ acc, int gp, int npers,
                                                    RowPush rp) {
    __shared__ XLds<WG> L;
    const int t = threadIdx.x;
    if (acc && blockIdx.x == 0 && t == 0) *acc = 0ull;  // CFL accumulator for this step's ysweep
    const std::conditional_t<BUF, BufPlanes<true>, Planes<true, const double>> src(Uo, g.ps);
    const std::conditional_t<BUF, BufPlanes<true>, Planes<true, double>> dst(Ut, g.ps);
    const int last = g.tnx * g.tny - 1;
    const int kb = NG * g.tnx - 2 + t;
    const bool looping = (int)blockIdx.x < gp;
    int s = looping ? (int)blockIdx.x : npers + ((int)blockIdx.x - gp);
    const int end = looping ? npers : s + 1;
    if (s >= end) return;
    int k = kb + s * (WG - 4);
    Cons un = src.load(min(k, last));  // clamp: tail lanes read valid memory
    // persistent blocks only: the one-segment blocks after them keep the default priority
    const Progress<MHD_PRIO_X != 0> prog((end - s + gp - 1) / gp, looping);
    for (int it = 0;; ++it) {
        prog.at(it);
        const int sn = s + gp;
        const int kn = kb + sn * (WG - 4);
        const Cons u = un;  // the segment loaded one iteration ago
        if (sn < end) un = src.load(min(kn, last));
        xsweep_segment<SOLVER, LIMITER, WG>(u, k, t, dst, g, ph, dtd, hdtd, L, rp);
        if (sn >= end) break;
        lds_barrier();  // sw is rewritten by the next segment
        s = sn;
        k = kn;
    }
}

template <int SOLVER, int LIMITER, bool BUF>
__global__ __launch_bounds__(256) MHD_Y_ATTR void ysweep_kernel(const double* __restrict__ Ut,
                                                                double* __restrict__ Un, Geom g, Phys ph,
                                                                double dtd, double hdtd, double damp, int seg,
                                                                double* __restrict__ frame,
                                                                unsigned long long* __restrict__ acc, int prio) {
    extern __shared__ double ring[];  // YRING_DOUBLES doubles per lane of the block (set at launch)
    // blockIdx.x: row segment, blockIdx.y: column block.  Workgroups go to the XCDs in turn
    // by linear index, so each XCD marches segments of every column block; with the
    // column block first it marched two column blocks only, and the work per column block
    // differs with the flow (4096^2 HLLD y sweep 1 % slower).
    const int i = NG + blockIdx.y * blockDim.x + threadIdx.x;
    const int j0 = NG + blockIdx.x * seg;
    const int j1 = min(j0 + seg, NG + g.ny);
    const double cand = ysweep_column<SOLVER, LIMITER, BUF>(Ut, Un, g, ph, dtd, hdtd, damp, frame, acc != nullptr, i,
                                                       i < NG + g.nx, j0, j1,
                                                       YRing::of(ring, (int)threadIdx.x), prio != 0);
    if (acc) wave_max_to(acc, cand);
}

__global__ __launch_bounds__(256) void bc_kernel(double* __restrict__ U, const double* __restrict__ prev, Geom g,
                                                 int bc0, int bc1, int bc2, int bc3,
                                                 const double* __restrict__ frame,
                                                 const unsigned long long* __restrict__ acc,
                                                 unsigned long long* acc_host, unsigned long long seq) {
    const int bc[4] = {bc0, bc1, bc2, bc3};
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ghost_count(g)) ghost_fill(U, prev, g, bc, frame, t);
    // Hand the previous launch's CFL maximum to the host: a system-scope store into pinned
    // memory replaces a device-to-host copy after the step, and the step's number, stored
    // after it with release order, tells a polling host that it has arrived.
    if (acc_host && t == 0) {
        __hip_atomic_store(acc_host, *acc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(acc_host + 1, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
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

// max over interior cells of |B| = sqrt(Bx^2 + By^2 + Bz^2), the driver's progress
// column (main_cpu.cpp:224-230); std::max order, so NaN cells are skipped as there.
__global__ __launch_bounds__(256) void max_b_kernel(const double* __restrict__ U, Geom g,
                                                    unsigned long long* __restrict__ acc) {
    const long long n = (long long)g.nx * g.ny;
    double m = 0.0;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
        const long long k = (long long)(NG + q / g.nx) * g.tnx + NG + q % g.nx;
        const double bx = U[F_BX * g.ps + k], by = U[F_BY * g.ps + k], bz = U[F_BZ * g.ps + k];
        m = std_max(m, __builtin_sqrt(bx * bx + by * by + bz * bz));
    }
    wave_max_to(acc, m);
}

// ---------------------------------------------------------------------------------------
// Streaming kernels of known traffic (mhd_probe_bandwidth): the HBM bandwidth the device
// reaches for a copy, a read and a write, against which the sweeps' figures are read.
// 16-byte lanes, grid-stride.
// ---------------------------------------------------------------------------------------
typedef double bw2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void bw_copy_kernel(const bw2* __restrict__ a, bw2* __restrict__ b, long long n) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) b[i] = a[i];
}
__global__ __launch_bounds__(256) void bw_read_kernel(const bw2* __restrict__ a, double* __restrict__ sink, long long n) {
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        s += a[i].x + a[i].y;
    if (s != s) sink[0] = s;  // keeps the loads; written only for NaN data
}
__global__ __launch_bounds__(256) void bw_write_kernel(bw2* __restrict__ b, double v, long long n) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        b[i] = bw2{v, v};
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

template <int S, int L, int WG, bool BUF>
static void launch_x(const StepArgs& a, const Geom& g, const Phys& ph, double dtdx, const RowPush& rp) {
    const long long flat = (long long)g.tnx * g.ny;
    const int nseg = (int)((flat + (WG - 4) - 1) / (WG - 4));
    // xgrid > 0: persistent grid of xgrid blocks per CU, plus xtail per mille of the
    // segments in one-segment blocks after it; 0: one block per segment
    const int gp = a.xgrid > 0 ? min(nseg, a.xgrid * a.cus) : nseg;
    const int tail = a.xgrid > 0 ? (int)((long long)nseg * a.xtail / 1000) : 0;
    const int npers = nseg - tail;
    hipLaunchKernelGGL((xsweep_kernel<S, L, WG, BUF>), dim3((unsigned)(gp + tail)), dim3(WG), 0, a.stream, a.Uold,
                       a.Utmp, g, ph, dtdx, 0.5 * dtdx, a.acc, gp, npers, rp);
}

template <int S, int L>
static hipError_t launch_step_sl(const StepArgs& a) {
    const Geom g{a.nx, a.ny, a.nx + 2 * NG, a.ny + 2 * NG, a.ps};
    Phys ph{a.gamma, a.gamma - 1.0, a.ch, a.ch * a.ch, 0.5 * a.ch, 0.5 / a.ch};
    const double dtdx = a.dt / a.dx, dtdy = a.dt / a.dy;
    hipError_t e;
    // Buffer-resource addressing for grids of at most 4 GiB (an OT 4096^2 grid is 1.2 GB).
    const bool buf = a.addr_buf && 9ull * (unsigned long long)a.ps * 8 <= BUF_MAX_BYTES;
    // Utmp's ghost cells: filled by a launch of their own for callers that read Utmp's ghost
    // columns (and across a Dirichlet bottom or top side, whose ghost rows come from Uold);
    // otherwise the x sweep writes the ghost rows the y sweep reads itself.
    const RowPush rp = a.fill_tmp ? RowPush{} : row_push(g.ny, a.bc[2], a.bc[3]);
    if (a.ev && (e = hipEventRecord(a.ev[0], a.stream)) != hipSuccess) return e;
    if (a.xwg == 64) buf ? launch_x<S, L, 64, true>(a, g, ph, dtdx, rp) : launch_x<S, L, 64, false>(a, g, ph, dtdx, rp);
    else if (a.xwg == 128)
        buf ? launch_x<S, L, 128, true>(a, g, ph, dtdx, rp) : launch_x<S, L, 128, false>(a, g, ph, dtdx, rp);
    else buf ? launch_x<S, L, 256, true>(a, g, ph, dtdx, rp) : launch_x<S, L, 256, false>(a, g, ph, dtdx, rp);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (a.ev && (e = hipEventRecord(a.ev[1], a.stream)) != hipSuccess) return e;
    const int nbc = 2 * NG * g.ny + 2 * NG * g.tnx;
    if (rp.n == 0) {
        hipLaunchKernelGGL(bc_kernel, dim3((nbc + 255) / 256), dim3(256), 0, a.stream, a.Utmp, a.Uold, g, a.bc[0],
                           a.bc[1], a.bc[2], a.bc[3], (const double*)nullptr, (const unsigned long long*)nullptr,
                           (unsigned long long*)nullptr, 0ull);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (a.x_only) return hipSuccess;
    if (a.ev && (e = hipEventRecord(a.ev[2], a.stream)) != hipSuccess) return e;
    // Rows per lane segment: enough column-waves x segments for 2 waves per SIMD
    // (8 per CU), at most 64 rows (the march's 2-row halo is then 3 %; 64 rows took 0.5 %
    // off the 4096^2 y sweep against 32) and at least 4.
    int seg = a.seg;
    if (seg <= 0) {
        const long long colwaves = (g.nx + 63) / 64;
        const long long rows = colwaves * g.ny / (8LL * (a.cus > 0 ? a.cus : 1));
        seg = (int)(rows < 4 ? 4 : (rows > 64 ? 64 : rows));
    }
    const int ywg = (a.ywg == 64 || a.ywg == 128) ? a.ywg : 256;
    dim3 gy((unsigned)((g.ny + seg - 1) / seg), (unsigned)((g.nx + ywg - 1) / ywg));
    // progress-ordered wave priority when the grid holds two or more rounds of waves at two
    // per SIMD (8 per CU)
    const long long ywaves = (long long)gy.x * gy.y * (ywg / 64);
    const int prio = a.cus > 0 && ywaves >= 2LL * 8 * a.cus;
    const size_t lds = (size_t)YRING_DOUBLES * ywg * sizeof(double);
    if (buf)
        hipLaunchKernelGGL((ysweep_kernel<S, L, true>), gy, dim3(ywg), lds, a.stream, a.Utmp, a.Unew, g, ph, dtdy,
                           0.5 * dtdy, a.damp, seg, a.frame, MHD_FUSED_CFL ? a.acc : nullptr, prio);
    else
        hipLaunchKernelGGL((ysweep_kernel<S, L, false>), gy, dim3(ywg), lds, a.stream, a.Utmp, a.Unew, g, ph, dtdy,
                           0.5 * dtdy, a.damp, seg, a.frame, MHD_FUSED_CFL ? a.acc : nullptr, prio);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (a.ev && (e = hipEventRecord(a.ev[3], a.stream)) != hipSuccess) return e;
    const bool to_host = MHD_FUSED_CFL && a.acc && a.acc_host;
    // Unew's ghost cells.  Those the reference copies from the previous grid (Utmp, through
    // a Dirichlet side) are read from Uold: Utmp's fill copied each of them from Uold's cell
    // at the same position, so the values are the same whether or not Utmp's ghost cells
    // were filled.
    hipLaunchKernelGGL(bc_kernel, dim3((nbc + 255) / 256), dim3(256), 0, a.stream, a.Unew, a.Uold, g, a.bc[0],
                       a.bc[1], a.bc[2], a.bc[3], (const double*)a.frame, (const unsigned long long*)a.acc,
                       to_host ? a.acc_host : (unsigned long long*)nullptr, a.seq);
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

hipError_t launch_max_b(const double* U, int nx, int ny, long long ps, unsigned long long* acc, hipStream_t s) {
    const Geom g{nx, ny, nx + 2 * NG, ny + 2 * NG, ps};
    long long nb = ((long long)nx * ny + 255) / 256;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(max_b_kernel, dim3((unsigned)nb), dim3(256), 0, s, U, g, acc);
    return hipGetLastError();
}

hipError_t launch_bandwidth(int kind, const double* a, double* b, long long bytes, int cus, hipStream_t s) {
    const long long n = bytes / 16;
    const dim3 grid((unsigned)(4 * (cus > 0 ? cus : 1)));  // 4 workgroups per CU: the best copy rate measured
    if (kind == 0)
        hipLaunchKernelGGL(bw_copy_kernel, grid, dim3(256), 0, s, (const bw2*)a, (bw2*)b, n);
    else if (kind == 1)
        hipLaunchKernelGGL(bw_read_kernel, grid, dim3(256), 0, s, (const bw2*)a, b, n);
    else
        hipLaunchKernelGGL(bw_write_kernel, grid, dim3(256), 0, s, (bw2*)b, 0.5, n);
    return hipGetLastError();
}

hipError_t launch_bc(double* U, const double* prev, int nx, int ny, long long ps, const int* bc, hipStream_t s) {
    const Geom g{nx, ny, nx + 2 * NG, ny + 2 * NG, ps};
    const int nbc = 2 * NG * g.ny + 2 * NG * g.tnx;
    hipLaunchKernelGGL(bc_kernel, dim3((nbc + 255) / 256), dim3(256), 0, s, U, prev, g, bc[0], bc[1], bc[2], bc[3],
                       (const double*)nullptr, (const unsigned long long*)nullptr, (unsigned long long*)nullptr, 0ull);
    return hipGetLastError();
}

}  // namespace mhd
