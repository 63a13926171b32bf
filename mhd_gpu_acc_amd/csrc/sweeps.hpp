// Device building blocks of the sweeps, shared by the plain kernels (kernels.hip)
// and the batched kernels (batch.hip): field planes, LDS staging, one x-sweep
// segment, one y-sweep column march, the closed-form ghost fill and the CFL
// reduction.  See kernels.hip for the memory layout.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "physics.hpp"

MHD_FP_CONTRACT

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

// Global-memory view of a device pointer: its accesses become global_*
// instructions.  A pointer read from memory (not a kernel argument) is otherwise
// accessed with flat_* instructions, which the compiler must order against LDS
// traffic.
template <class T> using GlobalT = __attribute__((address_space(1))) T;
template <class T> MHD_DI GlobalT<T>* as_global(T* p) { return (GlobalT<T>*)p; }

// Buffer-resource view of the nine planes of a grid of at most 4 GiB - 1 byte: one
// resource (4 SGPRs) for the grid and the plane offsets as scalar offsets, so the nine
// accesses of a cell share one 32-bit lane offset instead of each forming a 64-bit
// address.  Every access is in range by construction.
constexpr unsigned long long BUF_MAX_BYTES = 0xFFFFFFFFull;
template <bool XDIR> struct BufPlanes {
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    __amdgpu_buffer_rsrc_t rs;
    unsigned ps8;  // plane stride in bytes
    MHD_DI BufPlanes(const double* U, long long ps)
        : rs(__builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(U), 0, (int)(unsigned)(9ull * ps * 8), 0x00020000)),
          ps8((unsigned)(ps * 8)) {}
    template <int F> MHD_DI double ld(unsigned vo) const {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, (int)vo, (int)(F * ps8), 0));
    }
    template <int F> MHD_DI void st(unsigned vo, double v) const {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), rs, (int)vo, (int)(F * ps8), 0);
    }
    MHD_DI Cons load(int k) const {
        const unsigned vo = (unsigned)k * 8u;
        return Cons{ld<F_RHO>(vo), ld<XDIR ? F_MX : F_MY>(vo), ld<XDIR ? F_MY : F_MX>(vo), ld<F_MZ>(vo),
                    ld<XDIR ? F_BX : F_BY>(vo), ld<XDIR ? F_BY : F_BX>(vo), ld<F_BZ>(vo), ld<F_E>(vo),
                    ld<F_PSI>(vo)};
    }
    MHD_DI void store(int k, const Cons& u) const {
        const unsigned vo = (unsigned)k * 8u;
        st<F_RHO>(vo, u.r); st<XDIR ? F_MX : F_MY>(vo, u.mn); st<XDIR ? F_MY : F_MX>(vo, u.mt); st<F_MZ>(vo, u.mz);
        st<XDIR ? F_BX : F_BY>(vo, u.bn); st<XDIR ? F_BY : F_BX>(vo, u.bt); st<F_BZ>(vo, u.bz); st<F_E>(vo, u.e);
        st<F_PSI>(vo, u.psi);
    }
};

template <bool BUF, class V, class T> MHD_DI V plane_view(T* U, long long ps) {
    if constexpr (BUF) return V(U, ps);
    else return V(as_global(U), ps);
}

// A wave's issue priority ordered by its progress through its share of a sweep
// (s_setprio, the wave's own state): 3 at the start, one step lower at each quarter.  Two
// waves share each SIMD and the sequencer otherwise serves the older first, so the
// younger falls behind and ends alone on its SIMD; ordered by progress, the waves on a
// SIMD keep pace and finish closer together.  ON = false (or active = false, uniform)
// leaves the priority alone.
template <bool ON> struct Progress {
    int q1, q2, q3;  // done counts at which the priority steps down (uniform)
    bool active;
    MHD_DI explicit Progress(int total, bool act = true)
        : q1(total / 4), q2(total / 2), q3(3 * total / 4), active(act) {
        if constexpr (ON) if (active) __builtin_amdgcn_s_setprio(3);
    }
    MHD_DI void at(int done) const {
        if constexpr (ON) {
            if (!active) return;
            if (done == q1) __builtin_amdgcn_s_setprio(2);
            else if (done == q2) __builtin_amdgcn_s_setprio(1);
            else if (done == q3) __builtin_amdgcn_s_setprio(0);
        }
    }
};
// Knobs (compile time): progress-ordered priority in the y march where its launch asks for
// it (MHD_PRIO_Y, default on; the plain launch asks when its grid holds two or more
// rounds of waves: OT 4096^2 HLLD y sweep 0.874 -> 0.847 ms, while at 1024^2, one round,
// it made the sweep 3 % slower) and in the persistent x sweep (MHD_PRIO_X, default off:
// 1 % slower there).
#ifndef MHD_PRIO_Y
#define MHD_PRIO_Y 1
#endif
#ifndef MHD_PRIO_X
#define MHD_PRIO_X 0
#endif

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

template <int WG> struct XLds {
    Lds9<WG> sw;      // primitives, later face fluxes
    Lds9<WG> su;      // lower interface states
    Lds9<WG, 10> sv;  // their primitives
};

// Ghost rows of U* that the x sweep writes itself: a cell of interior row src[e] is also
// stored to ghost row dst[e] of its column (rho v and B_y negated when bit e of flip is
// set), as the ghost fill copies it for an interior column (entries e < n).  The y sweep
// reads only the interior columns of U*, so these rows are all of U*'s ghost cells it
// needs.  n = 0 when a ghost fill of U* follows the sweep instead.
struct RowPush {
    int n;
    int src[4], dst[4];
    unsigned flip;
};

// One flattened segment: lane t holds cell k of the source grid (state u).
// Lanes 2..WG-3 on an interior cell write the updated state to dst (and to the ghost
// rows of rp that copy their row).
template <int SOLVER, int LIMITER, int WG, class DstPlanes>
MHD_DI void xsweep_segment(const Cons& u, int k, int t, const DstPlanes& dst, const Geom& g,
                           const Phys& ph, double dtd, double hdtd, XLds<WG>& L, const RowPush& rp) {
    const Prim w = to_prim(u, ph);
    L.sw.put(t, w);
    lds_barrier();
    Face lo, hi;
    {
        const int tm = t > 0 ? t - 1 : t;
        const int tp = t < WG - 1 ? t + 1 : t;
        reconstruct<LIMITER>(L.sw.getp(tm), w, L.sw.getp(tp), hdtd, ph, lo, hi);
    }
    L.su.put(t, lo.u);
    L.sv.put(t, lo.w);
    lds_barrier();
    const int tp = t < WG - 1 ? t + 1 : t;
    const Cons f = riemann<SOLVER, true>(hi.u, hi.w, L.su.get(tp), L.sv.getp(tp), ph);  // face (this | next)
    L.sw.put(t, f);
    lds_barrier();
    if (t >= 2 && t < WG - 2 && k < g.tnx * g.tny) {
        const int i = k % g.tnx;
        const int j = k / g.tnx;
        if (i >= NG && i < NG + g.nx && j < NG + g.ny) {
            const Cons n = update(u, dtd, f, L.sw.get(t - 1));
            dst.store(k, n);
            // every source row of rp lies within two rows of an interior edge
            if (rp.n > 0 && (j < 2 * NG || j >= g.ny)) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    if (e < rp.n && j == rp.src[e]) {
                        Cons v = n;
                        if ((rp.flip >> e) & 1u) {  // x frame: m_t = rho v, B_t = B_y
                            v.mt = -v.mt;
                            v.bt = -v.bt;
                        }
                        dst.store(rp.dst[e] * g.tnx + i, v);
                    }
                }
            }
        }
    }
}

// The rows of rp for a grid whose bottom and top sides are of types bc_lo and bc_hi (each
// ghost row's source row and reflection from map_axis); n = 0 when a side is Dirichlet,
// whose ghost rows come from the previous grid.
MHD_DI bool map_axis(int c, int n, int type_lo, int type_hi, int& s, bool& flip);
MHD_DI RowPush row_push(int ny, int bc_lo, int bc_hi) {
    RowPush rp{};
    const int rows[4] = {0, 1, NG + ny, NG + ny + 1};
    for (int e = 0; e < 4; ++e) {
        int sj;
        bool flip;
        if (!map_axis(rows[e], ny, bc_lo, bc_hi, sj, flip)) return RowPush{};
        rp.src[e] = sj;
        rp.dst[e] = rows[e];
        if (flip) rp.flip |= 1u << e;
    }
    rp.n = 4;
    return rp;
}

// ---------------------------------------------------------------------------------------
// y sweep (column march)
// ---------------------------------------------------------------------------------------
// Undamped psi of the interior cells within NG of an edge, for the ghost fill
// that follows (the reference damps the interior after its BC, so ghosts carry
// undamped psi, solver_cpu.cpp:435-439).  Layout: [0, 2NG*nx) bottom/top strips
// (row r: r < NG -> jb + r, else je - 2NG + r), then [2NG*nx, +2NG*ny) left/right
// strips (column c: c < NG -> ib + c, else ie - 2NG + c).
template <class D> MHD_DI void frame_put(D* __restrict__ fr, const Geom& g, int i, int j, double psi) {
    const int li = i - NG, lj = j - NG;
    if (lj < NG) fr[lj * g.nx + li] = psi;
    if (lj >= g.ny - NG) fr[(lj - (g.ny - 2 * NG)) * g.nx + li] = psi;
    D* side = fr + 2 * NG * g.nx;
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

// Wave-wide maximum of non-negative doubles into *acc (IEEE order of such values is
// their bit order).  After the butterfly every lane holds the same maximum and all
// lanes issue the atomic (combined by the hardware): no lane-divergent branch.
MHD_DI void wave_max_to(unsigned long long* acc, double m) {
    for (int off = 32; off > 0; off >>= 1) m = std_max(m, __shfl_xor(m, off));
    atomicMax(acc, (unsigned long long)__double_as_longlong(m));
}

// Knob (compile time): MHD_Y_WAVES_PER_EU asks the compiler for that many resident
// waves per SIMD (0: its own choice, 256 VGPRs and 2 waves for HLLD).
#ifndef MHD_Y_WAVES_PER_EU
#define MHD_Y_WAVES_PER_EU 0
#endif

#if MHD_Y_WAVES_PER_EU > 0
#define MHD_Y_ATTR __attribute__((amdgpu_waves_per_eu(MHD_Y_WAVES_PER_EU)))
#else
#define MHD_Y_ATTR
#endif

// Per-lane LDS of the column march (field-major, lane-consecutive, so accesses are
// conflict-free; each lane touches only its own slots, so no barrier is needed):
//   - the raw conserved states of the last three rows the lane loaded: the update of
//     row r-2 reads them back instead of reloading global memory, which would expose
//     a memory latency per row and double the sweep's reads;
//   - the primitive of the row above the cell being reconstructed (the slope's lower
//     neighbour), which is not needed while the face flux is formed: keeping it here
//     instead of in registers leaves room for the next row's load to be in flight.
// Each wave owns a region of YRING_DOUBLES x 64 doubles, so a field's lanes are 512 bytes
// apart whatever the block's width: every access of a slot is an immediate offset from one
// address (paired as ds_read2st64 / ds_write2st64).
constexpr int YRING_DOUBLES = 36;  // per lane: 3 rows + 1 primitive, 9 doubles each
struct YRing {
    double* b;  // this lane's first double: ring + wave * YRING_DOUBLES * 64 + lane
    MHD_DI static YRing of(double* ring, int t) {
        return YRing{ring + (t >> 6) * (YRING_DOUBLES * 64) + (t & 63)};
    }
    MHD_DI void putw(const Prim& w) const {
        double* q = b + 27 * 64;
        q[0] = w.r; q[64] = w.un; q[2 * 64] = w.ut; q[3 * 64] = w.uz; q[4 * 64] = w.bn;
        q[5 * 64] = w.bt; q[6 * 64] = w.bz; q[7 * 64] = w.p; q[8 * 64] = w.psi;
    }
    MHD_DI Prim getw() const {
        const double* q = b + 27 * 64;
        return Prim{q[0], q[64], q[2 * 64], q[3 * 64], q[4 * 64], q[5 * 64], q[6 * 64], q[7 * 64], q[8 * 64], 0.0};
    }
    MHD_DI void put(int slot, const Cons& u) const {
        double* q = b + slot * (9 * 64);
        q[0] = u.r; q[64] = u.mn; q[2 * 64] = u.mt; q[3 * 64] = u.mz; q[4 * 64] = u.bn;
        q[5 * 64] = u.bt; q[6 * 64] = u.bz; q[7 * 64] = u.e; q[8 * 64] = u.psi;
    }
    MHD_DI Cons get(int slot) const {
        const double* q = b + slot * (9 * 64);
        return Cons{q[0], q[64], q[2 * 64], q[3 * 64], q[4 * 64], q[5 * 64], q[6 * 64], q[7 * 64], q[8 * 64]};
    }
};

// Column i, rows [j0, j1): returns the largest CFL candidate of the cells written
// (0 when want_cfl is false or the lane is idle).
// Ut / Un are the bases of the source and destination grids: the restrict
// qualifiers let the next rows' loads run ahead of this row's stores.
// BUF: address both grids through buffer resources (each grid at most BUF_MAX_BYTES).
template <int SOLVER, int LIMITER, bool BUF = false>
MHD_DI double ysweep_column(const double* __restrict__ Ut, double* __restrict__ Un, const Geom& g, const Phys& ph,
                            double dtd, double hdtd, double damp, double* __restrict__ frame, bool want_cfl, int i,
                            bool active, int j0, int j1, const YRing& ring, bool prio = false) {
    using Src = std::conditional_t<BUF, BufPlanes<false>, Planes<false, GlobalT<const double>>>;
    using Dst = std::conditional_t<BUF, BufPlanes<false>, Planes<false, GlobalT<double>>>;
    const Src src = plane_view<BUF, Src>(Ut, g.ps);
    const Dst dst = plane_view<BUF, Dst>(Un, g.ps);
    const int ic = active ? i : NG;  // idle lanes shadow column NG (no stores)
    const bool damping = !(damp >= 1.0);  // solver_cpu.cpp:265 skips only when factor >= 1
    double cand = 0.0;
    ring.putw(to_prim(src.load((j0 - 2) * g.tnx + ic), ph));
    Prim w1 = to_prim(src.load((j0 - 1) * g.tnx + ic), ph);
    Cons hu;     // upper interface state of the previous cell
    Prim hw;     // and its primitive
    Cons fprev;  // face flux below the cell being updated
    int slot = 0;  // ring slot of row r; row r-2 is in slot (slot + 1) % 3
    // Row r+1 is loaded during iteration r, once the face flux is formed: the update,
    // its stores and the CFL candidate then run while the load is in flight (the last
    // iteration loads row j1+1 again).  One load site for every iteration (the first has
    // no face flux) keeps most of the loaded row in the registers the next iteration
    // reads: two sites left it to a copy at the end of every iteration, and 13 more SGPRs
    // spilled.
    Cons ur = src.load(j0 * g.tnx + ic);
    const Progress<MHD_PRIO_Y != 0> prog(j1 + 2 - j0, prio);
    for (int r = j0; r <= j1 + 1; ++r) {
        prog.at(r - j0);
        ring.put(slot, ur);
        const Prim w2 = to_prim(ur, ph);
        Face lo, hi;
        reconstruct<LIMITER>(ring.getw(), w1, w2, hdtd, ph, lo, hi);  // cell r-1
        ring.putw(w1);
        Cons f;  // formed from the second iteration on
        if (r >= j0 + 1) f = riemann<SOLVER, false>(hu, hw, lo.u, lo.w, ph);  // face (r-2 | r-1)
        ur = src.load((r < j1 + 1 ? r + 1 : j1 + 1) * g.tnx + ic);
        if (r >= j0 + 1) {
            if (r >= j0 + 2) {
                const int j = r - 2;
                const int k = j * g.tnx + ic;
                Cons n = update(ring.get(slot == 2 ? 0 : slot + 1), dtd, f, fprev);
                if (active) {
                    if (frame && (j < 2 * NG || j >= g.ny || i < 2 * NG || i >= g.nx)) frame_put(as_global(frame), g, i, j, n.psi);
                    if (damping) n.psi *= damp;
                    dst.store(k, n);
                    if (want_cfl) {
                        Cons x = n;  // back to the x frame for the CFL candidate
                        x.mn = n.mt; x.mt = n.mn; x.bn = n.bt; x.bt = n.bn;
                        cand = std_max(cand, cfl_candidate(x, ph));
                    }
                }
            }
            fprev = f;
        }
        hu = hi.u;
        hw = hi.w;
        w1 = w2;
        slot = slot == 2 ? 0 : slot + 1;
    }
    return cand;
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

// Ghost cell number t of a grid: the 2NG ghost columns of the ny interior rows
// (row-major), then the 2NG ghost rows of tnx cells.
MHD_DI int ghost_count(const Geom& g) { return 2 * NG * g.ny + 2 * NG * g.tnx; }

// Position (gi, gj) of ghost cell number t (0 <= t < ghost_count(g)).
MHD_DI void ghost_cell(const Geom& g, int t, int& gi, int& gj) {
    const int n_side = 2 * NG * g.ny;
    if (t < n_side) {
        // four consecutive lanes per row: columns 0, 1, ie, ie+1
        const int c = t % (2 * NG);
        gj = NG + t / (2 * NG);
        gi = c < NG ? c : g.nx + c;
    } else {
        const int q = t - n_side;
        const int r = q / g.tnx;
        gi = q % g.tnx;
        gj = r < NG ? r : g.ny + r;
    }
}

MHD_DI void ghost_fill(double* __restrict__ U, const double* __restrict__ prev, const Geom& g, const int* bc,
                       const double* __restrict__ frame, int t) {
    int gi, gj;
    ghost_cell(g, t, gi, gj);
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

}  // namespace mhd
