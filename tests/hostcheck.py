"""ctypes binding of libmhd_hostcheck.so: the host build of the adjoint arithmetic
(mhd_gpu_acc_amd/csrc/hostcheck.cpp), used only by the tests."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

SO = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mhd_gpu_acc_amd",
                  "libmhd_hostcheck.so")
_dp = C.POINTER(C.c_double)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            raise FileNotFoundError(f"{SO} is not built (make -C mhd_gpu_acc_amd/csrc)")
        L = C.CDLL(SO)
        L.hc_riemann_adj.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp]
        L.hc_reconstruct_adj.argtypes = [_dp, _dp, _dp, C.c_double, C.c_int, C.c_double, C.c_double, C.c_int,
                                         _dp, _dp, _dp, _dp, _dp]
        L.hc_flux_adj.argtypes = [_dp, C.c_int, C.c_double, C.c_double, _dp, _dp]
        L.hc_cons_to_prim_adj.argtypes = [_dp, C.c_double, _dp, _dp]
        L.hc_prim_to_cons_adj.argtypes = [_dp, C.c_double, _dp, _dp]
        L.hc_fast_speed_adj.argtypes = [_dp, C.c_int, C.c_double, C.c_double, _dp]
        L.hc_limiter_adj.argtypes = [C.c_double, C.c_double, C.c_int, C.c_double, _dp, _dp]
        L.hc_cfl_candidates.argtypes = [_dp, C.c_int, C.c_double, C.c_double, _dp, _dp, C.POINTER(C.c_int)]
        L.hc_adjoint_step.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                      C.POINTER(C.c_int), _dp]
        _ip = C.POINTER(C.c_int)
        L.hc_row_push.argtypes = [C.c_int, C.c_int, C.c_int, _ip]
        L.hc_ghost_source.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip]
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _c9(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(9))


def riemann_adj(UL, UR, direction, solver, ch, gamma, Fb):
    ULb, URb = np.zeros(9), np.zeros(9)
    lib().hc_riemann_adj(_p(_c9(UL)), _p(_c9(UR)), direction, solver, ch, gamma, _p(_c9(Fb)), _p(ULb), _p(URb))
    return ULb, URb


def reconstruct_adj(Wm, Wc, Wp, dtd, direction, ch, gamma, limiter, Lb, Rb):
    out = [np.zeros(9) for _ in range(3)]
    lib().hc_reconstruct_adj(_p(_c9(Wm)), _p(_c9(Wc)), _p(_c9(Wp)), dtd, direction, ch, gamma, limiter,
                             _p(_c9(Lb)), _p(_c9(Rb)), *[_p(o) for o in out])
    return out


def flux_adj(U, direction, ch, gamma, Fb):
    Ub = np.zeros(9)
    lib().hc_flux_adj(_p(_c9(U)), direction, ch, gamma, _p(_c9(Fb)), _p(Ub))
    return Ub


def cons_to_prim_adj(U, gamma, Vb):
    Ub = np.zeros(9)
    lib().hc_cons_to_prim_adj(_p(_c9(U)), gamma, _p(_c9(Vb)), _p(Ub))
    return Ub


def prim_to_cons_adj(V, gamma, Ub):
    Vb = np.zeros(9)
    lib().hc_prim_to_cons_adj(_p(_c9(V)), gamma, _p(_c9(Ub)), _p(Vb))
    return Vb


def fast_speed_adj(V, direction, gamma, cb):
    Vb = np.zeros(9)
    lib().hc_fast_speed_adj(_p(_c9(V)), direction, gamma, cb, _p(Vb))
    return Vb


def limiter_adj(a, b, limiter, sb):
    ab, bb = C.c_double(), C.c_double()
    lib().hc_limiter_adj(a, b, limiter, sb, C.byref(ab), C.byref(bb))
    return ab.value, bb.value


def adjoint_step(Uold, Utmp, case, dt, ch, solver, limiter, lam, cr=0.18, bc=None):
    """One reverse step over AoS host grids; returns dL/dUold given lam = dL/d(step output)."""
    out = np.ascontiguousarray(lam.copy())
    bcs = (C.c_int * 4)(*(bc or case.bc))
    rc = lib().hc_adjoint_step(_p(np.ascontiguousarray(Uold)), _p(np.ascontiguousarray(Utmp)), case.nx, case.ny,
                               case.x_min, case.x_max, case.y_min, case.y_max, dt, ch, cr, case.gamma, solver,
                               limiter, bcs, _p(out))
    if rc != 0:
        raise ValueError("unsupported solver, limiter or boundary for the adjoint")
    return out


def cfl_candidates(U, gamma, ch):
    """The kernels' compute_dt candidate of each cell (rows of U: conserved, reference
    order), the exact candidate, and whether the speed test cleared the cell."""
    U = np.ascontiguousarray(U, dtype=np.float64).reshape(-1, 9)
    n = U.shape[0]
    fast, exact = np.zeros(n), np.zeros(n)
    hit = np.zeros(n, dtype=np.int32)
    lib().hc_cfl_candidates(_p(U), n, gamma, ch, _p(fast), _p(exact), hit.ctypes.data_as(C.POINTER(C.c_int)))
    return fast, exact, hit.astype(bool)


def row_push(ny, bc_lo, bc_hi):
    """The ghost rows the x sweep writes itself: (n, src rows, dst rows, flip bits)."""
    out = np.zeros(10, dtype=np.int32)
    lib().hc_row_push(ny, bc_lo, bc_hi, out.ctypes.data_as(C.POINTER(C.c_int)))
    return int(out[0]), [int(v) for v in out[1:5]], [int(v) for v in out[5:9]], int(out[9])


def ghost_source(gi, gj, nx, ny, bc):
    """The closed-form source of ghost cell (gi, gj): (i, j, prev, flip_x, flip_y)."""
    b = np.ascontiguousarray(bc, dtype=np.int32)
    out = np.zeros(5, dtype=np.int32)
    lib().hc_ghost_source(gi, gj, nx, ny, b.ctypes.data_as(C.POINTER(C.c_int)), out.ctypes.data_as(C.POINTER(C.c_int)))
    return int(out[0]), int(out[1]), bool(out[2]), bool(out[3]), bool(out[4])
