"""CPU checker for the GLM-MHD step: TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
import this package.  It loads

* ``liboracle.so`` -- this repository's C++17/OpenMP restatement of the
  reference CPU path (``mhd_oracle.cpp``), and
* ``_ref/libmhdref.so`` -- the reference CPU solver compiled from its own
  sources (present only where it was built; see ``Makefile``).

States are AoS float64 arrays of shape ``(ny + 2 ng, nx + 2 ng, 9)`` in the
reference field order (rho, rho*u, rho*v, rho*w, Bx, By, Bz, E, psi).
"""
from __future__ import annotations

import ctypes as C
import os
import time
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libmhdref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Geom(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("ng", C.c_int),
                ("x_min", C.c_double), ("x_max", C.c_double),
                ("y_min", C.c_double), ("y_max", C.c_double)]


class ICParams(C.Structure):
    _fields_ = [("gamma", C.c_double), ("b_scale", C.c_double), ("seed", C.c_uint64),
                ("n_modes", C.c_int), ("mode_amp", C.c_double * 4), ("mode_kx", C.c_double * 4),
                ("mode_ky", C.c_double * 4), ("mode_phase", C.c_double * 4),
                ("pert_amp", C.c_double), ("bump_amp", C.c_double), ("bump_sigma", C.c_double),
                ("bump_cx", C.c_double), ("bump_cy", C.c_double)]


def ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def ibuf(vals):
    return (C.c_int * len(vals))(*vals)


def _load(path: str):
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run `make -C oracle`)")
    return C.CDLL(path)


_orc = None
_ref = None


def orc():
    """The CPU restatement (always built by __graft_entry__.build())."""
    global _orc
    if _orc is None:
        lib = _load(ORACLE_SO)
        lib.orc_fast_speed.restype = C.c_double
        lib.orc_max_signal_speed.restype = C.c_double
        lib.orc_limiter.restype = C.c_double
        lib.orc_cell_uniform.restype = C.c_double
        lib.orc_splitmix64.restype = C.c_uint64
        lib.orc_limiter.argtypes = [C.c_double, C.c_double, C.c_int]
        lib.orc_cell_uniform.argtypes = [C.c_uint64, C.c_int, C.c_int]
        lib.orc_ic_sin.restype = C.c_double
        lib.orc_ic_sin.argtypes = [C.c_double]
        lib.orc_fast_speed.argtypes = [_dp, C.c_int, C.c_double]
        lib.orc_max_signal_speed.argtypes = [_dp, C.c_int, C.c_double, C.c_double]
        for name in ("orc_cons_to_prim", "orc_prim_to_cons"):
            getattr(lib, name).argtypes = [_dp, C.c_double, _dp]
        lib.orc_flux.argtypes = [_dp, C.c_int, C.c_double, C.c_double, _dp]
        lib.orc_solver_flux.argtypes = [C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_double, _dp]
        lib.orc_riemann_flux.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_double, C.c_double, _dp]
        lib.orc_apply_glm_flux.argtypes = [_dp, _dp, _dp, C.c_int, C.c_double]
        lib.orc_reconstruct.argtypes = [_dp, _dp, _dp, C.c_double, C.c_int, C.c_double, C.c_double,
                                        C.c_int, _dp, _dp]
        lib.orc_tan_cons_to_prim.argtypes = [_dp, _dp, C.c_double, _dp, _dp]
        lib.orc_tan_prim_to_cons.argtypes = [_dp, _dp, C.c_double, _dp, _dp]
        lib.orc_tan_flux.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_double, _dp, _dp]
        lib.orc_tan_fast_speed.argtypes = [_dp, _dp, C.c_int, C.c_double, _dp, _dp]
        lib.orc_tan_limiter.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _dp, _dp]
        lib.orc_tan_riemann.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp]
        lib.orc_tan_reconstruct.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_int, C.c_double,
                                            C.c_double, C.c_int, _dp, _dp, _dp, _dp]
        _orc = lib
    return _orc


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    """The compiled reference (only where /root/reference was mounted at build time)."""
    global _ref
    if _ref is None:
        lib = _load(REF_SO)
        for name in ("ref_fast_speed", "ref_max_signal_speed", "ref_minmod", "ref_get_ch",
                     "ref_get_gamma", "ref_grid_dx", "ref_grid_dy"):
            getattr(lib, name).restype = C.c_double
        lib.ref_set_gamma.argtypes = [C.c_double]
        lib.ref_set_ch.argtypes = [C.c_double]
        lib.ref_set_cr.argtypes = [C.c_double]
        lib.ref_minmod.argtypes = [C.c_double, C.c_double]
        lib.ref_fast_speed.argtypes = [_dp, C.c_int]
        lib.ref_max_signal_speed.argtypes = [_dp, C.c_int, C.c_double]
        lib.ref_flux.argtypes = [_dp, C.c_int, C.c_double, _dp]
        lib.ref_solver_flux.argtypes = [C.c_int, _dp, _dp, C.c_int, C.c_double, _dp]
        lib.ref_riemann_flux.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_double, _dp]
        lib.ref_apply_glm_flux.argtypes = [_dp, _dp, _dp, C.c_int, C.c_double]
        lib.ref_reconstruct.argtypes = [_dp, _dp, _dp, C.c_double, C.c_int, _dp, _dp]
        lib.ref_grid_new.restype = C.c_void_p
        lib.ref_grid_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                     C.c_double, C.c_double]
        lib.ref_make_n_grid.restype = C.c_void_p
        lib.ref_make_n_grid.argtypes = [C.c_int, C.c_int]
        for name in ("ref_grid_free", "ref_grid_get", "ref_grid_set", "ref_grid_dx", "ref_grid_dy"):
            getattr(lib, name).argtypes = [C.c_void_p] + ([_dp] if name in ("ref_grid_get", "ref_grid_set") else [])
        lib.ref_grid_copy.argtypes = [C.c_void_p, C.c_void_p]
        lib.ref_compute_dt.argtypes = [C.c_void_p, C.c_double, _dp]
        lib.ref_advance.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int, _ip]
        lib.ref_apply_boundary.argtypes = [C.c_void_p, _ip]
        lib.ref_copy_ghost_cells.argtypes = [C.c_void_p, C.c_void_p]
        lib.ref_case_config.argtypes = [C.c_int, C.c_int, _dp, _ip, _dp]
        lib.ref_initial_state_at.argtypes = [C.c_int, C.c_double, C.c_double, _dp]
        lib.ref_run_loop.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_int,
                                     _ip, C.c_int, _dp, _dp, _ip, _dp]
        _ref = lib
    return _ref


# ---------------------------------------------------------------------------------------
# Grid-level helpers over the oracle
# ---------------------------------------------------------------------------------------
@dataclass
class Case:
    nx: int
    ny: int
    ng: int
    x_min: float
    x_max: float
    y_min: float
    y_max: float
    cfl: float
    t_end: float
    gamma: float
    bc: tuple

    def geom(self) -> Geom:
        return Geom(self.nx, self.ny, self.ng, self.x_min, self.x_max, self.y_min, self.y_max)

    @property
    def shape(self):
        return (self.ny + 2 * self.ng, self.nx + 2 * self.ng, 9)


def case_config(case_id: int, n: int = 1) -> Case:
    cfg = np.zeros(10)
    bc = (C.c_int * 4)()
    if orc().orc_case_config(case_id, n, ptr(cfg), bc) != 0:
        raise ValueError(f"unknown case {case_id}")
    return Case(int(cfg[0]), int(cfg[1]), int(cfg[2]), *cfg[3:10].tolist(), tuple(bc))


def init_state(case_id: int, case: Case, params: ICParams | None = None) -> np.ndarray:
    U = np.zeros(case.shape)
    g = case.geom()
    rc = orc().orc_init_state(case_id, C.byref(g), C.byref(params) if params is not None else None, ptr(U))
    if rc != 0:
        raise ValueError(f"unknown IC {case_id}")
    return U


def compute_dt(U: np.ndarray, case: Case, cfl: float, ch: float):
    """Returns (status, dt, new c_h)."""
    g = case.geom()
    chv = C.c_double(ch)
    dt = C.c_double(0.0)
    rc = orc().orc_compute_dt(ptr(U), C.byref(g), C.c_double(cfl), C.c_double(case.gamma),
                              C.byref(chv), C.byref(dt))
    return rc, dt.value, chv.value


def advance(U: np.ndarray, case: Case, dt: float, ch: float, solver: int, limiter: int,
            bc=None, cr: float = 0.18):
    """One step; returns (Utmp, Unew)."""
    g = case.geom()
    Utmp = U.copy()
    Unew = U.copy()
    orc().orc_advance(ptr(U), ptr(Utmp), ptr(Unew), C.byref(g), C.c_double(dt), C.c_double(ch),
                      C.c_double(cr), C.c_double(case.gamma), solver, limiter, ibuf(bc or case.bc))
    return Utmp, Unew


def run_loop(U: np.ndarray, case: Case, solver: int, limiter: int, max_steps: int,
             t_end: float | None = None, cfl: float | None = None, ch0: float = 0.0,
             bc=None, cr: float = 0.18):
    """Driver loop (in place on a copy).  Returns dict(U, dts, chs, nsteps, t, code, ch)."""
    U = np.ascontiguousarray(U.copy())
    g = case.geom()
    dts = np.zeros(max(max_steps, 1))
    chs = np.zeros(max(max_steps, 1))
    ch = C.c_double(ch0)
    n = C.c_int(0)
    t = C.c_double(0.0)
    code = orc().orc_run_loop(ptr(U), C.byref(g), C.c_double(case.cfl if cfl is None else cfl),
                              C.c_double(case.t_end if t_end is None else t_end), max_steps,
                              C.c_double(case.gamma), C.c_double(cr), solver, limiter,
                              ibuf(bc or case.bc), C.byref(ch), ptr(dts), ptr(chs), C.byref(n), C.byref(t))
    k = n.value
    return dict(U=U, dts=dts[:k].copy(), chs=chs[:k].copy(), nsteps=k, t=t.value, code=code, ch=ch.value)


def run_loop_timed(U: np.ndarray, case: Case, solver: int, limiter: int, max_steps: int,
                   t_end: float | None = None, cr: float = 0.18):
    """run_loop with the wall time of the step loop only; returns (result dict, seconds)."""
    U = np.ascontiguousarray(U.copy())
    g = case.geom()
    dts = np.zeros(max(max_steps, 1))
    chs = np.zeros(max(max_steps, 1))
    ch = C.c_double(0.0)
    n = C.c_int(0)
    t = C.c_double(0.0)
    sec = C.c_double(0.0)
    code = orc().orc_run_loop_timed(ptr(U), C.byref(g), C.c_double(case.cfl),
                                    C.c_double(case.t_end if t_end is None else t_end), max_steps,
                                    C.c_double(case.gamma), C.c_double(cr), solver, limiter, ibuf(case.bc),
                                    C.byref(ch), ptr(dts), ptr(chs), C.byref(n), C.byref(t), C.byref(sec))
    k = n.value
    return dict(U=U, dts=dts[:k].copy(), chs=chs[:k].copy(), nsteps=k, t=t.value, code=code), sec.value


def ref_run_loop_timed(U: np.ndarray, case: Case, solver: int, max_steps: int,
                       t_end: float | None = None, cr: float = 0.18):
    """The compiled reference's own driver loop (ref_shim.cpp ref_run_loop: compute_dt +
    advance_second_order of src/cpu/solver_cpu.cpp, minmod slopes, its only limiter)
    from state U; returns (dict(U, dts, chs, nsteps, t, code), wall seconds of the
    loop alone)."""
    lib = ref()
    lib.ref_set_gamma(case.gamma)
    lib.ref_set_ch(0.0)
    lib.ref_set_cr(cr)
    U = np.ascontiguousarray(U, dtype=np.float64)
    h = [lib.ref_grid_new(case.nx, case.ny, case.ng, case.x_min, case.x_max, case.y_min, case.y_max)
         for _ in range(3)]
    try:
        for x in h:
            lib.ref_grid_set(x, ptr(U))
        dts = np.zeros(max(max_steps, 1))
        chs = np.zeros(max(max_steps, 1))
        ns = C.c_int()
        t = C.c_double()
        t0 = time.perf_counter()
        code = lib.ref_run_loop(h[0], h[1], h[2], case.cfl, case.t_end if t_end is None else t_end, solver,
                                ibuf(case.bc), max_steps, ptr(dts), ptr(chs), C.byref(ns), C.byref(t))
        sec = time.perf_counter() - t0
        Ur = np.zeros_like(U)
        lib.ref_grid_get(h[0], ptr(Ur))
    finally:
        for x in h:
            lib.ref_grid_free(x)
    k = ns.value
    return dict(U=Ur, dts=dts[:k].copy(), chs=chs[:k].copy(), nsteps=k, t=t.value, code=code), sec


def num_threads() -> int:
    return int(orc().orc_num_threads())


def tangent_run(U: np.ndarray, dU: np.ndarray, case: Case, dts, chs, solver: int, limiter: int,
                bc=None, cr: float = 0.18):
    """Forward-mode derivative through len(dts) steps with frozen dt/c_h; returns (U, dU)."""
    U = np.ascontiguousarray(U.copy())
    dU = np.ascontiguousarray(dU.copy())
    dts = np.ascontiguousarray(np.asarray(dts, dtype=np.float64))
    chs = np.ascontiguousarray(np.asarray(chs, dtype=np.float64))
    g = case.geom()
    orc().orc_tangent_run(ptr(U), ptr(dU), C.byref(g), len(dts), ptr(dts), ptr(chs), C.c_double(cr),
                          C.c_double(case.gamma), solver, limiter, ibuf(bc or case.bc))
    return U, dU


# ---------------------------------------------------------------------------------------
# Point tangents (forward mode, branches frozen at the primal values)
# ---------------------------------------------------------------------------------------
def _c9(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(9))


def tan_cons_to_prim(U, dU, gamma):
    V, dV = np.zeros(9), np.zeros(9)
    orc().orc_tan_cons_to_prim(ptr(_c9(U)), ptr(_c9(dU)), C.c_double(gamma), ptr(V), ptr(dV))
    return V, dV


def tan_prim_to_cons(V, dV, gamma):
    U, dU = np.zeros(9), np.zeros(9)
    orc().orc_tan_prim_to_cons(ptr(_c9(V)), ptr(_c9(dV)), C.c_double(gamma), ptr(U), ptr(dU))
    return U, dU


def tan_flux(U, dU, direction, ch, gamma):
    F, dF = np.zeros(9), np.zeros(9)
    orc().orc_tan_flux(ptr(_c9(U)), ptr(_c9(dU)), direction, C.c_double(ch), C.c_double(gamma), ptr(F), ptr(dF))
    return F, dF


def tan_fast_speed(V, dV, direction, gamma):
    c, dc = C.c_double(), C.c_double()
    orc().orc_tan_fast_speed(ptr(_c9(V)), ptr(_c9(dV)), direction, C.c_double(gamma), C.byref(c), C.byref(dc))
    return c.value, dc.value


def tan_limiter(a, da, b, db, limiter):
    v, dv = C.c_double(), C.c_double()
    orc().orc_tan_limiter(a, da, b, db, limiter, C.byref(v), C.byref(dv))
    return v.value, dv.value


def tan_riemann(UL, dUL, UR, dUR, direction, solver, ch, gamma):
    F, dF = np.zeros(9), np.zeros(9)
    orc().orc_tan_riemann(ptr(_c9(UL)), ptr(_c9(dUL)), ptr(_c9(UR)), ptr(_c9(dUR)), direction, solver,
                          C.c_double(ch), C.c_double(gamma), ptr(F), ptr(dF))
    return F, dF


def tan_reconstruct(Wm, dWm, Wc, dWc, Wp, dWp, dtd, direction, ch, gamma, limiter):
    L, dL, R, dR = (np.zeros(9) for _ in range(4))
    orc().orc_tan_reconstruct(ptr(_c9(Wm)), ptr(_c9(dWm)), ptr(_c9(Wc)), ptr(_c9(dWc)), ptr(_c9(Wp)),
                              ptr(_c9(dWp)), C.c_double(dtd), direction, C.c_double(ch), C.c_double(gamma),
                              limiter, ptr(L), ptr(dL), ptr(R), ptr(dR))
    return L, dL, R, dR


def frozen_fd(U: np.ndarray, probes, eps: float, case: Case, dts, chs, solver: int, limiter: int,
              bc=None, cr: float = 0.18):
    """Frozen-branch central-difference runs (orc_frozen_fd_run).  Returns (rho0, rho_pm):
    the final density of the unperturbed run, shape (tny, tnx), and of U +- eps e_q for
    each probe q = (row, column, field), shape (nprobe, 2, tny, tnx)."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    probes = [tuple(int(x) for x in q) for q in probes]
    pr = (C.c_int * (3 * len(probes)))(*[x for q in probes for x in q])
    dts = np.ascontiguousarray(np.asarray(dts, dtype=np.float64))
    chs = np.ascontiguousarray(np.asarray(chs, dtype=np.float64))
    tny, tnx = case.shape[:2]
    rho0 = np.zeros((tny, tnx))
    rho_pm = np.zeros((len(probes), 2, tny, tnx))
    g = case.geom()
    lib = orc()
    lib.orc_frozen_fd_run.argtypes = [_dp, _ip, C.c_int, C.c_double, C.POINTER(Geom), C.c_int, _dp, _dp, C.c_double,
                                      C.c_double, C.c_int, C.c_int, _ip, _dp, _dp]
    rc = lib.orc_frozen_fd_run(ptr(U), pr, len(probes), eps, C.byref(g), len(dts), ptr(dts), ptr(chs), cr,
                               case.gamma, solver, limiter, ibuf(bc or case.bc), ptr(rho0), ptr(rho_pm))
    if rc != 0:
        raise ValueError("at most 32 probes")
    return rho0, rho_pm
