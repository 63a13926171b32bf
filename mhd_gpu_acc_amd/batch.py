"""Batched form: many independent runs advanced together (mhd_batch_* of the C ABI).

Each run has its own grid size, gamma, c_r, CFL and BCs; a batch uses one Riemann
solver and one limiter.  Every step is one hipGraph replay covering all runs, and
each run keeps its own c_h ratchet and dt, so run r of a batch is bitwise the run
``Run.step_to`` would produce for it alone.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import ICParams, RunDesc, check, lib
from .solver import BoundaryConfig, BoundaryType, CaseConfig, CaseId, Limiter, RiemannSolver, get_case_config, \
    initial_state, parse_case_id

__all__ = ["Batch", "plan_ot_family", "ot_family_batch"]


def _bind():
    L = lib()
    if getattr(L, "_batch_bound", False):
        return L
    L.mhd_batch_create.argtypes = [C.POINTER(RunDesc), C.c_int, C.POINTER(C.c_void_p)]
    L.mhd_batch_destroy.argtypes = [C.c_void_p]
    L.mhd_batch_set_state.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int]
    L.mhd_batch_init_states.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.mhd_batch_get_state.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int]
    L.mhd_batch_get_ch.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
    L.mhd_batch_set_ch.argtypes = [C.c_void_p, C.c_int, C.c_double]
    L.mhd_batch_step.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.mhd_batch_tune.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_size_t]
    L.mhd_batch_set_shape.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
    L.mhd_batch_plan_ot_family.argtypes = [C.c_uint64, C.c_int, C.POINTER(C.c_int), C.POINTER(ICParams)]
    L._batch_bound = True
    return L


class Batch:
    def __init__(self, cfgs, solver=RiemannSolver.HLLD, limiter=Limiter.VAN_LEER, cr=0.18, stream=None):
        L = _bind()
        self.cfgs = list(cfgs)
        n = len(self.cfgs)
        crs = cr if isinstance(cr, (list, tuple)) else [cr] * n
        descs = (RunDesc * n)()
        for r, c in enumerate(self.cfgs):
            descs[r] = RunDesc(c.nx, c.ny, c.ng, c.x_min, c.x_max, c.y_min, c.y_max, c.gamma, crs[r], c.cfl,
                               int(solver), int(limiter), (C.c_int * 4)(*c.bc.as_list()), stream)
        h = C.c_void_p()
        check(L.mhd_batch_create(descs, n, C.byref(h)))
        self._h = h
        self.n = n

    def close(self):
        if getattr(self, "_h", None):
            lib().mhd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_state(self, r: int, U: np.ndarray):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.shape != self.cfgs[r].shape:
            raise ValueError(f"state shape {U.shape} != {self.cfgs[r].shape}")
        check(lib().mhd_batch_set_state(self._h, r, U.ctypes.data_as(C.c_void_p), 0))

    def init_states(self, case, params=None):
        """Every run's analytic initial state on the device in one launch
        (mhd_batch_init_states): bitwise set_state(r, initial_state(case, cfgs[r],
        params[r])) for every run.  ``params``: one ICParams per run, or None."""
        arr = None
        if params is not None:
            if len(params) != len(self.cfgs):
                raise ValueError(f"{len(params)} parameter sets for {len(self.cfgs)} runs")
            arr = (ICParams * len(params))(*params)
        check(lib().mhd_batch_init_states(self._h, int(parse_case_id(case)), arr))

    def get_state(self, r: int) -> np.ndarray:
        U = np.empty(self.cfgs[r].shape)
        check(lib().mhd_batch_get_state(self._h, r, U.ctypes.data_as(C.c_void_p), 0))
        return U

    def ch(self, r: int) -> float:
        v = C.c_double()
        check(lib().mhd_batch_get_ch(self._h, r, C.byref(v)))
        return v.value

    def set_ch(self, r: int, value: float):
        check(lib().mhd_batch_set_ch(self._h, r, float(value)))

    def step(self, nsteps: int, record: bool = True):
        """Every run takes nsteps steps; returns the (nsteps, n) dt table (or None)."""
        dts = np.zeros((nsteps, self.n)) if record else None
        done = C.c_int()
        rc = lib().mhd_batch_step(self._h, int(nsteps),
                                  dts.ctypes.data_as(C.POINTER(C.c_double)) if record else None, C.byref(done))
        check(rc)
        return dts

    def tune(self, reps: int = 3) -> str:
        buf = C.create_string_buffer(8192)
        check(lib().mhd_batch_tune(self._h, int(reps), buf, len(buf)))
        return buf.value.decode()

    def set_shape(self, xwg=256, xgrid=8, yseg=32, ygrid=8):
        check(lib().mhd_batch_set_shape(self._h, xwg, xgrid, yseg, ygrid))


def plan_ot_family(seed: int, n: int = 64):
    """(sizes, [ICParams]) of the Orszag-Tang family batch (mhd_batch_plan_ot_family)."""
    L = _bind()
    sizes = (C.c_int * n)()
    params = (ICParams * n)()
    check(L.mhd_batch_plan_ot_family(C.c_uint64(seed), n, sizes, params))
    return list(sizes), [params[r] for r in range(n)]


def ot_family_batch(seed: int, n: int = 64, solver=RiemannSolver.HLLD, limiter=Limiter.VAN_LEER, scale: int = 1,
                    device_init: bool = False):
    """The 64-run batch: Orszag-Tang family ICs on [0,1]^2, periodic, CFL 0.8.
    ``scale`` divides the grid sizes (for small test batches).  With ``device_init`` the
    states are computed on the device (Batch.init_states) and ``states`` is None.
    Returns (batch, cfgs, params, states)."""
    sizes, params = plan_ot_family(seed, n)
    base = get_case_config(CaseId.OrszagTangFamily)
    cfgs, states = [], []
    for s, p in zip(sizes, params):
        c = CaseConfig(s // scale, s // scale, 2, base.x_min, base.x_max, base.y_min, base.y_max, base.cfl, base.t_end,
                       p.gamma, BoundaryConfig(*[BoundaryType.Periodic] * 4))
        cfgs.append(c)
        if not device_init:
            states.append(initial_state(CaseId.OrszagTangFamily, c, p))
    b = Batch(cfgs, solver, limiter)
    if device_init:
        b.init_states(CaseId.OrszagTangFamily, params)
        return b, cfgs, params, None
    for r, U in enumerate(states):
        b.set_state(r, U)
    return b, cfgs, params, states
