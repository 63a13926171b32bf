"""mhd_advance_grids_noalloc (the entry point for callers that cannot allocate) against
mhd_advance_grids, through the C ABI: two calls in a row, the second on a smaller grid
with another solver, limiter and boundary mix, give the same U* and U_next bit for
bit, and the CFL maximum it returns is the c_h that mhd_compute_dt_grid ratchets to on
the same state.  A grid beyond the fixed scratch is refused."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import mhd_gpu_acc_amd as m
from mhd_gpu_acc_amd._lib import RunDesc, check, lib
from tests.util import assert_bitwise

pytestmark = pytest.mark.gpu


def _desc(cfg, solver, limiter):
    return RunDesc(cfg.nx, cfg.ny, 2, cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.gamma, 0.18, cfg.cfl,
                   solver, limiter, (C.c_int * 4)(*cfg.bc.as_list()), None)


def _one_call(L, nx, ny, bc, solver, limiter, seed):
    import torch
    from mhd_gpu_acc_amd.adjoint import from_planes, to_planes
    print(f"[gpu] noalloc {nx}x{ny} bc={bc} solver={solver} limiter={limiter}", flush=True)
    sys.stdout.flush()
    cfg = m.get_n_case_config("orszag_tang_unit", 1)
    cfg.nx, cfg.ny = nx, ny
    cfg.bc = m.BoundaryConfig(*[m.BoundaryType(b) for b in bc])
    U0 = m.initial_state(m.CaseId.OrszagTangUnit, cfg)
    U0 = U0 * (1.0 + 1e-3 * np.random.default_rng(seed).uniform(-1, 1, U0.shape))
    with m.Run(cfg, solver, limiter) as run:
        run.set_state(U0)
        dt = run.compute_dt()
        ch = run.ch
        Uold = to_planes(U0)
        Ta, Na = Uold.clone(), Uold.clone()
        Tb, Nb = Uold.clone(), Uold.clone()
        torch.cuda.synchronize()
        check(L.mhd_advance_grids(run.handle, C.c_void_p(Uold.data_ptr()), C.c_void_p(Ta.data_ptr()),
                                  C.c_void_p(Na.data_ptr()), dt))
        run.synchronize()
        d = _desc(cfg, solver, limiter)
        ms = C.c_double(-1.0)
        check(L.mhd_advance_grids_noalloc(C.byref(d), ch, C.c_void_p(Uold.data_ptr()), C.c_void_p(Tb.data_ptr()),
                                          C.c_void_p(Nb.data_ptr()), dt, C.byref(ms)))
        torch.cuda.synchronize()
        assert_bitwise(from_planes(Tb), from_planes(Ta), "U* (noalloc vs run)")
        assert_bitwise(from_planes(Nb), from_planes(Na), "U_next (noalloc vs run)")
        dt2 = C.c_double()
        check(L.mhd_compute_dt_grid(run.handle, C.c_void_p(Na.data_ptr()), C.byref(dt2)))
        assert ms.value == run.ch and ms.value >= ch, (ms.value, run.ch, ch)


def test_noalloc_matches_advance_grids_on_two_calls():
    L = lib()
    L.mhd_advance_grids_noalloc.argtypes = [C.POINTER(RunDesc), C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_double, C.POINTER(C.c_double)]
    _one_call(L, 300, 260, (0, 0, 0, 0), 2, 1, 1)
    _one_call(L, 41, 37, (3, 2, 1, 3), 1, 0, 2)  # smaller, Dirichlet / reflecting / transmissive sides


def test_noalloc_refuses_grids_beyond_its_scratch():
    L = lib()
    L.mhd_advance_grids_noalloc.argtypes = [C.POINTER(RunDesc), C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_double, C.POINTER(C.c_double)]
    cfg = m.get_n_case_config("orszag_tang_unit", 1)
    cfg.nx, cfg.ny = 32761, 8  # nx + ny above the scratch's limit; no grid is touched
    d = _desc(cfg, 2, 1)
    rc = L.mhd_advance_grids_noalloc(C.byref(d), 1.0, C.c_void_p(16), C.c_void_p(32), C.c_void_p(48), 1e-3, None)
    assert rc == 1 and b"MHD_NOALLOC_MAX_NXNY" in L.mhd_last_error()
