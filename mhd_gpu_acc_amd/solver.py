"""Python mirror of the reference solver interface over the HIP C ABI.

Names follow the reference (head/riemann.hpp, head/test_cases.hpp,
head/gpu/solver_gpu.cuh, head/init.hpp): ``RiemannSolver``, ``BoundaryType``,
``BoundaryConfig``, ``CaseId``, ``CaseConfig``, ``get_case_config`` /
``get_n_case_config``, ``initial_state``; and a ``Run`` object that owns the
device grids (Grid2DGPU x 3 + GpuWorkspace) with ``compute_dt`` /
``advance_second_order`` / ``step_to``.  Errors raise ``MHDError`` (a
``RuntimeError``), as the reference CPU path throws ``std::runtime_error``.

States crossing this interface are NumPy float64 arrays of shape
``(ny + 4, nx + 4, 9)`` (AoS, reference field order, ghosts included).
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field

import numpy as np

from ._lib import CaseCfg, ICParams, MHDError, RunDesc, check, lib

__all__ = ["RiemannSolver", "BoundaryType", "BoundaryConfig", "Limiter", "CaseId", "CaseConfig",
           "get_case_config", "get_n_case_config", "initial_state", "Run", "MHDError", "ICParams"]


class RiemannSolver(enum.IntEnum):   # head/riemann.hpp:23
    HLL = 0
    HLLC = 1
    HLLD = 2
    FORCE = 3


class BoundaryType(enum.IntEnum):    # head/test_cases.hpp:11-16
    Periodic = 0
    Transmissive = 1
    Reflecting = 2
    Dirichlet = 3


class Limiter(enum.IntEnum):
    MINMOD = 0    # src/cpu/solver_cpu.cpp:26-43
    VAN_LEER = 1  # not in the reference


class CaseId(enum.IntEnum):          # head/test_cases.hpp:29-34, then the added cases
    KelvinHelmholtz = 0
    ShockBubble = 1
    BrioWu = 2
    OrszagTang = 3
    OrszagTangUnit = 4
    ShockCloud = 5
    BrioWuTransmissive = 6
    OrszagTangFamily = 7
    OrszagTangBump = 8


_CASE_NAMES = {"kelvin_helmholtz": 0, "shock_bubble": 1, "brio_wu": 2, "orszag_tang": 3,
               "orszag_tang_unit": 4, "shock_cloud": 5, "brio_wu_transmissive": 6,
               "orszag_tang_family": 7, "orszag_tang_bump": 8}


def parse_case_id(name) -> CaseId:
    """parse_case_id (test_cases.cpp:465-475); raises ValueError on unknown names."""
    if isinstance(name, (int, CaseId)):
        return CaseId(int(name))
    if name not in _CASE_NAMES:
        raise ValueError(f"Unknown MHD test case: '{name}'. Valid names: {', '.join(_CASE_NAMES)}.")
    return CaseId(_CASE_NAMES[name])


@dataclass
class BoundaryConfig:                # head/test_cases.hpp:18-23 (default: all transmissive)
    left: BoundaryType = BoundaryType.Transmissive
    right: BoundaryType = BoundaryType.Transmissive
    bottom: BoundaryType = BoundaryType.Transmissive
    top: BoundaryType = BoundaryType.Transmissive

    def as_list(self):
        return [int(self.left), int(self.right), int(self.bottom), int(self.top)]


@dataclass
class CaseConfig:                    # head/test_cases.hpp:38-55
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
    bc: BoundaryConfig = field(default_factory=BoundaryConfig)

    @property
    def shape(self):
        return (self.ny + 2 * self.ng, self.nx + 2 * self.ng, 9)

    def _c(self) -> CaseCfg:
        return CaseCfg(self.nx, self.ny, self.ng, self.x_min, self.x_max, self.y_min, self.y_max,
                       self.cfl, self.t_end, self.gamma, (C.c_int * 4)(*self.bc.as_list()))


def get_n_case_config(case, n: int = 1) -> CaseConfig:
    """get_n_case_config (test_cases.cpp:275-282): the case table, nx and ny scaled by n."""
    if n < 1:
        raise ValueError("get_n_case_config: n must be >= 1.")
    c = CaseCfg()
    check(lib().mhd_case_config_get(int(parse_case_id(case)), n, C.byref(c)))
    return CaseConfig(c.nx, c.ny, c.ng, c.x_min, c.x_max, c.y_min, c.y_max, c.cfl, c.t_end, c.gamma,
                      BoundaryConfig(*[BoundaryType(b) for b in c.bc]))


def get_case_config(case) -> CaseConfig:
    return get_n_case_config(case, 1)


def init_state_device(case, cfg: CaseConfig, params: ICParams | None, ptr: int, stream: int = 0):
    """The initial state computed on the device into a caller's dense SoA grid at device
    address ``ptr`` (9 planes of (nx+4)(ny+4) doubles), asynchronously on ``stream``:
    bitwise ``initial_state`` for the cases that have a device state."""
    cc = cfg._c()
    check(lib().mhd_init_state_device(int(parse_case_id(case)), C.byref(cc),
                                      C.byref(params) if params is not None else None, C.c_void_p(ptr),
                                      C.c_void_p(stream) if stream else None))


def initial_state(case, cfg: CaseConfig, params: ICParams | None = None) -> np.ndarray:
    """initialise_grid (init.cpp:34-43): every cell, ghosts included, from the analytic IC."""
    U = np.zeros(cfg.shape)
    cc = cfg._c()
    check(lib().mhd_init_state_host(int(parse_case_id(case)), C.byref(cc),
                                    C.byref(params) if params is not None else None,
                                    U.ctypes.data_as(C.c_void_p)))
    return U


class Run:
    """Device grids and workspace of one run (Grid2DGPU x 3 + GpuWorkspace).

    ``solver``, ``limiter``, ``bc``, ``gamma``, ``cr`` and ``cfl`` are per run (the
    reference keeps gamma and c_h in process globals, head/physics.hpp:26-43).
    ``stream`` is a hipStream_t as an int (e.g. torch.cuda.current_stream().cuda_stream).
    """

    def __init__(self, cfg: CaseConfig, solver=RiemannSolver.HLLD, limiter=Limiter.MINMOD,
                 cr: float = 0.18, stream: int | None = None, bc: BoundaryConfig | None = None):
        self.cfg = cfg
        bcl = (bc or cfg.bc).as_list()
        d = RunDesc(cfg.nx, cfg.ny, cfg.ng, cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.gamma, cr,
                    cfg.cfl, int(solver), int(limiter), (C.c_int * 4)(*bcl), stream)
        h = C.c_void_p()
        check(lib().mhd_run_create(C.byref(d), C.byref(h)))
        self._h = h
        self._stream = int(stream or 0)

    def close(self):
        if getattr(self, "_h", None):
            for t in list(getattr(self, "_tapes", ())):  # adjoint tapes hold the run's buffers
                t.close()
            lib().mhd_run_destroy(self._h)
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

    @property
    def handle(self):
        return self._h

    @property
    def stream(self) -> int:
        """hipStream_t (as an int, 0 = the null stream) the run's work is queued on."""
        return self._stream

    def set_stream(self, stream: int):
        """Queue the run's work on ``stream`` from now on (mhd_run_set_stream): the new
        stream first waits, on the device, for what is already queued on the old one."""
        stream = int(stream or 0)
        check(lib().mhd_run_set_stream(self._h, C.c_void_p(stream) if stream else None))
        self._stream = stream

    @property
    def plane_stride(self) -> int:
        return int(lib().mhd_plane_stride(self._h))

    @property
    def state_ptr(self) -> int:
        """Device address of the current state (9 SoA planes, ``plane_stride`` apart)."""
        return int(lib().mhd_state_ptr(self._h))

    def set_state(self, U: np.ndarray):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.shape != self.cfg.shape:
            raise ValueError(f"state shape {U.shape} != {self.cfg.shape}")
        check(lib().mhd_set_state(self._h, U.ctypes.data_as(C.c_void_p), 0))

    def get_state(self) -> np.ndarray:
        U = np.empty(self.cfg.shape)
        check(lib().mhd_get_state(self._h, U.ctypes.data_as(C.c_void_p), 0))
        return U

    def init_state(self, case, params: ICParams | None = None):
        """The analytic initial state computed on the device (mhd_run_init_state): bitwise
        ``set_state(initial_state(case, cfg, params))`` for the cases that have one
        (Brio-Wu, shock-bubble, shock-cloud, the batch Orszag-Tang family)."""
        check(lib().mhd_run_init_state(self._h, int(parse_case_id(case)),
                                       C.byref(params) if params is not None else None))

    def set_state_device(self, ptr: int):
        """Copy from a dense device SoA buffer (9 x cells) at address ``ptr``."""
        check(lib().mhd_set_state(self._h, C.c_void_p(ptr), 2))

    def get_state_device(self, ptr: int):
        check(lib().mhd_get_state(self._h, C.c_void_p(ptr), 2))

    @property
    def ch(self) -> float:
        v = C.c_double()
        check(lib().mhd_get_ch(self._h, C.byref(v)))
        return v.value

    @ch.setter
    def ch(self, value: float):
        check(lib().mhd_set_ch(self._h, float(value)))

    def compute_dt(self) -> float:
        """compute_dt (solver_cpu.cpp:283-312): ratchets c_h, returns cfl min(dx,dy)/c_h."""
        v = C.c_double()
        check(lib().mhd_compute_dt(self._h, C.byref(v)))
        return v.value

    def advance_second_order(self, dt: float):
        """advance_second_order (solver_cpu.cpp:327-440); asynchronous."""
        check(lib().mhd_advance(self._h, float(dt)))

    advance = advance_second_order

    def apply_boundary(self):
        check(lib().mhd_apply_boundary(self._h))

    def synchronize(self):
        check(lib().mhd_sync(self._h))

    def step_to(self, t_end: float, max_steps: int, record: bool = True):
        """Driver loop of main_cpu.cpp:196-222.  Returns dict(t, steps, code, dts, chs);
        ``code`` is 0 or the stop status (3 no signal, 4 bad dt, 5 dt floor)."""
        n = max(int(max_steps), 1)
        dts = np.zeros(n) if record else None
        chs = np.zeros(n) if record else None
        t = C.c_double()
        s = C.c_int()
        dp = C.POINTER(C.c_double)
        rc = lib().mhd_step_to(self._h, float(t_end), int(max_steps), C.byref(t), C.byref(s),
                               dts.ctypes.data_as(dp) if record else None,
                               chs.ctypes.data_as(dp) if record else None)
        if rc not in (0, 3, 4, 5):
            check(rc)
        k = s.value
        return dict(t=t.value, steps=k, code=rc, dts=dts[:k].copy() if record else None,
                    chs=chs[:k].copy() if record else None)
