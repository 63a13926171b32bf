"""ctypes binding of libmhd_hip.so (the C ABI declared in include/mhd_hip.h).

Loading fails loudly when the library is not built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# MHD_HIP_LIB: another build of the same library (a measurement knob, e.g. one built with
# EXTRA=-DMHD_FP_CONTRACT_FAST=1); the default is the in-tree build.
LIB_PATH = os.environ.get("MHD_HIP_LIB") or os.path.join(HERE, "libmhd_hip.so")

MHD_OK = 0
ERRORS = {1: "invalid argument", 2: "HIP runtime error", 3: "non-positive maximum wave speed",
          4: "dt non-finite or <= 0", 5: "dt below floor (blow-up)", 6: "no HIP device"}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class RunDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("ng", C.c_int),
                ("x_min", C.c_double), ("x_max", C.c_double), ("y_min", C.c_double), ("y_max", C.c_double),
                ("gamma", C.c_double), ("cr", C.c_double), ("cfl", C.c_double),
                ("solver", C.c_int), ("limiter", C.c_int), ("bc", C.c_int * 4), ("stream", C.c_void_p)]


class CaseCfg(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("ng", C.c_int),
                ("x_min", C.c_double), ("x_max", C.c_double), ("y_min", C.c_double), ("y_max", C.c_double),
                ("cfl", C.c_double), ("t_end", C.c_double), ("gamma", C.c_double), ("bc", C.c_int * 4)]


class ICParams(C.Structure):
    _fields_ = [("gamma", C.c_double), ("b_scale", C.c_double), ("seed", C.c_uint64),
                ("n_modes", C.c_int), ("mode_amp", C.c_double * 4), ("mode_kx", C.c_double * 4),
                ("mode_ky", C.c_double * 4), ("mode_phase", C.c_double * 4),
                ("pert_amp", C.c_double), ("bump_amp", C.c_double), ("bump_sigma", C.c_double),
                ("bump_cx", C.c_double), ("bump_cy", C.c_double)]


class MHDError(RuntimeError):
    """A non-OK status of the C ABI (the reference throws std::runtime_error)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}: {ERRORS.get(code, '?')}] {msg}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                          "(make -C mhd_gpu_acc_amd/csrc)")
    # One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64.so with
    # the same SONAME as /opt/rocm's.  Loading torch's first makes the library
    # bind to it, so device pointers and streams from torch are valid here (the
    # other order leaves torch without a usable device).
    import torch  # noqa: F401
    L = C.CDLL(LIB_PATH)
    L.mhd_last_error.restype = C.c_char_p
    L.mhd_plane_stride.restype = C.c_size_t
    L.mhd_plane_stride.argtypes = [C.c_void_p]
    L.mhd_state_ptr.restype = C.c_void_p
    L.mhd_state_ptr.argtypes = [C.c_void_p]
    L.mhd_run_create.argtypes = [C.POINTER(RunDesc), C.POINTER(C.c_void_p)]
    L.mhd_run_destroy.argtypes = [C.c_void_p]
    L.mhd_run_set_stream.argtypes = [C.c_void_p, C.c_void_p]
    L.mhd_set_state.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    L.mhd_get_state.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    L.mhd_set_ch.argtypes = [C.c_void_p, C.c_double]
    L.mhd_get_ch.argtypes = [C.c_void_p, _dp]
    L.mhd_compute_dt.argtypes = [C.c_void_p, _dp]
    L.mhd_advance.argtypes = [C.c_void_p, C.c_double]
    L.mhd_step_to.argtypes = [C.c_void_p, C.c_double, C.c_int, _dp, _ip, _dp, _dp]
    L.mhd_apply_boundary.argtypes = [C.c_void_p]
    L.mhd_compute_dt_grid.argtypes = [C.c_void_p, C.c_void_p, _dp]
    L.mhd_advance_grids.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]
    L.mhd_sync.argtypes = [C.c_void_p]
    L.mhd_case_config_get.argtypes = [C.c_int, C.c_int, C.POINTER(CaseCfg)]
    L.mhd_init_state_host.argtypes = [C.c_int, C.POINTER(CaseCfg), C.POINTER(ICParams), C.c_void_p]
    L.mhd_init_state_device.argtypes = [C.c_int, C.POINTER(CaseCfg), C.POINTER(ICParams), C.c_void_p, C.c_void_p]
    L.mhd_run_init_state.argtypes = [C.c_void_p, C.c_int, C.POINTER(ICParams)]
    L.mhd_device_count.argtypes = [_ip]
    L.mhd_profile.argtypes = [C.c_void_p, C.c_int]
    L.mhd_profile_read.argtypes = [C.c_void_p, _dp, _ip]
    L.mhd_probe_bandwidth.argtypes = [C.c_void_p, C.c_void_p, C.c_longlong, C.c_int, _dp]
    _lib = L
    return L


def check(rc: int):
    if rc != MHD_OK:
        raise MHDError(rc, lib().mhd_last_error().decode())
    return rc
