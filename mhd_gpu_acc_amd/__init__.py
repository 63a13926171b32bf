"""MI355X (gfx950) GLM-MHD MUSCL-Hancock step: hand-written HIP kernels behind a C ABI.

The compute path is ``libmhd_hip.so`` (mhd_gpu_acc_amd/csrc); this package is the
host-side mirror of the reference's solver interface (see ``solver``).
"""
from ._lib import MHDError, lib
from .solver import (BoundaryConfig, BoundaryType, CaseConfig, CaseId, ICParams, Limiter, RiemannSolver, Run,
                     get_case_config, get_n_case_config, init_state_device, initial_state, parse_case_id)

__all__ = ["MHDError", "lib", "BoundaryConfig", "BoundaryType", "CaseConfig", "CaseId", "ICParams", "Limiter",
           "RiemannSolver", "Run", "get_case_config", "get_n_case_config", "init_state_device", "initial_state",
           "parse_case_id"]
