"""Run driver and output formats (mhd_drive and the snapshot writers of the C ABI).

The time loop of the reference's drivers (scripts/cpu/main_cpu.cpp:176-276,
scripts/gpu/main_gpu.cu:198-280) on the GPU step: dt clamped to the next snapshot,
the dt guards, progress lines with max |B|, the snapshot CSVs
(main_cpu.cpp:36-91, byte for byte the reference writer's), the [TIMING] line, and
binary FP64 state dumps for restarts and exact comparisons.

    python -m mhd_gpu_acc_amd.driver orszag_tang --n 1 --solver hlld --out output

takes the reference CLI's arguments (main_cpu.cpp:103-131) plus ``--limiter``,
``--max-steps`` and ``--bin`` (also dump each snapshot as <prefix>.bin).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import ICParams, check, lib
from .solver import _CASE_NAMES, CaseId, Limiter, RiemannSolver, Run, get_n_case_config, initial_state, parse_case_id

DRIVE_DONE, DRIVE_BAD_DT, DRIVE_DT_FLOOR, DRIVE_MAX_STEPS = 0, 1, 2, 3
BIN_HEADER_BYTES = 128
_DEVICE_IC = {CaseId.ShockBubble, CaseId.BrioWu, CaseId.BrioWuTransmissive, CaseId.ShockCloud,
              CaseId.OrszagTangFamily}


class DriveOpts(C.Structure):
    _fields_ = [("case_name", C.c_char_p), ("platform", C.c_char_p), ("out_dir", C.c_char_p),
                ("log_path", C.c_char_p), ("n_scale", C.c_int), ("t_end", C.c_double), ("n_snapshots", C.c_int),
                ("snapshot_times", C.POINTER(C.c_double)), ("snapshot_tags", C.POINTER(C.c_char_p)),
                ("print_interval", C.c_double), ("max_steps", C.c_int), ("verbose", C.c_int),
                ("t_start", C.c_double), ("step_start", C.c_longlong)]


class DriveResult(C.Structure):
    _fields_ = [("t", C.c_double), ("steps", C.c_longlong), ("elapsed_s", C.c_double), ("last_dt", C.c_double),
                ("status", C.c_int)]


_bound = False


def _bind():
    global _bound
    L = lib()
    if not _bound:
        L.mhd_case_snapshots.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double), C.c_char_p, C.c_int, C.c_int,
                                         C.POINTER(C.c_int)]
        L.mhd_write_fields_csv.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p]
        L.mhd_write_fields_csv_host.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_char_p, C.c_char_p]
        L.mhd_write_state_bin.argtypes = [C.c_void_p, C.c_char_p, C.c_double, C.c_longlong]
        L.mhd_read_state_bin.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]
        L.mhd_drive.argtypes = [C.c_void_p, C.POINTER(DriveOpts), C.POINTER(DriveResult)]
        _bound = True
    return L


def case_snapshots(case, n_scale: int = 1):
    """(times, tags) of a case's snapshot schedule (CaseConfig::snapshot_times / _tags)."""
    L = _bind()
    cnt = C.c_int()
    check(L.mhd_case_snapshots(int(parse_case_id(case)), n_scale, None, None, 0, 0, C.byref(cnt)))
    n = cnt.value
    times = (C.c_double * max(n, 1))()
    tags = C.create_string_buffer(16 * max(n, 1))
    if n:
        check(L.mhd_case_snapshots(int(parse_case_id(case)), n_scale, times, tags, 16, n, C.byref(cnt)))
    return [times[k] for k in range(n)], [tags.raw[16 * k:16 * (k + 1)].split(b"\0")[0].decode() for k in range(n)]


def write_fields_csv(run: Run, out_dir: str, prefix: str):
    """write_all_fields (main_cpu.cpp:66-91) of the run's current state."""
    check(_bind().mhd_write_fields_csv(run.handle, out_dir.encode(), prefix.encode()))


def write_fields_csv_host(U: np.ndarray, gamma: float, out_dir: str, prefix: str):
    """write_all_fields of a host state in the reference layout ((ny+4, nx+4, 9))."""
    U = np.asarray(U, dtype=np.float64)
    soa = np.ascontiguousarray(np.moveaxis(U, -1, 0))
    ny, nx = U.shape[0] - 4, U.shape[1] - 4
    check(_bind().mhd_write_fields_csv_host(soa.ctypes.data_as(C.c_void_p), nx, ny, float(gamma), out_dir.encode(),
                                            prefix.encode()))


def write_state_bin(run: Run, path: str, t: float = 0.0, step: int = 0):
    check(_bind().mhd_write_state_bin(run.handle, path.encode(), float(t), int(step)))


def read_state_bin(run: Run, path: str):
    """Load a dump into the run (state and c_h); returns (t, step)."""
    t, step = C.c_double(), C.c_longlong()
    check(_bind().mhd_read_state_bin(run.handle, path.encode(), C.byref(t), C.byref(step)))
    return t.value, step.value


def load_bin(path: str):
    """Read a state dump without a run: (header dict, state (ny+4, nx+4, 9) float64)."""
    with open(path, "rb") as fh:
        raw = fh.read(BIN_HEADER_BYTES)
        if len(raw) != BIN_HEADER_BYTES or raw[:8] != b"MHDSTATE":
            raise ValueError(f"{path}: not an MHD state dump")
        version, hbytes, nx, ny, ng, nf = np.frombuffer(raw[8:32], dtype="<u4,<u4,<i4,<i4,<i4,<i4")[0]
        t, step, gamma, ch = np.frombuffer(raw[32:64], dtype="<f8,<i8,<f8,<f8")[0]
        layout = int(np.frombuffer(raw[64:68], dtype="<u4")[0])
        if version != 1 or hbytes != BIN_HEADER_BYTES or layout != 1 or nf != 9:
            raise ValueError(f"{path}: unsupported dump (version {version}, layout {layout})")
        tnx, tny = nx + 2 * ng, ny + 2 * ng
        planes = np.fromfile(fh, dtype="<f8", count=9 * tnx * tny)
    if planes.size != 9 * tnx * tny:
        raise ValueError(f"{path}: short file")
    head = dict(nx=int(nx), ny=int(ny), ng=int(ng), t=float(t), step=int(step), gamma=float(gamma), ch=float(ch))
    return head, np.moveaxis(planes.reshape(9, tny, tnx), 0, -1).copy()


def drive(run: Run, case_name: str, t_end: float, snapshots=((), ()), out_dir: str | None = None,
          log_path: str | None = None, print_interval: float = 0.1, n_scale: int = 1, platform: str = "gpu",
          max_steps: int = 0, verbose: bool = True, t_start: float = 0.0, step_start: int = 0) -> dict:
    """mhd_drive: the reference time loop on the run from its current state and c_h
    (a restart passes the t and step of the state it loaded)."""
    times, tags = snapshots
    n = len(times)
    ct = (C.c_double * max(n, 1))(*times)
    tg = (C.c_char_p * max(n, 1))(*[t.encode() for t in tags])
    o = DriveOpts(case_name.encode(), platform.encode(), out_dir.encode() if out_dir else None,
                  log_path.encode() if log_path else None, n_scale, float(t_end), n, ct, tg, float(print_interval),
                  int(max_steps), int(bool(verbose)), float(t_start), int(step_start))
    res = DriveResult()
    check(_bind().mhd_drive(run.handle, C.byref(o), C.byref(res)))
    return dict(t=res.t, steps=res.steps, elapsed_s=res.elapsed_s, last_dt=res.last_dt, status=res.status)


def run_case(case, n_scale: int = 1, solver=RiemannSolver.HLLD, limiter=Limiter.MINMOD, out_dir: str | None = "output",
             log_path: str | None = None, params: ICParams | None = None, print_interval: float = 0.1,
             max_steps: int = 0, verbose: bool = True, bin_dumps: bool = False) -> dict:
    """The reference driver's run of a case (main_cpu.cpp:136-276) on the GPU: the case
    table, the analytic initial state (on the device where the case has one), the
    snapshot schedule and the time loop."""
    cid = parse_case_id(case)
    name = [k for k, v in _CASE_NAMES.items() if v == int(cid)][0]
    cfg = get_n_case_config(cid, n_scale)
    snaps = case_snapshots(cid, n_scale)
    with Run(cfg, solver, limiter) as run:
        if cid in _DEVICE_IC and (cid != CaseId.OrszagTangFamily or params is not None):
            run.init_state(cid, params)
        else:
            run.set_state(initial_state(cid, cfg, params))
        run.ch = 0.0
        res = drive(run, name, cfg.t_end, snaps, out_dir, log_path, print_interval, n_scale, "gpu", max_steps,
                    verbose)
        if bin_dumps and out_dir:
            write_state_bin(run, os.path.join(out_dir, f"{name}_gpu_final.bin"), res["t"], res["steps"])
    return res


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="MHD GLM solver on the GPU (the reference driver's loop)")
    ap.add_argument("case_name", nargs="?", default="kelvin_helmholtz")
    ap.add_argument("--n", type=int, default=1)
    ap.add_argument("--solver", default="hlld", choices=["hll", "hllc", "hlld", "force"])
    ap.add_argument("--limiter", default="minmod", choices=["minmod", "vanleer"])
    ap.add_argument("--out", default="output")
    ap.add_argument("--no-out", action="store_true")
    ap.add_argument("--max-steps", type=int, default=0)
    ap.add_argument("--bin", action="store_true", help="also dump the final state as <case>_gpu_final.bin")
    a = ap.parse_args(argv)
    if a.n < 1:
        raise SystemExit("--n must be >= 1")
    solver = {"hll": RiemannSolver.HLL, "hllc": RiemannSolver.HLLC, "hlld": RiemannSolver.HLLD,
              "force": RiemannSolver.FORCE}[a.solver]
    limiter = Limiter.MINMOD if a.limiter == "minmod" else Limiter.VAN_LEER
    res = run_case(a.case_name, a.n, solver, limiter, None if a.no_out else a.out, max_steps=a.max_steps,
                   bin_dumps=a.bin)
    return 0 if res["status"] == DRIVE_DONE else 1


if __name__ == "__main__":
    raise SystemExit(main())
