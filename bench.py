"""Benchmark of the plain FP64 step: Orszag-Tang 4096^2 (unit box), HLLD, van Leer.

One step = compute_dt (device max-reduction fused into the previous y sweep, 8-byte
read-back, host dt/c_h arithmetic) + advance (x sweep, ghost fill, y sweep with fused
psi damping and CFL candidate, ghost fill).  The state is resident in HBM before the
timed region starts.  Prints ONE JSON line.

    python bench.py [--gpus 1] [--steps 100] [--warmup 5] [--n 4096] [--solver hlld]
                    [--limiter vanleer] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
BYTES_PER_CELL_SWEEP = 144  # 9 FP64 read + 9 FP64 written per cell per directional sweep
BYTES_PER_CELL_STEP = 288   # two sweeps
SOLVERS = {"hll": 0, "hllc": 1, "hlld": 2, "force": 3}
LIMITERS = {"minmod": 0, "vanleer": 1}
# Intervals between the HIP events a profiled step records: the x sweep, the gap to the y sweep
# (holding U*'s ghost fill only where U* is read afterwards: a run's own steps skip it, the x
# sweep writing the ghost rows the y sweep reads), the y sweep, the closing ghost fill.
KERNELS = ("xsweep", "bc_after_x", "ysweep", "bc_after_y")
# VALU lane-instructions per cell of the plain sweeps (rocprofv3 SQ_INSTS_VALU x 64 / cells,
# profiles/valu_sweeps_hlld_vanleer.md) and the MI355X's VALU issue peak: 1024 SIMDs of 16
# FP64 lanes, one wave64 instruction per 4 cycles, at the 2.4 GHz peak clock.
VALU_PER_CELL = {("hlld", "vanleer"): {"ysweep": 1264, "xsweep": 1145}}
VALU_PEAK = 1024 * 16 * 2.4e9  # lane-instructions/s


def _progress(msg):
    """One line on stderr per finished leg, so a long run shows where it is."""
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def read_traffic(n, solver, limiter):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled
    by a copy-kernel calibration), from profiles/pmc_*.txt when present."""
    path = os.path.join(ROOT, "profiles", f"pmc_ot{n}_{solver}_{limiter}.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.startswith("#"):
            k, v = line.strip().split("=", 1)
            vals[k.strip()] = v.strip()
    return vals


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--solver", default="hlld", choices=sorted(SOLVERS))
    ap.add_argument("--limiter", default="vanleer", choices=sorted(LIMITERS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--config-cpu-seconds", type=float, default=2.0,
                    help="CPU sample per configuration for its CPU figure and parity columns")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-form measurement")
    ap.add_argument("--batch-steps", type=int, default=40)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint measurement")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other plain configurations (all solvers at 4096^2 and 1024^2, a 4096x512 "
                         "slab, shock-cloud, Brio-Wu)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="before the timed steps, copy BYTES between two device buffers once (a kernel "
                         "of known traffic for calibrating rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU; the path shards by y-slabs but no multi-GPU "
              "run is part of it (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import mhd_gpu_acc_amd as m

    torch.cuda.init()
    n = args.n
    solver, limiter = SOLVERS[args.solver], LIMITERS[args.limiter]
    cfg = m.get_n_case_config("orszag_tang_unit", 1)
    cfg.nx = cfg.ny = n
    U0 = m.initial_state(m.CaseId.OrszagTangUnit, cfg)
    cells = n * n
    run = m.Run(cfg, solver, limiter)
    run.set_state(U0)
    L = m.lib()
    import ctypes as C

    if args.copy_probe > 0:
        src = torch.ones(args.copy_probe // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # warm-up (also the first steps of the parity sample)
    w = run.step_to(cfg.t_end, args.warmup)
    assert w["code"] == 0, w
    L.mhd_profile(run.handle, 1)
    torch.cuda.synchronize()
    run.synchronize()
    t0 = time.perf_counter()
    r = run.step_to(cfg.t_end, args.steps)
    run.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    assert r["code"] == 0 and r["steps"] == args.steps, r
    ms4 = (C.c_double * 4)()
    nst = C.c_int()
    L.mhd_profile_read(run.handle, ms4, C.byref(nst))
    L.mhd_profile(run.handle, 0)
    elapsed = t1 - t0
    value = args.steps * cells / elapsed
    kern_ms = {k: ms4[i] / max(nst.value, 1) for i, k in enumerate(KERNELS)}
    dom = max(("xsweep", "ysweep"), key=lambda k: kern_ms[k])
    achieved = BYTES_PER_CELL_SWEEP * cells / (kern_ms[dom] * 1e-3)
    traffic = read_traffic(n, args.solver, args.limiter)
    gpu_dts = np.concatenate([w["dts"], r["dts"]])
    run.close()

    # The HBM bandwidth streaming kernels of known traffic reach on this device, in this run:
    # the practical ceiling the sweeps' bytes per second are read against (a sweep reads and
    # writes the same number of bytes, as the copy does).
    probe_bytes = 1 << 31
    pa = torch.zeros(probe_bytes // 8, dtype=torch.float64, device="cuda")
    pb = torch.empty_like(pa)
    torch.cuda.synchronize()
    tbps = (C.c_double * 3)()
    from mhd_gpu_acc_amd._lib import check
    check(L.mhd_probe_bandwidth(pa.data_ptr(), pb.data_ptr(), probe_bytes, 10, tbps))
    del pa, pb
    hbm_probe = {"copy": tbps[0], "read": tbps[1], "write": tbps[2], "unit": "TB/s", "bytes": probe_bytes,
                 "kernels": "16-byte lanes, 4 workgroups of 256 per CU, grid-stride; copy moves 2 x bytes "
                            "(mhd_probe_bandwidth)"}

    out = {
        "metric": "cell-updates/s + %HBM peak, OT 4096² HLLD FP64; batched & adjoint vs plain",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: analytic Orszag-Tang initial state (no dataset)",
        "config": {"workload": f"orszag_tang_unit {n}x{n} periodic, {args.solver.upper()}, {args.limiter}, "
                               f"CFL {cfg.cfl}, plain step (compute_dt + advance)",
                   "nx": n, "ny": n, "solver": args.solver, "limiter": args.limiter, "cfl": cfg.cfl},
        "hbm_frac_step": value * BYTES_PER_CELL_STEP / HBM_PEAK,
        "copy_probe_bytes": args.copy_probe,
        "kernel_ms": kern_ms,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": (float(traffic["bytes_per_launch_" + dom]) if traffic and
                                 ("bytes_per_launch_" + dom) in traffic else None),
                     "algorithmic_bytes_per_launch": BYTES_PER_CELL_SWEEP * cells,
                     "traffic_source": (f"profiles/pmc_ot{n}_{args.solver}_{args.limiter}.txt (rocprofv3 --pmc "
                                        "FETCH_SIZE / WRITE_SIZE runs of this kernel, copy-calibrated), not "
                                        "measured in this run") if traffic else None,
                     "frac_of_measured_copy": achieved / (tbps[0] * 1e12)},
        "hbm_probe": hbm_probe,
    }
    per_cell = VALU_PER_CELL.get((args.solver, args.limiter), {}).get(dom)
    if per_cell:
        # The sweeps are bound by FP64 VALU issue, not by HBM: the same kernel against the
        # VALU issue peak (instruction count from the committed counters, time from this run).
        rate = per_cell * cells / (kern_ms[dom] * 1e-3)
        out["valu_roofline"] = {"bound": "valu", "kernel": dom, "lane_instructions_per_cell": per_cell,
                                "achieved": rate / 1e12, "peak": VALU_PEAK / 1e12, "unit": "T lane-instructions/s",
                                "frac": rate / VALU_PEAK,
                                "source": "profiles/valu_sweeps_hlld_vanleer.md (rocprofv3 SQ_INSTS_VALU x 64 / "
                                          "cells); peak at 2.4 GHz, FP64 rcp/rsq (31 per cell) counted as one "
                                          "issue each though they take four"}

    _progress(f"plain {n}^2 {args.solver} {args.limiter}: {out['ms_per_step']:.3f} ms/step")
    if not args.no_batch:
        out["batched"] = bench_batch(args.batch_steps)
        _progress("batched form done")
    if not args.no_adjoint:
        out["adjoint"] = bench_adjoint()
        _progress("adjoint form done")
    if not args.no_configs:
        out["configs"] = bench_configs(0.0 if args.no_cpu else args.config_cpu_seconds)
        _progress("configurations done")

    if not args.no_cpu:
        import oracle
        case = oracle.Case(n, n, 2, cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.cfl, cfg.t_end, cfg.gamma,
                           tuple(cfg.bc.as_list()))
        _, t_one = oracle.run_loop_timed(U0, case, solver, limiter, 1)
        k = int(max(1, min(40, round(args.cpu_seconds / max(t_one, 1e-9)))))
        res, secs = oracle.run_loop_timed(U0, case, solver, limiter, k)
        kk = min(k, len(gpu_dts))
        same = bool(np.array_equal(res["dts"][:kk].view(np.uint64), gpu_dts[:kk].view(np.uint64)))
        out["cpu_baseline"] = {"value": k * cells / secs, "unit": "cell-updates/s",
                               "cores": oracle.num_threads(), "kind": "port",
                               "sample": f"first {k} steps of the same {n}^2 run ({secs:.1f} s), "
                                         f"C++/OpenMP restatement of the reference CPU path"}
        out["parity"] = {"dt_sequence_bitwise_vs_cpu": same, "steps_compared": kk}
        if oracle.ref_available():
            # The reference's own CPU solver (compiled from its sources into oracle/_ref) on
            # the same grid.  It has no van Leer limiter, so this sample uses its minmod
            # slopes; the port runs the same minmod sample for comparison.
            kr = int(max(1, min(40, round(0.5 * args.cpu_seconds / max(t_one, 1e-9)))))
            rref, sref = oracle.ref_run_loop_timed(U0, case, solver, kr)
            rprt, sprt = oracle.run_loop_timed(U0, case, solver, 0, kr)
            kq = min(rref["nsteps"], rprt["nsteps"])
            out["cpu_reference"] = {
                "value": kr * cells / sref, "unit": "cell-updates/s", "cores": oracle.num_threads(),
                "kind": "reference",
                "sample": f"first {kr} steps of the same {n}^2 Orszag-Tang run with {args.solver.upper()} and "
                          f"minmod slopes (the reference has no van Leer limiter), {sref:.1f} s: the reference "
                          f"CPU solver compiled from its own sources",
                "port_same_sample": kr * cells / sprt,
                "dt_sequence_bitwise_vs_port": bool(np.array_equal(rref["dts"][:kq].view(np.uint64),
                                                                   rprt["dts"][:kq].view(np.uint64))),
                "state_bitwise_vs_port": bool(np.array_equal(rref["U"].view(np.uint64),
                                                             rprt["U"].view(np.uint64)))}
    print(json.dumps(out), flush=True)


_IC_CACHE = {}


def _time_plain(case_name, nx, ny, solver, limiter, warmup, steps, t_end=None, params=None, cfl=None,
                cpu_seconds=0.0, y_max=None):
    """Plain step of one configuration: warm-up, then `steps` timed steps (or the run to
    t_end when steps is None), with the sweep kernels timed by HIP events.  A run the
    scheme cannot finish (the reference's blow-up stop, dt below 1e-8 t_end) reports
    its stop code and the time reached, as the reference driver does.  y_max replaces the
    case's upper y bound (a slab of the case's domain)."""
    import ctypes as C
    import time

    import mhd_gpu_acc_amd as m

    cfg = m.get_n_case_config(case_name, 1)
    cfg.nx, cfg.ny = nx, ny
    if case_name == "brio_wu_transmissive":
        cfg.cfl = 0.8
    if cfl is not None:
        cfg.cfl = cfl
    if y_max is not None:
        cfg.y_max = y_max
    key = (case_name, nx, ny, cfg.y_max)
    if key not in _IC_CACHE:
        _IC_CACHE.clear()
        _IC_CACHE[key] = m.initial_state(m.parse_case_id(case_name), cfg, params)
    U0 = _IC_CACHE[key]
    L = m.lib()
    with m.Run(cfg, SOLVERS[solver], LIMITERS[limiter]) as run:
        run.set_state(U0)
        if warmup:
            run.step_to(cfg.t_end, warmup)
        L.mhd_profile(run.handle, 1)
        run.synchronize()
        t0 = time.perf_counter()
        r = run.step_to(cfg.t_end if t_end is None else t_end, steps if steps is not None else 10 ** 7)
        run.synchronize()
        sec = time.perf_counter() - t0
        ms4 = (C.c_double * 4)()
        nst = C.c_int()
        L.mhd_profile_read(run.handle, ms4, C.byref(nst))
        L.mhd_profile(run.handle, 0)
    k = max(nst.value, 1)
    ysweep = ms4[2] / k
    slab = f" on [{cfg.x_min:g},{cfg.x_max:g}]x[{cfg.y_min:g},{cfg.y_max:g}]" if y_max is not None else ""
    out = {"workload": f"{case_name} {nx}x{ny}{slab}, {solver.upper()}, {limiter}, CFL {cfg.cfl}", "steps": r["steps"],
           "t": r["t"], "stop_code": r["code"], "value": r["steps"] * nx * ny / sec, "unit": "cell-updates/s",
           "ms_per_step": 1e3 * sec / max(r["steps"], 1), "xsweep_ms": ms4[0] / k, "ysweep_ms": ysweep,
           "ysweep_hbm_frac": BYTES_PER_CELL_SWEEP * nx * ny / (ysweep * 1e-3) / HBM_PEAK}
    if cpu_seconds > 0:
        out.update(_cpu_parity(U0, cfg, solver, limiter, cfg.t_end if t_end is None else t_end, cpu_seconds))
    return out


def _cpu_parity(U0, cfg, solver, limiter, t_end, budget):
    """The oracle (C++/OpenMP restatement of the reference CPU path) on the first steps of
    the same run from the same initial state, about `budget` seconds of it, on the host's
    cores; then the GPU over the same steps from the same state.  Returns the CPU
    cell-updates/s and the parity columns: rel-L1 of the GPU state against the CPU's over
    the nine interior fields, sum |gpu - cpu| / sum |cpu|, and the largest relative
    deviation of the dt sequence."""
    import numpy as np

    import mhd_gpu_acc_amd as m
    import oracle

    case = oracle.Case(cfg.nx, cfg.ny, 2, cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.cfl, t_end, cfg.gamma,
                       tuple(cfg.bc.as_list()))
    s, lim = SOLVERS[solver], LIMITERS[limiter]
    _, t1 = oracle.run_loop_timed(U0, case, s, lim, 1)
    k = int(max(1, min(10 ** 7, round(budget / max(t1, 1e-9)))))
    res, secs = oracle.run_loop_timed(U0, case, s, lim, k)
    n = res["nsteps"]
    with m.Run(cfg, SOLVERS[solver], LIMITERS[limiter]) as run:
        run.set_state(U0)
        g = run.step_to(t_end, n)
        Ug = run.get_state()
    Uc = res["U"]
    a, b = Ug[2:-2, 2:-2], Uc[2:-2, 2:-2]
    den = float(np.abs(b).sum())
    same_dt = g["steps"] == n
    dev = float(np.max(np.abs(g["dts"] - res["dts"]) / res["dts"])) if same_dt and n else float("nan")
    return {"cpu_value": n * cfg.nx * cfg.ny / secs, "cpu_cores": oracle.num_threads(), "cpu_kind": "port",
            "cpu_sample": f"first {n} steps of the same run ({secs:.2f} s)",
            "rel_l1": float(np.abs(a - b).sum()) / den if den > 0 else float("nan"),
            "dt_max_rel_dev": dev,
            "bitwise_vs_cpu": bool(same_dt and np.array_equal(Ug.view(np.uint64), Uc.view(np.uint64)))}


def bench_configs(cpu_seconds: float = 2.0):
    """The other plain configurations: Orszag-Tang at 4096^2 and 1024^2 with each Riemann
    solver (van Leer, CFL 0.8, 5 warm-up + 100 timed steps), the 1024^2 HLLD run to
    t = 0.5 (CFL 0.4), one 4096x512 y slab of the 4096^2 domain (HLLD, 100 steps), the
    seeded shock-cloud 2048x1024 (HLLD, 200 steps) and Brio-Wu 800x4 (HLLD, minmod,
    CFL 0.8, to t = 0.1).  Each one also times the oracle on about
    cpu_seconds of its first steps on the host's cores and reports the GPU's parity over
    those steps (cpu_value, rel_l1, dt_max_rel_dev; 0 skips that leg)."""
    import mhd_gpu_acc_amd as m

    c = cpu_seconds
    res = []
    for n in (4096, 1024):
        for s in SOLVERS:
            res.append(_time_plain("orszag_tang_unit", n, n, s, "vanleer", 5, 100, cpu_seconds=c))
    # the accuracy run to t = 0.5: HLLD reaches it at the reference's own Orszag-Tang CFL
    # of 0.4; at 0.8 its run (like the reference's) stops at the blow-up guard near t = 0.14
    res.append(_time_plain("orszag_tang_unit", 1024, 1024, "hlld", "vanleer", 0, None, t_end=0.5, cfl=0.4,
                           cpu_seconds=c))
    # one GPU's share of the 4096^2 run split into 8 y slabs: 4096x512 cells of the same
    # size (y in [0, 1/8], periodic), the shape each GPU of a slab decomposition steps
    res.append(_time_plain("orszag_tang_unit", 4096, 512, "hlld", "vanleer", 5, 100, cpu_seconds=c, y_max=0.125))
    p = m.ICParams()
    p.seed = 20261203
    res.append(_time_plain("shock_cloud", 2048, 1024, "hlld", "vanleer", 5, 200, params=p, cpu_seconds=c))
    res.append(_time_plain("brio_wu_transmissive", 800, 4, "hlld", "minmod", 0, None, cpu_seconds=c))
    _IC_CACHE.clear()
    return res


def bench_batch(steps: int, seed: int = 20261203):
    """64 Orszag-Tang-family runs (128^2 x32, 256^2 x16, 512^2 x12, 1024^2 x4; HLLD, van Leer):
    the same runs stepped one after another with the plain path, then as one batch
    (autotuned, one graph replay per step).  Only the step loops are timed."""
    import hashlib

    import numpy as np
    import torch

    import mhd_gpu_acc_amd as m
    from mhd_gpu_acc_amd.batch import ot_family_batch

    def h(U):
        return hashlib.sha256(np.ascontiguousarray(U).tobytes()).hexdigest()

    b, cfgs, params, states = ot_family_batch(seed)
    cells = sum(c.nx * c.ny for c in cfgs)
    # initial states: the C++ host initialiser plus one upload per run, against one
    # device launch for the whole batch (the second of two calls is timed)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    host_states = [m.initial_state(m.CaseId.OrszagTangFamily, c, p) for c, p in zip(cfgs, params)]
    for r, U in enumerate(host_states):
        b.set_state(r, U)
    torch.cuda.synchronize()
    t_host_init = time.perf_counter() - t0
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.init_states(m.CaseId.OrszagTangFamily, params)
        torch.cuda.synchronize()
        t_dev_init = time.perf_counter() - t0
    init_same = all(h(b.get_state(r)) == h(host_states[r]) for r in range(len(cfgs)))
    del host_states
    report = b.tune(2)
    b.step(1)  # warm-up (graph upload), then re-seed the states
    for r, U in enumerate(states):
        b.set_state(r, U)
        b.set_ch(r, 0.0)
    b.step(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.step(steps, record=False)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    check = [0, 1, 2] + [r for r in range(len(cfgs)) if cfgs[r].nx == 1024][:1]
    bh = {r: h(b.get_state(r)) for r in check}
    b.close()

    t_plain = 0.0
    same = True
    for r, (c, U0) in enumerate(zip(cfgs, states)):
        with m.Run(c, m.RiemannSolver.HLLD, m.Limiter.VAN_LEER) as run:
            run.set_state(U0)
            run.step_to(1.0, 1)  # warm-up step, then restart from the IC
            run.set_state(U0)
            run.ch = 0.0
            run.synchronize()
            t0 = time.perf_counter()
            rr = run.step_to(1.0, steps, record=False)
            assert rr["code"] == 0 and rr["steps"] == steps, rr
            run.synchronize()
            t_plain += time.perf_counter() - t0
            if r in bh:
                same = same and (h(run.get_state()) == bh[r])
    best = report.strip().splitlines()[-1]
    return {"value": steps * cells / t_batch, "unit": "cell-updates/s", "runs": len(cfgs), "cells_per_step": cells,
            "steps": steps, "sequential_plain": steps * cells / t_plain, "speedup": t_plain / t_batch,
            "bitwise_equal_to_plain_runs_checked": same, "runs_checked": check, "tuned": best,
            "init_host_and_upload_ms": 1e3 * t_host_init, "init_device_ms": 1e3 * t_dev_init,
            "init_device_bitwise_equal_to_host_all_runs": init_same}


def bench_adjoint(n: int = 512, steps: int = 200, K: int = 20, seed: int = 20261203):
    """Adjoint workload: unit-box Orszag-Tang n^2, HLLD, van Leer, `steps` steps, checkpoint
    every K.  Loss 1/2 |rho_N - rho_target|^2 with the target from the same run on the
    state plus a seeded Gaussian density bump; gradient w.r.t. all 9 initial fields
    through the torch.autograd.Function.  Cost = (forward with checkpoints + backward)
    / plain forward, each timed alone with the state resident on the device.

    Accuracy at 32 seeded probes, against tests/golden/adjoint_ot512.json (CPU):
    on the Orszag-Tang state, the oracle's tangent model and the CPU adjoint (SHA-256
    of the whole gradient); on the same state with seeded 1e-3 noise (no exact ties
    of the scheme, so differences are well posed), the tangent model and
    frozen-branch central differences (two-point at eps 1e-6, 1e-7, 2e-6 and the
    fourth-order stencil at eps 1e-6), plus plain GPU central
    differences (eps 1e-6, dt / c_h frozen, branches free) for comparison."""
    import hashlib

    import numpy as np
    import torch

    import mhd_gpu_acc_amd as m
    from mhd_gpu_acc_amd.adjoint import (Tape, adjoint_workload, from_planes, generic_state, probe_pairs, replay,
                                         run_steps, to_planes)

    cfg, U0, Ubump = adjoint_workload(n, seed)
    golden = os.path.join(ROOT, "tests", "golden", "adjoint_ot512.json")
    ref = json.load(open(golden)) if (n, steps, K) == (512, 200, 20) and os.path.exists(golden) else None
    pairs = probe_pairs(seed, n, n)
    with m.Run(cfg, m.RiemannSolver.HLLD, m.Limiter.VAN_LEER) as run:
        run.set_state(Ubump)
        assert run.step_to(1.0, steps)["steps"] == steps
        target = to_planes(run.get_state())[0, 2:-2, 2:-2].clone()

        def plain():
            run.set_state(U0)
            run.ch = 0.0
            run.synchronize()
            t0 = time.perf_counter()
            r = run.step_to(1.0, steps, record=False)
            run.synchronize()
            assert r["steps"] == steps
            return time.perf_counter() - t0

        def adjoint(U):
            run.ch = 0.0
            U0t = to_planes(U).requires_grad_(True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            UN = run_steps(U0t, run, steps, K)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            res = UN[0, 2:-2, 2:-2] - target
            loss = 0.5 * (res * res).sum()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            loss.backward()
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            return t1 - t0, t3 - t2, float(loss.detach()), from_planes(U0t.grad)

        plain()
        adjoint(U0)  # warm-up of both paths
        t_fwd = min(plain() for _ in range(3))
        runs = [adjoint(U0) for _ in range(3)]
        t_ckpt = min(r[0] for r in runs)
        t_bwd = min(r[1] for r in runs)
        loss, grad = runs[-1][2], runs[-1][3]

        Ug = generic_state(U0, 7)
        _, _, loss_g, grad_g = adjoint(Ug)
        run.set_state(Ug)
        run.ch = 0.0
        with Tape(run, steps, K) as tape:
            dts, chs = tape.forward()
        base = to_planes(Ug)
        eps = 1e-6
        fd = []
        for j, i, f in pairs:
            up, um = base.clone(), base.clone()
            up[f, j, i] += eps
            um[f, j, i] -= eps
            rp = replay(run, up, dts, chs)[0, 2:-2, 2:-2] - target
            rm = replay(run, um, dts, chs)[0, 2:-2, 2:-2] - target
            fd.append(float(((rp - rm) * (rp + rm)).sum() * 0.5) / (2 * eps))
        fd = np.array(fd)

    def probes(g):
        return np.array([g[j, i, f] for j, i, f in pairs])

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    adj, adj_g = probes(grad), probes(grad_g)
    out = {"workload": f"orszag_tang_unit {n}x{n} periodic, HLLD, vanleer, {steps} steps, checkpoint every {K}",
           "loss": loss, "forward_ms": 1e3 * t_fwd, "forward_with_checkpoints_ms": 1e3 * t_ckpt,
           "backward_ms": 1e3 * t_bwd, "cost_ratio": (t_ckpt + t_bwd) / t_fwd, "probes": len(pairs),
           "generic_state_gpu_plain_central_differences_eps1e-06_norm_rel": rel(adj_g, fd)}
    if ref is not None:
        ex, ge = ref["exact"], ref.get("generic", {})
        out["cpu_tangent_norm_rel"] = rel(adj, np.array(ex["tangent"]["probes"]))
        out["cpu_adjoint_bitwise"] = (hashlib.sha256(np.ascontiguousarray(grad).tobytes()).hexdigest()
                                      == ex["cpu_adjoint"]["sha256"])
        if ge:
            out["generic_state_cpu_tangent_norm_rel"] = rel(adj_g, np.array(ge["tangent"]["probes"]))
            ffd = ge.get("frozen_fd", {})
            for key, vals in sorted(ffd.items()):
                out[f"generic_state_frozen_central_differences_eps{key}_norm_rel"] = rel(adj_g, np.array(vals))
            if "1e-06" in ffd and "2e-06" in ffd:
                fd4 = (4.0 * np.array(ffd["1e-06"]) - np.array(ffd["2e-06"])) / 3.0
                out["generic_state_frozen_central_differences_4th_order_eps1e-06_norm_rel"] = rel(adj_g, fd4)
    return out


if __name__ == "__main__":
    main()
