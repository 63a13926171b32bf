"""The run driver on the GPU against the reference's own CPU driver
(tests/golden/driver_reference.json, from oracle/_ref/main_cpu): for Brio-Wu,
Orszag-Tang (snapshot schedule), shock-bubble (nine snapshots) and Kelvin-Helmholtz
(4 695 steps), every snapshot
CSV is byte for byte the reference's, the progress lines (step, t, dt, max|B|) and
[snap] lines are the reference's verbatim, and the step count is the same.  A run
restarted from a state dump ends in the same files as the uninterrupted run; a
snapshot case without output reaches t_end (the reference stops at the first
snapshot there)."""
from __future__ import annotations

import hashlib
import json
import math
import os
import re
import sys

import numpy as np
import pytest

import mhd_gpu_acc_amd as m
from mhd_gpu_acc_amd.driver import (DRIVE_DONE, DRIVE_MAX_STEPS, case_snapshots, drive, load_bin, read_state_bin,
                                    run_case, write_state_bin)

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "driver_reference.json")
PROGRESS = re.compile(r"^\s+\d+\s+\d+\.\d{5}\s+\S+\s+\S+$")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def _check_files(out_dir, ref, only=None):
    for name, meta in ref["files"].items():
        if only and not any(tag in name for tag in only):
            continue
        ours = os.path.join(out_dir, name.replace("_cpu_", "_gpu_"))
        assert os.path.exists(ours), name
        assert sha(ours) == meta["sha256"], name


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["brio_wu", "orszag_tang", "shock_bubble", "kelvin_helmholtz"])
def test_driver_matches_reference_driver(case, tmp_path):
    ref = json.load(open(GOLDEN))[case]
    say(f"[gpu] driver {case} hlld minmod n=1 ({ref['steps']} steps)")
    log = tmp_path / "log.txt"
    res = run_case(case, 1, m.RiemannSolver.HLLD, m.Limiter.MINMOD, str(tmp_path / "out"), str(log))
    assert res["status"] == DRIVE_DONE and res["steps"] == ref["steps"]
    _check_files(tmp_path / "out", ref)
    assert len(os.listdir(tmp_path / "out")) == len(ref["files"])
    lines = log.read_text().splitlines()
    assert [ln for ln in lines if PROGRESS.match(ln)] == ref["progress"]
    assert [ln for ln in lines if ln.startswith("  [snap]")] == ref["snaps"]
    timing = [ln for ln in lines if ln.startswith("[TIMING]")]
    assert len(timing) == 1 and f" steps={ref['steps']} " in timing[0]


@pytest.mark.gpu
def test_restart_from_state_dump_is_exact(tmp_path):
    """Brio-Wu: 400 steps, dump, a fresh run loads the dump (state and c_h) and drives on
    from (t, step): its final files equal the uninterrupted reference run's."""
    ref = json.load(open(GOLDEN))["brio_wu"]
    cfg = m.get_n_case_config("brio_wu", 1)
    say("[gpu] driver restart brio_wu at step 400")
    with m.Run(cfg, m.RiemannSolver.HLLD, m.Limiter.MINMOD) as run:
        run.init_state("brio_wu")
        run.ch = 0.0
        r1 = drive(run, "brio_wu", cfg.t_end, max_steps=400, verbose=False)
        assert r1["status"] == DRIVE_MAX_STEPS and r1["steps"] == 400
        dump = str(tmp_path / "s400.bin")
        write_state_bin(run, dump, r1["t"], r1["steps"])
        head, U = load_bin(dump)
        assert (head["t"], head["step"], head["ch"]) == (r1["t"], 400, run.ch)
        assert np.array_equal(U.view(np.uint64), run.get_state().view(np.uint64))
    with m.Run(cfg, m.RiemannSolver.HLLD, m.Limiter.MINMOD) as run2:
        t, step = read_state_bin(run2, dump)
        r2 = drive(run2, "brio_wu", cfg.t_end, out_dir=str(tmp_path / "out"), verbose=False, t_start=t,
                   step_start=step)
    assert r2["status"] == DRIVE_DONE and r2["steps"] == ref["steps"]
    _check_files(tmp_path / "out", ref, only=["_t10_"])
    assert not any("_t0_" in n for n in os.listdir(tmp_path / "out"))


@pytest.mark.gpu
def test_snapshot_case_without_output_reaches_t_end():
    ref = json.load(open(GOLDEN))["orszag_tang"]
    cfg = m.get_n_case_config("orszag_tang", 1)
    say("[gpu] driver orszag_tang without output")
    with m.Run(cfg, m.RiemannSolver.HLLD, m.Limiter.MINMOD) as run:
        run.set_state(m.initial_state("orszag_tang", cfg))
        run.ch = 0.0
        res = drive(run, "orszag_tang", cfg.t_end, case_snapshots("orszag_tang"), out_dir=None, verbose=False)
    # the same clamped dt sequence as the run that writes its snapshots
    assert res["status"] == DRIVE_DONE and res["steps"] == ref["steps"] and res["t"] >= 2 * math.pi
