"""Output formats on the CPU: the snapshot CSV writer (byte for byte the reference
writer's on the reference's own initial states, and "%.8e" of the right quantity in
the right place on a random state), the snapshot schedule against the case table,
the state-dump reader, and the comparer (closed forms, PASS rule, exit codes)."""
from __future__ import annotations

import hashlib
import io
import json
import math
import os
import struct

import numpy as np
import pytest

import mhd_gpu_acc_amd as m
import oracle
from mhd_gpu_acc_amd import compare
from mhd_gpu_acc_amd.driver import BIN_HEADER_BYTES, case_snapshots, load_bin, write_fields_csv_host
from oracle import ptr

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "driver_reference.json")
FIELDS = ("rho", "u", "v", "Bx", "By", "Bz", "E", "psi", "p")


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


@pytest.mark.parametrize("case", ["brio_wu", "orszag_tang", "shock_bubble"])
def test_csv_writer_matches_reference_writer_on_initial_state(case, tmp_path):
    """The reference driver's <case>_cpu_t0_<f>.csv files (tests/golden/driver_reference.json,
    written by oracle/_ref/main_cpu) are reproduced byte for byte from the host initial state."""
    ref = json.load(open(GOLDEN))[case]
    cfg = m.get_n_case_config(case, 1)
    U = m.initial_state(case, cfg)
    write_fields_csv_host(U, cfg.gamma, str(tmp_path), f"{case}_cpu_t0")
    for f in FIELDS:
        name = f"{case}_cpu_t0_{f}.csv"
        assert sha(tmp_path / name) == ref["files"][name]["sha256"], name
    assert not list(tmp_path.glob("*.tmp"))


def test_csv_writer_values_and_layout(tmp_path):
    """Random state with special values: each file holds, row j = ny-1 first, "%.8e" of
    rho, m_x / rho, m_y / rho, Bx, By, Bz, E, psi and the oracle's cons_to_prim pressure."""
    nx, ny, gamma = 13, 7, 1.4
    rng = np.random.default_rng(3)
    U = rng.uniform(-2.0, 2.0, (ny + 4, nx + 4, 9))
    U[..., 0] = rng.uniform(0.1, 3.0, U.shape[:2])
    U[..., 7] = rng.uniform(5.0, 9.0, U.shape[:2])
    specials = [0.0, -0.0, 5e-324, 1e-300, 1e300, -1e-7, math.inf, -math.inf, math.nan, 123456789.5]
    for k, v in enumerate(specials):
        U[2 + k % ny, 2 + k, 4] = v
    write_fields_csv_host(U, gamma, str(tmp_path), "x")
    lib = oracle.orc()
    V = np.zeros(9)
    pres = np.empty((ny, nx))
    for j in range(ny):
        for i in range(nx):
            lib.orc_cons_to_prim(ptr(np.ascontiguousarray(U[j + 2, i + 2])), gamma, ptr(V))
            pres[j, i] = V[7]
    inner = U[2:-2, 2:-2]
    expect = {"rho": inner[..., 0], "u": inner[..., 1] / inner[..., 0], "v": inner[..., 2] / inner[..., 0],
              "Bx": inner[..., 4], "By": inner[..., 5], "Bz": inner[..., 6], "E": inner[..., 7],
              "psi": inner[..., 8], "p": pres}

    def fmt(v):
        if math.isnan(v):
            return "-nan" if math.copysign(1.0, v) < 0 else "nan"
        return "%.8e" % v

    for f in FIELDS:
        lines = (tmp_path / f"x_{f}.csv").read_text().split("\n")
        assert lines[-1] == "" and len(lines) == ny + 1
        for r, line in enumerate(lines[:-1]):
            j = ny - 1 - r
            assert line == ",".join(fmt(v) for v in expect[f][j]), (f, j)


def test_snapshot_schedule_matches_case_table():
    times, tags = case_snapshots("shock_bubble")
    assert tags == ["t006", "t012", "t018", "t030", "t046", "t062", "t078", "t126", "t190"]
    cfg = m.get_n_case_config("shock_bubble", 1)
    assert times[-1] == cfg.t_end  # the case table's t_end is the last snapshot (test_cases.cpp:232)
    assert all(a < b for a, b in zip(times, times[1:]))
    # equal steps of the dimensionless time: T = T0 + t_paper tau
    tau = (times[-1] - times[0]) / (19.0 - 0.6)
    assert times[3] == pytest.approx(times[0] + 2.4 * tau, rel=1e-12)
    t2, g2 = case_snapshots("orszag_tang")
    assert t2 == [math.pi, 2 * math.pi] and g2 == ["tpi", "t2pi"]
    assert case_snapshots("brio_wu") == ([], [])


def test_state_dump_reader(tmp_path):
    """load_bin reads the documented layout: 128-byte header, 9 planes with ghosts."""
    nx, ny, ng = 5, 3, 2
    U = np.random.default_rng(0).standard_normal((ny + 4, nx + 4, 9))
    head = b"MHDSTATE" + struct.pack("<IIiiii", 1, BIN_HEADER_BYTES, nx, ny, ng, 9) + \
        struct.pack("<dqdd", 0.25, 17, 1.4, 2.5) + struct.pack("<I", 1)
    head += b"\0" * (BIN_HEADER_BYTES - len(head))
    p = tmp_path / "s.bin"
    p.write_bytes(head + np.ascontiguousarray(np.moveaxis(U, -1, 0)).astype("<f8").tobytes())
    h, V = load_bin(str(p))
    assert h == dict(nx=nx, ny=ny, ng=ng, t=0.25, step=17, gamma=1.4, ch=2.5)
    assert np.array_equal(V, U)
    p.write_bytes(head[:40])
    with pytest.raises(ValueError):
        load_bin(str(p))


def test_compare_stats_closed_forms():
    a = np.arange(1.0, 13.0).reshape(3, 4)
    st = compare.diff_stats(a, a.copy())
    assert (st["max_abs"], st["mean_abs"], st["max_rel"], st["l2_abs"], st["rel_l1"]) == (0, 0, 0, 0, 0)
    b = a.copy()
    b[1, 2] += 0.5  # one cell off by 0.5
    st = compare.diff_stats(a, b)
    assert st["max_abs"] == 0.5 and st["mean_abs"] == 0.5 / 12
    assert st["l2_abs"] == pytest.approx(0.5 / math.sqrt(12), rel=1e-15)
    assert st["max_rel"] == 0.5 / 7.5
    assert st["rel_l1"] == 0.5 / b.sum()
    z = np.zeros((2, 2))
    assert compare.diff_stats(z, z)["max_rel"] == 0.0   # 0 where both are 0
    assert compare.diff_stats(a, a[:2]) is None          # shape mismatch


def _write_run(d, arrays):
    d.mkdir(parents=True, exist_ok=True)
    for name, arr in arrays.items():
        np.savetxt(d / name, arr, delimiter=",", fmt="%.8e")


def test_compare_dirs_pass_rule_and_exit_codes(tmp_path):
    base = np.random.default_rng(1).uniform(1.0, 2.0, (6, 5))
    _write_run(tmp_path / "a", {"c_gpu_t10_rho.csv": base, "c_gpu_t10_p.csv": base + 1, "extra_rho.csv": base})
    _write_run(tmp_path / "b", {"c_gpu_t10_rho.csv": base, "c_gpu_t10_p.csv": base + 1})
    out = io.StringIO()
    assert compare.compare_dirs(tmp_path / "a", tmp_path / "b", out=out) == 0
    assert "extra_rho.csv" in out.getvalue() and "Overall result     : PASS" in out.getvalue()
    # max_abs > tol but max_rel <= rtol passes (either criterion, compare_runs.py:176)
    _write_run(tmp_path / "c", {"c_gpu_t10_rho.csv": base * (1 + 1e-7), "c_gpu_t10_p.csv": base + 1})
    assert compare.compare_dirs(tmp_path / "c", tmp_path / "b", out=io.StringIO()) == 0
    # ... unless the rel-L1 bound is asked for
    assert compare.compare_dirs(tmp_path / "c", tmp_path / "b", rel_l1=1e-12, out=io.StringIO()) == 1
    _write_run(tmp_path / "d", {"c_gpu_t10_rho.csv": base * 1.01, "c_gpu_t10_p.csv": base + 1})
    out = io.StringIO()
    assert compare.compare_dirs(tmp_path / "d", tmp_path / "b", verbose=True, out=out) == 1
    assert "FAIL" in out.getvalue()
    assert compare.compare_dirs(tmp_path / "d", tmp_path / "b", fields="p", out=io.StringIO()) == 0
    _write_run(tmp_path / "e", {"c_gpu_t10_rho.csv": base[:, :4], "c_gpu_t10_p.csv": base + 1})
    assert compare.compare_dirs(tmp_path / "e", tmp_path / "b", out=io.StringIO()) == 1  # shape mismatch
    assert compare.compare_dirs(tmp_path / "nope", tmp_path / "b", out=io.StringIO()) == 2
    assert compare.main([str(tmp_path / "a"), str(tmp_path / "b")]) == 0


def test_writer_and_reader_errors(tmp_path):
    """A blocked output directory is an I/O error (no partial files); foreign or
    truncated dumps are refused."""
    from mhd_gpu_acc_amd._lib import MHDError
    blocker = tmp_path / "taken"
    blocker.write_text("a file where the directory should be")
    U = np.ones((8, 9, 9))
    with pytest.raises(MHDError, match="create_directories"):
        write_fields_csv_host(U, 1.4, str(blocker / "sub"), "x")
    out = tmp_path / "ok"
    write_fields_csv_host(U, 1.4, str(out), "x")
    assert sorted(p.name for p in out.iterdir()) == sorted(f"x_{f}.csv" for f in FIELDS)
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"NOTASTATE" + b"\0" * 200)
    with pytest.raises(ValueError, match="not an MHD state dump"):
        load_bin(str(bad))
    head = b"MHDSTATE" + struct.pack("<IIiiii", 1, BIN_HEADER_BYTES, 4, 4, 2, 9) + \
        struct.pack("<dqdd", 0.0, 0, 1.4, 0.0) + struct.pack("<I", 1)
    head += b"\0" * (BIN_HEADER_BYTES - len(head))
    bad.write_bytes(head + b"\0" * 64)  # 8 of 9*8*8 doubles
    with pytest.raises(ValueError, match="short file"):
        load_bin(str(bad))
