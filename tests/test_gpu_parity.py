"""GPU parity of the plain step: the HIP path against the reference's golden
vectors and against the CPU oracle, bitwise (FP64 results, dt and c_h
sequences, stop codes).

Every test here runs the product through the C ABI (libmhd_hip.so); the oracle
is only the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np
import pytest

import mhd_gpu_acc_amd as m
import oracle
from tests.util import assert_bitwise, load_grid, load_meta, parse_float, rel_l1, state_hash

pytestmark = pytest.mark.gpu

G = load_grid()
META = load_meta()


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def cfg_of(spec) -> m.CaseConfig:
    x0, x1 = spec["x"]
    y0, y1 = spec["y"]
    return m.CaseConfig(spec["nx"], spec["ny"], 2, x0, x1, y0, y1, spec["cfl"], spec["t_end"], spec["gamma"],
                        m.BoundaryConfig(*[m.BoundaryType(b) for b in spec["bc"]]))


def apply_faults(U, faults):
    for f in faults or []:
        if f[0] == "all":
            U[2:-2, 2:-2, f[1]] = parse_float(f[2])
        else:
            i, j, k, v = f
            U[j + 2, i + 2, k] = parse_float(v)


def gpu_run(cfg, U0, solver, limiter, steps, t_end=None):
    with m.Run(cfg, solver, limiter) as run:
        run.set_state(U0)
        r = run.step_to(cfg.t_end if t_end is None else t_end, steps)
        r["U"] = run.get_state()
        r["ch"] = run.ch
    return r


@pytest.mark.parametrize("name", sorted(META["trajectories"]))
def test_reference_trajectory(name):
    """The reference's own runs (golden vectors): dt/c_h per step, stop code, final state."""
    tr = META["trajectories"][name]
    spec = tr["spec"]
    cfg = cfg_of(spec)
    U0 = m.initial_state(spec["case"], cfg)
    apply_faults(U0, spec.get("faults"))
    assert state_hash(U0) == tr["ic_hash"]
    say(f"[gpu] trajectory {name}: {cfg.nx}x{cfg.ny} solver={spec['solver']} steps<={spec['steps']}")
    r = gpu_run(cfg, U0, spec["solver"], 0, spec["steps"])
    codes = {0: 0, 1: 3, 2: 4, 3: 5}  # reference loop stop -> C ABI status
    assert r["code"] == codes[tr["code"]], (r["code"], tr["code"])
    assert r["steps"] == tr["nsteps"]
    assert_bitwise(r["dts"], G[f"{name}__dts"], "dt sequence")
    assert_bitwise(r["chs"], G[f"{name}__chs"], "c_h sequence")
    assert r["t"].hex() == tr["t"]
    if state_hash(r["U"]) != tr["final_hash"]:
        # diagnose against the oracle (pinned to the same hash)
        case = oracle.Case(cfg.nx, cfg.ny, 2, cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.cfl, cfg.t_end,
                           cfg.gamma, tuple(spec["bc"]))
        o = oracle.run_loop(U0, case, spec["solver"], 0, spec["steps"])
        assert_bitwise(r["U"], o["U"], "final state vs oracle")
        raise AssertionError("final state hash differs")


def ocase(cfg):
    return oracle.Case(cfg.nx, cfg.ny, 2, cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.cfl, cfg.t_end,
                       cfg.gamma, tuple(cfg.bc.as_list()))


@pytest.mark.parametrize("solver", [0, 1, 2, 3])
@pytest.mark.parametrize("limiter", [0, 1])
def test_orszag_tang_unit_vs_oracle(solver, limiter):
    cfg = m.get_n_case_config("orszag_tang_unit", 1)
    cfg.nx, cfg.ny = 96, 80
    U0 = m.initial_state(4, cfg)
    say(f"[gpu] OT unit 96x80 solver={solver} limiter={limiter} 25 steps")
    r = gpu_run(cfg, U0, solver, limiter, 25)
    o = oracle.run_loop(U0, ocase(cfg), solver, limiter, 25)
    assert r["steps"] == o["nsteps"] == 25
    assert_bitwise(r["dts"], o["dts"], "dt")
    assert_bitwise(r["chs"], o["chs"], "c_h")
    assert_bitwise(r["U"], o["U"], "state")


@pytest.mark.parametrize("variant", ["scaled", "tiny_bz"])
@pytest.mark.parametrize("solver", [0, 1, 2, 3])
@pytest.mark.parametrize("limiter", [0, 1])
def test_quotient_range_fallbacks_vs_oracle(variant, solver, limiter):
    """States whose division operands leave the range of the shared-reciprocal quotients
    (frexp exponents beyond +-380): rho, momenta and E scaled by 1e120 everywhere (operands
    near 2^399: every face and reconstruction takes the plain divisions), or a tiny B_z of
    1e-150 (side-star numerators near 2^-498 in some lanes of a wave, in range elsewhere).
    The run must still be the oracle's bit for bit."""
    cfg = m.get_n_case_config("orszag_tang_unit", 1)
    cfg.nx, cfg.ny = 96, 80
    U0 = m.initial_state(4, cfg)
    if variant == "scaled":
        U0[:, :, [0, 1, 2, 3, 7]] *= 1e120
        assert np.frexp(U0[20, 20, 0])[1] > 380
    else:
        x = np.arange(U0.shape[1])
        U0[:, :, 6] = 1e-150 * np.sin(0.3 * x)[None, :]
        assert np.frexp(1e-150 * 0.3)[1] < -380
    say(f"[gpu] OT unit 96x80 {variant}, solver={solver} limiter={limiter} 10 steps")
    r = gpu_run(cfg, U0, solver, limiter, 10)
    o = oracle.run_loop(U0, ocase(cfg), solver, limiter, 10)
    assert r["steps"] == o["nsteps"] == 10
    assert_bitwise(r["dts"], o["dts"], "dt")
    assert_bitwise(r["chs"], o["chs"], "c_h")
    assert_bitwise(r["U"], o["U"], "state")


def test_shock_cloud_vs_oracle():
    cfg = m.get_n_case_config("shock_cloud", 1)
    cfg.nx, cfg.ny = 160, 80
    p = m.ICParams()
    p.seed = 20261203
    po = oracle.ICParams()
    po.seed = 20261203
    U0 = m.initial_state(5, cfg, p)
    assert_bitwise(U0, oracle.init_state(5, ocase(cfg), po), "IC")
    say("[gpu] shock-cloud 160x80 HLLD van Leer 40 steps")
    r = gpu_run(cfg, U0, 2, 1, 40)
    o = oracle.run_loop(U0, ocase(cfg), 2, 1, 40)
    assert_bitwise(r["dts"], o["dts"], "dt")
    assert_bitwise(r["U"], o["U"], "state")


def test_ot_family_vs_oracle():
    cfg = m.get_n_case_config("orszag_tang_family", 1)
    cfg.nx, cfg.ny = 64, 64
    p = m.ICParams()
    po = oracle.ICParams()
    for obj in (p, po):
        obj.gamma, obj.b_scale, obj.n_modes, obj.pert_amp = 1.93, 1.41, 4, 0.05
        for q, (a, kx, ky, ph) in enumerate(((0.5, 1, 1, 0.2), (-0.8, 3, 0, 1.7), (0.2, -2, 2, 4.0),
                                              (0.9, 0, -4, 3.3))):
            obj.mode_amp[q], obj.mode_kx[q], obj.mode_ky[q], obj.mode_phase[q] = a, kx, ky, ph
    cfg.gamma = 1.93
    U0 = m.initial_state(7, cfg, p)
    say("[gpu] OT family 64^2 HLLD van Leer 30 steps")
    r = gpu_run(cfg, U0, 2, 1, 30)
    o = oracle.run_loop(U0, ocase(cfg), 2, 1, 30)
    assert_bitwise(r["dts"], o["dts"], "dt")
    assert_bitwise(r["U"], o["U"], "state")


def test_boundary_kats():
    """apply_boundary in place on random grids, every BC type on every side."""
    Gs = G["bc_G"]
    ny, nx = Gs.shape[0] - 4, Gs.shape[1] - 4
    for c, combo in enumerate(G["bc_combos"]):
        cfg = m.CaseConfig(nx, ny, 2, 0.0, 1.0, 0.0, 1.0, 0.5, 1.0, 1.4,
                           m.BoundaryConfig(*[m.BoundaryType(int(b)) for b in combo]))
        say(f"[gpu] boundary {list(combo)}")
        with m.Run(cfg) as run:
            run.set_state(Gs)
            run.apply_boundary()
            assert_bitwise(run.get_state(), G["bc_out"][c], f"apply_boundary{list(combo)}")


@pytest.mark.parametrize("bc", [(0, 0, 0, 0), (1, 2, 2, 1), (3, 1, 0, 0), (2, 3, 3, 2), (0, 3, 1, 3)])
def test_all_bc_mixes_step_vs_oracle(bc):
    """Ghost fills composed through every side, Dirichlet included, over several steps."""
    cfg = m.get_n_case_config("kelvin_helmholtz", 1)
    cfg.nx, cfg.ny = 40, 36
    cfg.bc = m.BoundaryConfig(*[m.BoundaryType(b) for b in bc])
    U0 = m.initial_state(0, cfg)
    say(f"[gpu] KH 40x36 bc={bc} HLLD van Leer 12 steps")
    r = gpu_run(cfg, U0, 2, 1, 12)
    o = oracle.run_loop(U0, ocase(cfg), 2, 1, 12)
    assert_bitwise(r["dts"], o["dts"], "dt")
    assert_bitwise(r["U"], o["U"], "state incl. ghosts")


@pytest.mark.parametrize("knobs", [dict(MHD_YSEG=1), dict(MHD_YSEG=3), dict(MHD_YSEG=17), dict(MHD_YSEG=1000),
                                   dict(MHD_XWG=64), dict(MHD_XWG=128), dict(MHD_XGRID=1), dict(MHD_XGRID=3),
                                   dict(MHD_XWG=64, MHD_XGRID=2), dict(MHD_YWG=64), dict(MHD_YWG=128, MHD_YSEG=5),
                                   dict(MHD_ADDR="flat"), dict(MHD_ADDR="flat", MHD_YSEG=7),
                                   dict(MHD_ADDR="flat", MHD_XWG=128, MHD_XGRID=0), dict(MHD_XTAIL=0),
                                   dict(MHD_XTAIL=300), dict(MHD_XTAIL=1000, MHD_XGRID=1),
                                   dict(MHD_XTAIL=77, MHD_XWG=64, MHD_XGRID=2)])
def test_launch_shape_independence(knobs, monkeypatch):
    """Workgroup widths, persistent grids, the share of x segments in one-segment blocks,
    segment lengths and the sweeps' addressing (flat, as for grids above 4 GiB, or buffer
    resources) change the launch, never the result."""
    for k, v in knobs.items():
        monkeypatch.setenv(k, str(v))
    cfg = m.get_n_case_config("orszag_tang_unit", 1)
    cfg.nx, cfg.ny = 70, 90
    U0 = m.initial_state(4, cfg)
    say(f"[gpu] OT 70x90 {knobs}")
    r = gpu_run(cfg, U0, 2, 1, 8)
    o = oracle.run_loop(U0, ocase(cfg), 2, 1, 8)
    assert_bitwise(r["U"], o["U"], f"{knobs}")


@pytest.mark.parametrize("knobs", [dict(MHD_YWG=64), dict(MHD_YWG=128), dict(MHD_YWG=64, MHD_YSEG=16),
                                   dict(MHD_XWG=128, MHD_XGRID=2, MHD_XTAIL=150)])
def test_launch_shapes_at_scale(knobs, monkeypatch):
    """At 1024^2 the grids of both sweeps hold one or more full rounds of workgroups on the
    chip (the 70x90 grid above holds a fraction of one): the narrow y workgroups, with their
    LDS ring sized by the launch, and other x shapes give the default launch's bits."""
    cfg = m.get_n_case_config("orszag_tang_unit", 1)
    cfg.nx = cfg.ny = 1024
    U0 = m.initial_state(4, cfg)
    say("[gpu] OT 1024^2 default launch")
    ref = gpu_run(cfg, U0, 2, 1, 6)
    for k, v in knobs.items():
        monkeypatch.setenv(k, str(v))
    say(f"[gpu] OT 1024^2 {knobs}")
    r = gpu_run(cfg, U0, 2, 1, 6)
    assert_bitwise(r["dts"], ref["dts"], f"dt {knobs}")
    assert_bitwise(r["U"], ref["U"], f"state {knobs}")


def test_small_and_odd_grids():
    for nx, ny in ((2, 2), (3, 5), (5, 2), (65, 3), (257, 7)):
        cfg = m.get_n_case_config("orszag_tang_unit", 1)
        cfg.nx, cfg.ny = nx, ny
        cfg.bc = m.BoundaryConfig(m.BoundaryType.Periodic, m.BoundaryType.Transmissive,
                                  m.BoundaryType.Reflecting, m.BoundaryType.Periodic)
        U0 = m.initial_state(4, cfg)
        say(f"[gpu] odd grid {nx}x{ny}")
        r = gpu_run(cfg, U0, 2, 0, 6)
        o = oracle.run_loop(U0, ocase(cfg), 2, 0, 6)
        assert_bitwise(r["dts"], o["dts"], f"dt {nx}x{ny}")
        assert_bitwise(r["U"], o["U"], f"state {nx}x{ny}")


def test_compute_dt_and_advance_api():
    """The split API (compute_dt / advance / set_ch) equals the loop, and c_h ratchets."""
    cfg = m.get_n_case_config("orszag_tang_unit", 1)
    cfg.nx, cfg.ny = 48, 48
    U0 = m.initial_state(4, cfg)
    say("[gpu] split API")
    with m.Run(cfg, m.RiemannSolver.HLLD, m.Limiter.VAN_LEER) as run:
        run.set_state(U0)
        dts = []
        for _ in range(5):
            dt = run.compute_dt()
            assert run.compute_dt() == dt  # idempotent
            run.advance(dt)
            dts.append(dt)
        U = run.get_state()
        ch = run.ch
    o = oracle.run_loop(U0, ocase(cfg), 2, 1, 5)
    assert_bitwise(np.array(dts), o["dts"], "dt")
    assert_bitwise(U, o["U"], "state")
    assert ch == o["ch"]


def test_cfl_maximum_handoff_after_steps():
    """compute_dt after a step takes the CFL maximum that the step's closing ghost fill
    stored into pinned host memory.  It must equal compute_dt of the same state and c_h
    through the stand-alone reduction, also after two steps with no compute_dt between."""
    cfg = m.get_n_case_config("orszag_tang_unit", 1)
    cfg.nx, cfg.ny = 96, 80
    U0 = m.initial_state(4, cfg)
    say("[gpu] CFL hand-off")
    with m.Run(cfg, m.RiemannSolver.HLLD, m.Limiter.VAN_LEER) as run, \
            m.Run(cfg, m.RiemannSolver.HLLD, m.Limiter.VAN_LEER) as fresh:
        run.set_state(U0)
        dt = run.compute_dt()
        for k in range(4):
            run.advance(dt)
            if k % 2:
                run.advance(0.5 * dt)
            ch = run.ch
            dt = run.compute_dt()
            fresh.set_state(run.get_state())
            fresh.ch = ch
            dt_fresh = fresh.compute_dt()
            assert dt.hex() == dt_fresh.hex(), (k, dt, dt_fresh)
            assert run.ch.hex() == fresh.ch.hex()


def test_grid_level_api_with_torch_buffers():
    """mhd_advance_grids / mhd_compute_dt_grid over caller-owned device grids."""
    import torch
    cfg = m.get_n_case_config("orszag_tang_unit", 1)
    cfg.nx, cfg.ny = 40, 30
    U0 = m.initial_state(4, cfg)
    cells = (cfg.nx + 4) * (cfg.ny + 4)
    soa = torch.from_numpy(np.ascontiguousarray(U0.reshape(cells, 9).T)).cuda()
    tmp = soa.clone()
    new = soa.clone()
    L = m.lib()
    say("[gpu] grid-level API")
    with m.Run(cfg, m.RiemannSolver.HLLC, m.Limiter.MINMOD) as run:
        dts = []
        a, b = soa, new
        for _ in range(4):
            dt = C.c_double()
            assert L.mhd_compute_dt_grid(run.handle, C.c_void_p(a.data_ptr()), C.byref(dt)) == 0
            assert L.mhd_advance_grids(run.handle, C.c_void_p(a.data_ptr()), C.c_void_p(tmp.data_ptr()),
                                       C.c_void_p(b.data_ptr()), dt.value) == 0
            assert L.mhd_sync(run.handle) == 0
            dts.append(dt.value)
            a, b = b, a
        U = a.cpu().numpy().T.reshape(U0.shape)
    o = oracle.run_loop(U0, ocase(cfg), 1, 0, 4)
    assert_bitwise(np.array(dts), o["dts"], "dt")
    assert_bitwise(U, o["U"], "state")


def test_device_state_roundtrip_and_invalid_args():
    import torch
    cfg = m.get_n_case_config("brio_wu", 1)
    U0 = m.initial_state(2, cfg)
    cells = (cfg.nx + 4) * (cfg.ny + 4)
    say("[gpu] state transfer")
    with m.Run(cfg) as run:
        d = torch.from_numpy(np.ascontiguousarray(U0.reshape(cells, 9).T)).cuda()
        run.set_state_device(d.data_ptr())
        assert_bitwise(run.get_state(), U0, "device->host")
        out = torch.empty_like(d)
        run.get_state_device(out.data_ptr())
        assert torch.equal(out, d)
        with pytest.raises(ValueError):
            run.set_state(np.zeros((3, 3, 9)))
        with pytest.raises(m.MHDError):
            m.lib().mhd_set_state(run.handle, None, 0) and m._lib.check(1)


def test_no_signal_is_reported():
    cfg = m.get_n_case_config("orszag_tang_unit", 1)
    cfg.nx, cfg.ny = 8, 8
    U0 = m.initial_state(4, cfg)
    U0[2:-2, 2:-2, 7] = -1.0  # every interior pressure negative
    say("[gpu] no-signal")
    with m.Run(cfg) as run:
        run.set_state(U0)
        with pytest.raises(m.MHDError) as e:
            run.compute_dt()
        assert e.value.code == 3


@pytest.mark.parametrize("n", [1024])
def test_full_size_orszag_tang_vs_oracle(n):
    """The measured size class: OT 1024^2, HLLD van Leer, 3 steps, bitwise vs the oracle."""
    cfg = m.get_n_case_config("orszag_tang_unit", 1)
    cfg.nx = cfg.ny = n
    U0 = m.initial_state(4, cfg)
    say(f"[gpu] OT {n}^2 HLLD van Leer 3 steps")
    r = gpu_run(cfg, U0, 2, 1, 3)
    o = oracle.run_loop(U0, ocase(cfg), 2, 1, 3)
    assert_bitwise(r["dts"], o["dts"], "dt")
    assert rel_l1(r["U"], o["U"]) == 0.0
    assert_bitwise(r["U"], o["U"], "state")


def test_full_size_properties_4096():
    """OT 4096^2 (the headline shape): a uniform state is preserved bitwise and the
    y-uniform Brio-Wu rows stay identical, at full size."""
    n = 4096
    cfg = m.CaseConfig(n, n, 2, 0.0, 1.0, 0.0, 1.0, 0.8, 1.0, 5.0 / 3.0,
                       m.BoundaryConfig(*[m.BoundaryType.Periodic] * 4))
    U = np.empty(cfg.shape)
    U[...] = np.array([1.2, 0.3 * 1.2, -0.1 * 1.2, 0.05 * 1.2, 0.4, -0.2, 0.1, 3.0, 0.0])
    say("[gpu] uniform 4096^2 HLLD van Leer 2 steps")
    r = gpu_run(cfg, U, 2, 1, 2)
    assert (r["U"][..., :8] == U[..., :8]).all()
    # Brio-Wu stretched to 4096 x 64 rows: every row identical
    bw = m.get_n_case_config("brio_wu_transmissive", 1)
    bw.nx, bw.ny, bw.y_max = n, 64, 64.0 / n
    U0 = m.initial_state(6, bw)
    say("[gpu] Brio-Wu 4096x64 rows")
    r = gpu_run(bw, U0, 2, 0, 20)
    rows = r["U"][2:-2, 2:-2]
    assert (rows == rows[:1]).all()


def test_bandwidth_probe():
    """mhd_probe_bandwidth: the copy leaves b equal to a, the three rates are positive
    and below the 8 TB/s peak, and bad arguments are refused."""
    import torch

    from mhd_gpu_acc_amd._lib import check
    L = m.lib()
    n = 8 << 20  # doubles: 64 MiB
    a = torch.arange(n, dtype=torch.float64, device="cuda")
    b = torch.zeros_like(a)
    torch.cuda.synchronize()
    t = (C.c_double * 3)()
    check(L.mhd_probe_bandwidth(a.data_ptr(), b.data_ptr(), 8 * n, 2, t))
    say(f"[gpu] bandwidth probe on 64 MiB: copy {t[0]:.2f}, read {t[1]:.2f}, write {t[2]:.2f} TB/s")
    assert torch.equal(a, b)
    assert all(0.0 < t[k] < 8.0 for k in range(3))
    assert L.mhd_probe_bandwidth(a.data_ptr(), a.data_ptr(), 8 * n, 2, t) != 0
    assert L.mhd_probe_bandwidth(a.data_ptr(), b.data_ptr(), 1000, 2, t) != 0
