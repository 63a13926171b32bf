"""Initial states computed on the device (mhd_run_init_state, mhd_init_state_device,
mhd_batch_init_states) are bitwise the host initial states (mhd_init_state_host), for
every case that has one, on odd and even grid sizes, and for the whole 64-run batch
family; cases whose host state needs the C library's transcendentals are refused."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import mhd_gpu_acc_amd as m
from mhd_gpu_acc_amd._lib import MHDError
from mhd_gpu_acc_amd.batch import Batch, ot_family_batch, plan_ot_family
from tests.util import assert_bitwise


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _family_params(k=0):
    sizes, params = plan_ot_family(20261203)
    return params[k]


def _cases():
    sc = m.ICParams()
    sc.seed = 987654321
    return [
        ("brio_wu", m.CaseId.BrioWu, (800, 4), None),
        ("brio_wu_t", m.CaseId.BrioWuTransmissive, (203, 5), None),
        ("shock_bubble", m.CaseId.ShockBubble, (130, 33), None),
        ("shock_cloud", m.CaseId.ShockCloud, (256, 128), sc),
        ("shock_cloud odd", m.CaseId.ShockCloud, (101, 67), sc),
        ("ot_family", m.CaseId.OrszagTangFamily, (128, 128), _family_params(0)),
        ("ot_family odd", m.CaseId.OrszagTangFamily, (97, 61), _family_params(5)),
    ]


@pytest.mark.gpu
@pytest.mark.parametrize("name,case,shape,params", _cases(), ids=[c[0] for c in _cases()])
def test_run_init_state_bitwise_host(name, case, shape, params):
    import torch
    cfg = m.get_n_case_config(case, 1)
    cfg.nx, cfg.ny = shape
    if params is not None and case == m.CaseId.OrszagTangFamily:
        cfg.gamma = params.gamma
    say(f"[gpu] device IC {name} {shape}")
    host = m.initial_state(case, cfg, params)
    with m.Run(cfg, m.RiemannSolver.HLLD, m.Limiter.MINMOD) as run:
        run.init_state(case, params)
        assert_bitwise(run.get_state(), host, f"{name}: run state")
        # the three grids start equal: a step from the device state equals one from the host state
        res = run.step_to(cfg.t_end, 3)
        U_dev = run.get_state()
        run.set_state(host)
        run.ch = 0.0
        res2 = run.step_to(cfg.t_end, 3)
        assert_bitwise(U_dev, run.get_state(), f"{name}: 3 steps")
        assert_bitwise(res["dts"], res2["dts"], f"{name}: dt")
    # into a caller's device grid (9 planes, dense)
    T = torch.full((9, cfg.ny + 4, cfg.nx + 4), np.nan, dtype=torch.float64, device="cuda")
    m.init_state_device(case, cfg, params, T.data_ptr())
    torch.cuda.synchronize()
    assert_bitwise(np.moveaxis(T.cpu().numpy(), 0, -1), host, f"{name}: caller grid")


@pytest.mark.gpu
def test_batch_init_states_family_bitwise_host():
    """The full 64-run Orszag-Tang family (8.9 M cells) initialised by one launch."""
    say("[gpu] batch device IC, 64-run OT family")
    b, cfgs, params, states = ot_family_batch(20261203, device_init=True)
    assert states is None
    try:
        for r, (c, p) in enumerate(zip(cfgs, params)):
            assert_bitwise(b.get_state(r), m.initial_state(m.CaseId.OrszagTangFamily, c, p), f"run {r}")
        # stepping from device states equals stepping from host states
        dts = b.step(2)
        b2, _, _, _ = ot_family_batch(20261203, scale=1)
        try:
            dts2 = b2.step(2)
            assert_bitwise(dts, dts2, "dt tables")
            for r in (0, 7, 63):
                assert_bitwise(b.get_state(r), b2.get_state(r), f"run {r} after 2 steps")
        finally:
            b2.close()
    finally:
        b.close()


@pytest.mark.gpu
def test_batch_init_states_mixed_sizes_and_shock_cloud():
    cfgs = []
    for nx, ny in ((64, 32), (50, 70), (128, 16)):
        c = m.get_n_case_config(m.CaseId.ShockCloud, 1)
        c.nx, c.ny = nx, ny
        cfgs.append(c)
    ps = []
    for k in range(3):
        p = m.ICParams()
        p.seed = 1000 + k
        ps.append(p)
    say("[gpu] batch device IC, shock-cloud 3 sizes")
    with Batch(cfgs, m.RiemannSolver.HLLD, m.Limiter.VAN_LEER) as b:
        b.init_states(m.CaseId.ShockCloud, ps)
        for r, (c, p) in enumerate(zip(cfgs, ps)):
            assert_bitwise(b.get_state(r), m.initial_state(m.CaseId.ShockCloud, c, p), f"run {r}")


@pytest.mark.gpu
@pytest.mark.parametrize("case", [m.CaseId.KelvinHelmholtz, m.CaseId.OrszagTang, m.CaseId.OrszagTangUnit,
                                  m.CaseId.OrszagTangBump])
def test_device_init_refuses_libm_cases(case):
    cfg = m.get_n_case_config(m.CaseId.OrszagTangUnit, 1)
    cfg.nx = cfg.ny = 32
    p = m.ICParams()
    with m.Run(cfg, m.RiemannSolver.HLLD, m.Limiter.MINMOD) as run:
        with pytest.raises(MHDError, match="no device initial state"):
            run.init_state(case, p)
    with pytest.raises(MHDError):
        with Batch([cfg], m.RiemannSolver.HLLD, m.Limiter.MINMOD) as b:
            b.init_states(case, [p])
    with m.Run(cfg, m.RiemannSolver.HLLD, m.Limiter.MINMOD) as run:  # family without parameters
        with pytest.raises(MHDError, match="parameters"):
            run.init_state(m.CaseId.OrszagTangFamily, None)
