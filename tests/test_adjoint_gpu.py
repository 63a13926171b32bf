"""Discrete adjoint on the GPU (through the C ABI):

* one reverse step (mhd_adjoint_step_grids) equals the host build of the same
  arithmetic bit for bit, on grids that span several tiles and every BC mix;
* the checkpointed tape (forward with the c_h ratchet, segment recompute, reverse)
  equals the host chain bit for bit, and passes the dot-product test against the
  oracle's tangent-linear model (an independent derivation);
* the torch.autograd.Function gives the tape's gradient and, on a state without
  rounding-level ties, matches plain GPU central differences (dt / c_h frozen);
* the adjoint workload (Orszag-Tang 512^2, HLLD, van Leer, 200 steps, K = 20)
  matches the oracle's CPU reference gradient (tests/golden/adjoint_ot512.json:
  tangent model, and frozen-branch central differences at eps = 1e-6) on the 32
  seeded probes.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

import mhd_gpu_acc_amd as m
import oracle
from tests import hostcheck as hc
from tests.test_adjoint_host import perturbed_ot, small_case
from tests.util import assert_bitwise

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "adjoint_ot512.json")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def cfg_of(case):
    return m.CaseConfig(case.nx, case.ny, 2, case.x_min, case.x_max, case.y_min, case.y_max, case.cfl, case.t_end,
                        case.gamma, m.BoundaryConfig(*[m.BoundaryType(b) for b in case.bc]))


STEP_CASES = [(13, 10, (0, 0, 0, 0), 2, 1), (37, 29, (2, 1, 0, 0), 2, 0), (64, 40, (1, 1, 2, 2), 0, 1),
              (251, 7, (0, 0, 1, 2), 2, 1), (5, 90, (2, 2, 0, 0), 0, 0), (70, 70, (0, 0, 0, 0), 2, 1),
              (41, 23, (0, 0, 0, 0), 1, 1), (19, 66, (2, 1, 0, 0), 1, 0), (70, 31, (1, 1, 2, 2), 3, 1),
              (33, 47, (0, 0, 1, 2), 3, 0), (29, 23, (3, 3, 3, 3), 2, 1), (66, 9, (3, 1, 0, 0), 0, 1),
              (17, 70, (2, 3, 1, 3), 1, 1), (70, 40, (0, 0, 3, 2), 3, 1)]


@pytest.mark.gpu
@pytest.mark.parametrize("nx,ny,bc,solver,limiter", STEP_CASES)
def test_reverse_step_matches_host(nx, ny, bc, solver, limiter):
    import torch
    from mhd_gpu_acc_amd.adjoint import adjoint_step_grids, from_planes, to_planes
    say(f"[gpu] reverse step {nx}x{ny} bc={bc} solver={solver} limiter={limiter}")
    case = small_case(nx, ny, bc)
    U = perturbed_ot(case, nx + 7 * ny)
    rc, dt, ch = oracle.compute_dt(U, case, case.cfl, 0.0)
    assert rc == 0
    Utmp, _ = oracle.advance(U, case, dt, ch, solver, limiter)
    lam = np.random.default_rng(nx * ny).normal(0, 1, U.shape)
    expected = hc.adjoint_step(U, Utmp, case, dt, ch, solver, limiter, lam)
    with m.Run(cfg_of(case), solver, limiter) as run:
        dU, dT, dL = to_planes(U), to_planes(Utmp), to_planes(lam)
        dB = torch.empty_like(dL)
        torch.cuda.synchronize()
        adjoint_step_grids(run, dU.data_ptr(), dT.data_ptr(), dt, ch, dL.data_ptr(), dB.data_ptr())
        run.synchronize()
        got = from_planes(dL)
    assert_bitwise(got, expected, "reverse step")


@pytest.mark.gpu
@pytest.mark.parametrize("nx,ny,bc,solver,limiter", [(70, 70, (0, 0, 0, 0), 2, 1), (19, 66, (2, 1, 0, 0), 1, 0)])
def test_reverse_step_flat_addressing_matches_host(nx, ny, bc, solver, limiter, monkeypatch):
    """The flat-addressed reverse kernels (the path of grids above 4 GiB), forced."""
    monkeypatch.setenv("MHD_ADDR", "flat")
    test_reverse_step_matches_host(nx, ny, bc, solver, limiter)


def _oracle_chain(case, U0, nsteps, solver, limiter):
    states, tmps, dts, chs = [U0], [], [], []
    ch, U = 0.0, U0
    for _ in range(nsteps):
        rc, dt, ch = oracle.compute_dt(U, case, case.cfl, ch)
        assert rc == 0
        Utmp, U = oracle.advance(U, case, dt, ch, solver, limiter)
        states.append(U)
        tmps.append(Utmp)
        dts.append(dt)
        chs.append(ch)
    return states, tmps, np.array(dts), np.array(chs)


@pytest.mark.gpu
@pytest.mark.parametrize("solver,limiter,bc,K", [(2, 1, (0, 0, 0, 0), 5), (0, 0, (1, 2, 2, 1), 12),
                                                 (2, 0, (2, 2, 0, 0), 1), (1, 1, (0, 0, 0, 0), 4),
                                                 (3, 0, (1, 1, 2, 2), 6), (2, 1, (3, 1, 3, 2), 5),
                                                 (2, 0, (3, 3, 3, 3), 4)])
@pytest.mark.parametrize("sets", ["2", "1"])
def test_tape_matches_host_chain_and_tangent(solver, limiter, bc, K, sets, monkeypatch):
    """sets: the tape's segment buffer sets (2: the next segment is recomputed on a
    second stream during the current one's reverse steps; 1: one after the other)."""
    from mhd_gpu_acc_amd.adjoint import Tape
    monkeypatch.setenv("MHD_TAPE_SETS", sets)
    say(f"[gpu] tape 40x36 solver={solver} limiter={limiter} bc={bc} K={K} sets={sets}")
    nsteps = 12
    case = small_case(40, 36, bc)
    U0 = perturbed_ot(case, 31)
    states, tmps, dts, chs = _oracle_chain(case, U0, nsteps, solver, limiter)
    rng = np.random.default_rng(3)
    lam = rng.normal(0, 1, U0.shape)
    v = rng.normal(0, 1, U0.shape)
    with m.Run(cfg_of(case), solver, limiter) as run:
        run.set_state(U0)
        with Tape(run, nsteps, K) as tape:
            gdts, gchs = tape.forward()
            assert_bitwise(gdts, dts, "dt sequence")
            assert_bitwise(gchs, chs, "c_h sequence")
            assert_bitwise(run.get_state(), states[-1], "final state")
            g = tape.backward(lam)
            g2 = tape.backward(lam)  # a second backward is the same
    A = lam
    for s in range(nsteps - 1, -1, -1):
        A = hc.adjoint_step(states[s], tmps[s], case, dts[s], chs[s], solver, limiter, A)
    assert_bitwise(g, A, "tape vs host chain")
    assert_bitwise(g2, g, "repeated backward")
    _, dUN = oracle.tangent_run(U0, v, case, dts, chs, solver, limiter)
    lhs, rhs = np.sum(lam * dUN), np.sum(g * v)
    assert abs(lhs - rhs) <= 1e-11 * (np.sum(np.abs(lam * dUN)) + np.sum(np.abs(g * v))), (lhs, rhs)


@pytest.mark.gpu
def test_tape_argument_checks():
    from mhd_gpu_acc_amd.adjoint import Tape
    with m.Run(cfg_of(small_case(8, 8, (3, 0, 0, 0))), 2, 1) as run:  # Dirichlet sides are supported
        Tape(run, 4, 2).close()
        with pytest.raises(m.MHDError):
            Tape(run, 0, 2)
        with pytest.raises(m.MHDError):
            Tape(run, 4, 0)


def _loss_from_planes(UN, target):
    r = UN[0, 2:-2, 2:-2] - target
    return 0.5 * (r * r).sum()


def _fd(run, U0t, dts, chs, target, probe, eps):
    """Central difference of the loss along one (row, col, field) probe, dt / c_h frozen.
    L(+) - L(-) is formed as sum (r+ - r-)(r+ + r-)/2 to avoid cancellation."""
    from mhd_gpu_acc_amd.adjoint import replay
    j, i, f = probe
    Up, Um = U0t.clone(), U0t.clone()
    Up[f, j, i] += eps
    Um[f, j, i] -= eps
    rp = replay(run, Up, dts, chs)[0, 2:-2, 2:-2] - target
    rm = replay(run, Um, dts, chs)[0, 2:-2, 2:-2] - target
    return float(((rp - rm) * (rp + rm)).sum() * 0.5) / (2 * eps)


@pytest.mark.gpu
def test_autograd_function_and_finite_differences():
    """The Function's gradient is the tape's, and on a state without rounding-level
    ties (Orszag-Tang with seeded noise) it matches plain GPU central differences."""
    import torch
    from mhd_gpu_acc_amd.adjoint import Tape, from_planes, run_steps, to_planes
    say("[gpu] autograd noisy OT 48^2 HLLD van Leer 20 steps K=7")
    case = small_case(48, 48, (0, 0, 0, 0))
    U0 = perturbed_ot(case, 77)
    nsteps, K = 20, 7
    with m.Run(cfg_of(case), 2, 1) as run:
        run.set_state(U0 * np.r_[1.01, 1, 1, 1, 1, 1, 1, 1.01, 1])  # any target
        run.step_to(1.0, nsteps)
        target = to_planes(run.get_state())[0, 2:-2, 2:-2].clone()
        run.ch = 0.0
        U0t = to_planes(U0).requires_grad_(True)
        UN = run_steps(U0t, run, nsteps, K)
        loss = _loss_from_planes(UN, target)
        loss.backward()
        grad = from_planes(U0t.grad)
        assert np.all(np.isfinite(grad)) and np.any(grad != 0)
        # the same gradient through the tape API
        lam = np.zeros(case.shape)
        lam[2:-2, 2:-2, 0] = from_planes(UN.detach())[2:-2, 2:-2, 0] - target.cpu().numpy()
        run.ch = 0.0
        run.set_state(U0)
        with Tape(run, nsteps, K) as tape:
            dts, chs = tape.forward()
            g2 = tape.backward(lam)
        assert_bitwise(grad, g2, "autograd vs tape")
        base = to_planes(U0)
        probes = [(20, 30, 0), (40, 11, 1), (9, 45, 2), (33, 33, 4), (45, 20, 5), (12, 12, 7), (44, 44, 8),
                  (2, 40, 3), (30, 1, 0), (35, 49, 7)]
        adj = np.array([grad[j, i, f] for j, i, f in probes])
        fd = np.array([_fd(run, base, dts, chs, target, q, 1e-6) for q in probes])
    rel = np.abs(adj - fd) / np.maximum(np.abs(fd), 1e-3 * np.max(np.abs(fd)))
    say(f"  per-probe rel {np.array2string(rel, precision=1)}")
    assert np.linalg.norm(adj - fd) <= 1e-6 * np.linalg.norm(fd), rel


@pytest.mark.gpu
def test_autograd_repeated_backward_and_tape_reuse():
    """A second backward of one graph (retain_graph) gives the first gradient bit for bit;
    once a later forward has taken that graph's tape from the idle pool, a further
    backward of the old graph is refused with an error instead of using the new
    recording.  Two graphs alive at once each keep their own tape."""
    import torch
    from mhd_gpu_acc_amd.adjoint import run_steps, to_planes
    say("[gpu] autograd repeated backward, OT 32^2 HLLD van Leer 6 steps K=2")
    case = small_case(32, 32, (0, 0, 0, 0))
    U0 = perturbed_ot(case, 5)
    with m.Run(cfg_of(case), 2, 1) as run:
        run.ch = 0.0
        a = to_planes(U0).requires_grad_(True)
        la = (run_steps(a, run, 6, 2)[0] ** 2).sum()
        g1, = torch.autograd.grad(la, a, retain_graph=True)
        g2, = torch.autograd.grad(la, a, retain_graph=True)
        assert torch.equal(g1, g2)
        # two graphs alive at once: b takes the idle tape, c gets a tape of its own
        run.ch = 0.0
        b = to_planes(U0 * np.r_[1.001, 1, 1, 1, 1, 1, 1, 1, 1]).requires_grad_(True)
        lb = (run_steps(b, run, 6, 2)[0] ** 2).sum()
        run.ch = 0.0
        c = to_planes(U0).requires_grad_(True)
        lc = (run_steps(c, run, 6, 2)[0] ** 2).sum()
        gc, = torch.autograd.grad(lc, c)
        assert torch.equal(gc, g1)  # c recorded the same trajectory as a
        gb, = torch.autograd.grad(lb, b)
        assert not torch.equal(gb, g1)
        with pytest.raises(RuntimeError, match="reused by a later forward"):
            torch.autograd.grad(la, a)


@pytest.mark.gpu
def test_backward_of_an_older_graph_leaves_the_run_as_it_was():
    """Forward A, forward B (the c_h ratchet carried on from A), backward A: the run's
    c_h stays B's, and compute_dt and the next steps are what they would have been
    without A's backward; A's gradient is that of a fresh forward + backward."""
    import torch
    from mhd_gpu_acc_amd.adjoint import run_steps, to_planes
    say("[gpu] autograd: backward of an older graph, OT 32^2 HLLD van Leer 6 steps K=2")
    case = small_case(32, 32, (0, 0, 0, 0))
    U0 = perturbed_ot(case, 9)
    with m.Run(cfg_of(case), 2, 1) as run:
        run.ch = 0.0
        a = to_planes(U0).requires_grad_(True)
        la = (run_steps(a, run, 6, 2)[0] ** 2).sum()
        b = to_planes(U0 * np.r_[1.002, 1, 1, 1, 1, 1, 1, 1, 1]).requires_grad_(True)
        run_steps(b, run, 6, 2)  # no c_h reset: B continues A's ratchet
        dt_b = run.compute_dt()  # ratchets c_h for the next step
        ch_b, state_b = run.ch, run.get_state()
        ga, = torch.autograd.grad(la, a)
        assert run.ch == ch_b
        assert run.compute_dt() == dt_b
        assert_bitwise(run.get_state(), state_b, "state after the backward")
        after = run.step_to(1.0, 3)
        run.set_state(state_b)
        run.ch = ch_b
        control = run.step_to(1.0, 3)
        assert_bitwise(after["dts"], control["dts"], "dt sequence after the backward")
        run.ch = 0.0
        a2 = to_planes(U0).requires_grad_(True)
        ga2, = torch.autograd.grad((run_steps(a2, run, 6, 2)[0] ** 2).sum(), a2)
        assert torch.equal(ga, ga2)


@pytest.mark.gpu
def test_autograd_is_stream_ordered_on_a_side_stream():
    """The Function queues its work on torch's current stream: on a side stream, with no
    host synchronisation between the producing op, the forward, the loss and the
    backward, the gradient equals the default-stream one bit for bit, and the run is
    handed back to its own stream."""
    import torch
    from mhd_gpu_acc_amd.adjoint import run_steps, to_planes
    say("[gpu] autograd on a side stream, OT 40^2 HLLD van Leer 8 steps K=3")
    case = small_case(40, 40, (0, 0, 0, 0))
    U0 = perturbed_ot(case, 21)
    scale = torch.tensor(np.r_[1.0, 1, 1, 1, 1, 1, 1, 1, 1], dtype=torch.float64, device="cuda")[:, None, None]

    def grad_on(stream):
        with torch.cuda.stream(stream):
            base = to_planes(U0)
            x = (base * scale).requires_grad_(True)  # produced on this stream just before
            run.ch = 0.0
            loss = (run_steps(x, run, 8, 3)[0] ** 2).sum()
            g, = torch.autograd.grad(loss, x)
            out = g.clone()
        stream.synchronize()
        return out

    with m.Run(cfg_of(case), 2, 1) as run:
        own = run.stream
        g_default = grad_on(torch.cuda.current_stream())
        side = torch.cuda.Stream()
        g_side = grad_on(side)
        assert run.stream == own
    assert torch.equal(g_default, g_side)
    assert torch.isfinite(g_side).all() and (g_side != 0).any()


def _workload_gradient(run, U0, target, steps):
    from mhd_gpu_acc_amd.adjoint import from_planes, run_steps, to_planes
    run.ch = 0.0
    U0t = to_planes(U0).requires_grad_(True)
    UN = run_steps(U0t, run, steps, 20)
    loss = _loss_from_planes(UN, target)
    loss.backward()
    return float(loss.detach()), from_planes(U0t.grad)


@pytest.mark.gpu
def test_adjoint_workload_ot512():
    """The adjoint workload against the CPU references of tests/golden/adjoint_ot512.json:
    on the Orszag-Tang state the gradient equals the oracle's tangent model at the 32
    seeded probes and on a whole-grid projection, and the CPU adjoint bit for bit; on
    the generic (noise-perturbed) state it also matches frozen-branch central
    differences: the fourth-order stencil at eps = 1e-6, (4 D(eps) - D(2 eps)) / 3, to
    1e-6, and the two-point D(eps) at eps = 1e-6 and 1e-7 to the bound their
    truncation error O(eps^2) allows."""
    import hashlib
    from mhd_gpu_acc_amd.adjoint import adjoint_workload, generic_state, to_planes
    if not os.path.exists(GOLDEN):
        pytest.skip("adjoint_ot512.json not generated")
    ref = json.load(open(GOLDEN))
    w = ref["workload"]
    say("[gpu] adjoint workload OT 512^2 HLLD van Leer 200 steps K=20")
    cfg, U0, Ubump = adjoint_workload(w["nx"], w["seed"])
    assert (cfg.cfl, cfg.gamma, cfg.bc.as_list()) == (w["cfl"], w["gamma"], [0, 0, 0, 0])
    pairs = [tuple(q) for q in ref["pairs"]]
    v = np.random.default_rng(w["seed"]).standard_normal(U0.shape)
    with m.Run(cfg, 2, 1, cr=w["cr"]) as run:
        run.set_state(Ubump)
        assert run.step_to(1.0, w["steps"])["steps"] == w["steps"]
        target = to_planes(run.get_state())[0, 2:-2, 2:-2].clone()
        parts = {"exact": U0}
        if "generic" in ref:
            parts["generic"] = generic_state(U0, w["generic_seed"])
        for name, U in parts.items():
            r = ref[name]
            loss, grad = _workload_gradient(run, U, target, w["steps"])
            assert loss == pytest.approx(r["loss"], rel=1e-12)
            adj = np.array([grad[j, i, f] for j, i, f in pairs])
            tan = np.array(r["tangent"]["probes"])
            say(f"  {name}: vs CPU tangent max abs {np.max(np.abs(adj - tan)):.3e} (max |g| {np.max(np.abs(tan)):.3e})")
            assert np.allclose(adj, tan, rtol=1e-9, atol=1e-12 * np.max(np.abs(tan)))
            assert float(np.sum(grad * v)) == pytest.approx(r["tangent"]["direction_dot"], rel=1e-9)
            if "cpu_adjoint" in r:
                assert hashlib.sha256(np.ascontiguousarray(grad).tobytes()).hexdigest() == r["cpu_adjoint"]["sha256"]
            ffd = r.get("frozen_fd", {})
            for key, bound in (("1e-06", 3e-6), ("1e-07", 3e-7)):
                if key in ffd:
                    fd = np.array(ffd[key])
                    nrel = np.linalg.norm(adj - fd) / np.linalg.norm(fd)
                    say(f"  {name}: vs frozen-branch central differences eps {key}: norm rel {nrel:.3e}")
                    assert nrel <= bound
            if "1e-06" in ffd and "2e-06" in ffd:
                fd4 = (4.0 * np.array(ffd["1e-06"]) - np.array(ffd["2e-06"])) / 3.0
                nrel = np.linalg.norm(adj - fd4) / np.linalg.norm(fd4)
                say(f"  {name}: vs fourth-order frozen-branch central differences eps 1e-6: norm rel {nrel:.3e}")
                assert nrel <= 1e-6
