"""Hand-written adjoint, host build (CPU): dot-product tests against the oracle's
tangent-linear model, which is the oracle's own code instantiated on forward-mode
dual numbers (an independent derivation).  For a map y = f(x) with tangent
dy = J dx and adjoint xb = J^T yb, <yb, J dx> must equal <J^T yb, dx> to rounding
for every (dx, yb); a dropped term, a wrong sign or a transposed operand in any
hand-written derivative breaks the equality for random directions."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests import hostcheck as hc

GAMMA = 5.0 / 3.0


def close(lhs, rhs, scale, rtol=1e-11):
    assert np.isfinite(lhs) and np.isfinite(rhs), (lhs, rhs)
    assert abs(lhs - rhs) <= rtol * scale + 1e-300, (lhs, rhs, scale)


def prims(rng, n=1):
    out = []
    for _ in range(n):
        out.append(np.array([rng.uniform(0.5, 2.0), *rng.normal(0, 1, 3), *rng.normal(0, 1, 3),
                             rng.uniform(0.3, 3.0), rng.normal(0, 0.1)]))
    return out


def cons(V, gamma=GAMMA):
    U = np.zeros(9)
    oracle.orc().orc_prim_to_cons(oracle.ptr(np.ascontiguousarray(V)), gamma, oracle.ptr(U))
    return U


def riemann_pairs(rng, direction):
    """(regime, UL, UR) pairs covering the HLLD regions and fallbacks."""
    n = 1 if direction == 0 else 2  # slot of the normal velocity
    bn = 4 if direction == 0 else 5
    pairs = []
    for k in range(60):
        VL, VR = prims(rng, 2)
        pairs.append(("generic", VL, VR))
    for sgn in (1.0, -1.0):
        for k in range(8):
            VL, VR = prims(rng, 2)
            VL[n] = VR[n] = 8.0 * sgn
            pairs.append(("supersonic", VL, VR))
    for k in range(12):
        VL = prims(rng)[0]
        VR = VL + rng.normal(0, 0.05, 9)
        pairs.append(("weak", VL, VR))
    for k in range(12):
        VL, VR = prims(rng, 2)
        VL[bn] = VR[bn] = 0.0
        pairs.append(("bn_zero", VL, VR))
    for k in range(12):
        VL, VR = prims(rng, 2)
        VL[bn] = VR[bn] = 2.5
        for s in range(4, 7):
            if s != bn:
                VL[s] = VR[s] = 0.0
        pairs.append(("field_aligned", VL, VR))
    for k in range(8):
        VL = prims(rng)[0]
        pairs.append(("uniform", VL, VL.copy()))
    for k in range(8):
        VL, VR = prims(rng, 2)
        VL[4:7] *= 6.0
        pairs.append(("strong_field", VL, VR))
    out = []
    for name, VL, VR in pairs:
        out.append((name, cons(VL), cons(VR)))
    for k in range(6):  # negative pressure on one side: HLLD falls back to HLL, HLL to the mean
        VL, VR = prims(rng, 2)
        UL = cons(VL)
        UL[7] -= 3.0 * (UL[7] - 0.5 * VL[0] * np.sum(VL[1:4] ** 2) - 0.5 * np.sum(VL[4:7] ** 2)) + 1.0
        out.append(("unphysical", UL, cons(VR)))
    return out


@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("solver", [0, 1, 2, 3])
def test_riemann_adjoint_dot(direction, solver):
    rng = np.random.default_rng(11 + direction + 10 * solver)
    for ch in (0.0, 0.7, 3.0):
        for name, UL, UR in riemann_pairs(rng, direction):
            dUL, dUR, Fb = rng.normal(0, 1, (3, 9))
            F, dF = oracle.tan_riemann(UL, dUL, UR, dUR, direction, solver, ch, GAMMA)
            F0 = np.zeros(9)
            oracle.orc().orc_riemann_flux(oracle.ptr(UL), oracle.ptr(UR), direction, solver, ch, GAMMA,
                                          oracle.ptr(F0))
            assert np.array_equal(F, F0)
            ULb, URb = hc.riemann_adj(UL, UR, direction, solver, ch, GAMMA, Fb)
            lhs = Fb @ dF
            rhs = ULb @ dUL + URb @ dUR
            scale = np.abs(Fb) @ np.abs(dF) + np.abs(ULb) @ np.abs(dUL) + np.abs(URb) @ np.abs(dUR)
            close(lhs, rhs, scale)


def test_riemann_adjoint_matches_finite_differences():
    """Spot check of the tangent itself on smooth points: central differences."""
    rng = np.random.default_rng(5)
    for name, UL, UR in riemann_pairs(rng, 0)[:40]:
        dUL, dUR = rng.normal(0, 1, (2, 9))
        F, dF = oracle.tan_riemann(UL, dUL, UR, dUR, 0, 2, 0.9, GAMMA)
        eps = 1e-6
        Fp = np.zeros(9)
        Fm = np.zeros(9)
        a, b = UL + eps * dUL, UR + eps * dUR
        oracle.orc().orc_riemann_flux(oracle.ptr(a), oracle.ptr(b), 0, 2, 0.9, GAMMA, oracle.ptr(Fp))
        a, b = UL - eps * dUL, UR - eps * dUR
        oracle.orc().orc_riemann_flux(oracle.ptr(a), oracle.ptr(b), 0, 2, 0.9, GAMMA, oracle.ptr(Fm))
        fd = (Fp - Fm) / (2 * eps)
        # a region switch inside [-eps, eps] would show as a large gap; smooth points agree
        if np.max(np.abs(fd - dF)) > 1e-5 * (1 + np.max(np.abs(dF))):
            continue
        assert np.allclose(fd, dF, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("direction", [0, 1])
def test_flux_and_conversions_adjoint_dot(direction):
    rng = np.random.default_rng(3 + direction)
    for V in prims(rng, 40):
        U = cons(V)
        dU, Fb, dV, Vb = rng.normal(0, 1, (4, 9))
        F, dF = oracle.tan_flux(U, dU, direction, 0.8, GAMMA)
        Ub = hc.flux_adj(U, direction, 0.8, GAMMA, Fb)
        close(Fb @ dF, Ub @ dU, np.abs(Fb) @ np.abs(dF) + np.abs(Ub) @ np.abs(dU))
        _, dW = oracle.tan_cons_to_prim(U, dU, GAMMA)
        Ub = hc.cons_to_prim_adj(U, GAMMA, Vb)
        close(Vb @ dW, Ub @ dU, np.abs(Vb) @ np.abs(dW) + np.abs(Ub) @ np.abs(dU))
        _, dUc = oracle.tan_prim_to_cons(V, dV, GAMMA)
        Vb2 = hc.prim_to_cons_adj(V, GAMMA, Fb)
        close(Fb @ dUc, Vb2 @ dV, np.abs(Fb) @ np.abs(dUc) + np.abs(Vb2) @ np.abs(dV))
        c, dc = oracle.tan_fast_speed(V, dV, direction, GAMMA)
        Vb3 = hc.fast_speed_adj(V, direction, GAMMA, 1.7)
        close(1.7 * dc, Vb3 @ dV, abs(1.7 * dc) + np.abs(Vb3) @ np.abs(dV))


def test_fast_speed_degenerate_discriminant():
    """a^2 = b^2 with B along the normal: the clamped discriminant is 0 and both the
    tangent and the adjoint take sqrt's derivative there as 0 (finite, equal)."""
    V = np.array([1.0, 0.1, 0.2, 0.0, 1.0, 0.0, 0.0, 0.6, 0.0])  # gamma p = |B|^2 -> a^2 = b^2 = 1
    dV = np.linspace(0.1, 0.9, 9)
    c, dc = oracle.tan_fast_speed(V, dV, 0, GAMMA)
    Vb = hc.fast_speed_adj(V, 0, GAMMA, 1.0)
    assert np.isfinite(dc) and np.all(np.isfinite(Vb))
    close(dc, Vb @ dV, abs(dc) + np.abs(Vb) @ np.abs(dV))


@pytest.mark.parametrize("limiter", [0, 1])
def test_limiter_adjoint(limiter):
    rng = np.random.default_rng(7)
    pairs = [(1.0, 2.0), (2.0, 1.0), (-1.0, -3.0), (-3.0, -1.0), (1.0, -1.0), (0.0, 2.0), (1.5, 1.5),
             (-0.5, -0.5)] + [tuple(x) for x in rng.normal(0, 1, (30, 2))]
    for a, b in pairs:
        da, db = rng.normal(0, 1, 2)
        v, dv = oracle.tan_limiter(a, da, b, db, limiter)
        assert v == oracle.orc().orc_limiter(a, b, limiter)
        ab, bb = hc.limiter_adj(a, b, limiter, 1.3)
        close(1.3 * dv, ab * da + bb * db, abs(1.3 * dv) + abs(ab * da) + abs(bb * db))


@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("limiter", [0, 1])
def test_reconstruct_adjoint_dot(direction, limiter):
    rng = np.random.default_rng(17 + direction + 2 * limiter)
    cases = []
    for k in range(50):
        base = prims(rng)[0]
        cases.append([base + rng.normal(0, 0.2, 9) * np.r_[0.3, 1, 1, 1, 1, 1, 1, 0.3, 0.1] for _ in range(3)])
    for k in range(6):  # a pressure jump that makes one extrapolated state unphysical (fallback to Wc)
        W = prims(rng, 3)
        W[0][7], W[1][7], W[2][7] = 0.05, 1.0, 8.0
        cases.append(W)
    for W in cases:
        for dtd in (0.02, 0.4, 3.0):  # the largest pushes half steps out of the physical set
            dWm, dWc, dWp, Lb, Rb = rng.normal(0, 1, (5, 9))
            L, dL, R, dR = oracle.tan_reconstruct(W[0], dWm, W[1], dWc, W[2], dWp, dtd, direction, 0.8, GAMMA,
                                                  limiter)
            Wmb, Wcb, Wpb = hc.reconstruct_adj(W[0], W[1], W[2], dtd, direction, 0.8, GAMMA, limiter, Lb, Rb)
            lhs = Lb @ dL + Rb @ dR
            rhs = Wmb @ dWm + Wcb @ dWc + Wpb @ dWp
            scale = (np.abs(Lb) @ np.abs(dL) + np.abs(Rb) @ np.abs(dR) + np.abs(Wmb) @ np.abs(dWm)
                     + np.abs(Wcb) @ np.abs(dWc) + np.abs(Wpb) @ np.abs(dWp))
            close(lhs, rhs, scale)


def small_case(nx, ny, bc, gamma=GAMMA):
    return oracle.Case(nx, ny, 2, 0.0, 1.0, 0.0, 1.0, 0.8, 1.0, gamma, tuple(bc))


def perturbed_ot(case, seed):
    """Unit-box Orszag-Tang on the case's grid with seeded noise (breaks the IC's
    symmetries, so no face sits exactly on a branch tie)."""
    U = oracle.init_state(4, case)
    rng = np.random.default_rng(seed)
    U[..., 0] *= 1.0 + 0.05 * rng.uniform(-1, 1, U.shape[:2])
    U[..., 1:7] += 0.02 * rng.normal(0, 1, U.shape[:2] + (6,))
    U[..., 8] = 0.01 * rng.normal(0, 1, U.shape[:2])
    return U


BCS = [(0, 0, 0, 0), (1, 1, 1, 1), (2, 2, 2, 2), (0, 0, 2, 1), (2, 1, 0, 0), (1, 2, 2, 0),
       # Dirichlet sides: ghosts copied from the previous grid (alone, beside each other type)
       (3, 3, 3, 3), (3, 1, 0, 0), (0, 0, 3, 2), (2, 3, 1, 3), (1, 3, 3, 1)]


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("solver,limiter", [(2, 1), (2, 0), (0, 1), (0, 0), (1, 1), (1, 0), (3, 1), (3, 0)])
def test_step_adjoint_dot(bc, solver, limiter):
    case = small_case(13, 10, bc)
    U = perturbed_ot(case, 100 + bc[0] + 3 * bc[2])
    rc, dt, ch = oracle.compute_dt(U, case, case.cfl, 0.0)
    assert rc == 0
    Utmp, Unew = oracle.advance(U, case, dt, ch, solver, limiter)
    rng = np.random.default_rng(1)
    for _ in range(2):
        v = rng.normal(0, 1, U.shape)
        lam = rng.normal(0, 1, U.shape)
        U1, dU1 = oracle.tangent_run(U, v, case, [dt], [ch], solver, limiter)
        assert np.array_equal(U1, Unew)
        A = hc.adjoint_step(U, Utmp, case, dt, ch, solver, limiter, lam)
        close(np.sum(lam * dU1), np.sum(A * v), np.sum(np.abs(lam * dU1)) + np.sum(np.abs(A * v)))


def test_step_adjoint_ghost_rows_and_corners_zero():
    """The input's y-ghost rows and corners are overwritten by the boundary fill
    before anything reads them: their adjoint is exactly zero."""
    case = small_case(9, 7, (0, 0, 0, 0))
    U = perturbed_ot(case, 4)
    rc, dt, ch = oracle.compute_dt(U, case, case.cfl, 0.0)
    Utmp, _ = oracle.advance(U, case, dt, ch, 2, 1)
    lam = np.random.default_rng(2).normal(0, 1, U.shape)
    A = hc.adjoint_step(U, Utmp, case, dt, ch, 2, 1, lam)
    assert np.all(A[:2] == 0) and np.all(A[-2:] == 0)
    assert np.any(A[2:-2, :2] != 0)  # x ghosts are read by the x sweep


@pytest.mark.parametrize("solver,limiter,bc", [(2, 1, (0, 0, 0, 0)), (0, 0, (1, 2, 0, 0)), (1, 1, (0, 0, 2, 2)),
                                               (3, 0, (2, 1, 0, 0)), (2, 1, (3, 1, 3, 0)), (1, 0, (3, 3, 3, 3))])
def test_multistep_adjoint_dot(solver, limiter, bc):
    """Five steps with the c_h ratchet and damping: reverse steps chained over the
    stored states against the tangent run with the recorded dt / c_h."""
    case = small_case(12, 12, bc)
    U0 = perturbed_ot(case, 9)
    states, tmps, dts, chs = [U0], [], [], []
    ch = 0.0
    U = U0
    for _ in range(5):
        rc, dt, ch = oracle.compute_dt(U, case, case.cfl, ch)
        Utmp, U = oracle.advance(U, case, dt, ch, solver, limiter)
        states.append(U)
        tmps.append(Utmp)
        dts.append(dt)
        chs.append(ch)
    rng = np.random.default_rng(8)
    v = rng.normal(0, 1, U0.shape)
    lam = rng.normal(0, 1, U0.shape)
    UN, dUN = oracle.tangent_run(U0, v, case, dts, chs, solver, limiter)
    assert np.array_equal(UN, states[-1])
    A = lam
    for s in range(4, -1, -1):
        A = hc.adjoint_step(states[s], tmps[s], case, dts[s], chs[s], solver, limiter, A)
    close(np.sum(lam * dUN), np.sum(A * v), np.sum(np.abs(lam * dUN)) + np.sum(np.abs(A * v)))


def test_unsupported_configurations_rejected():
    U = perturbed_ot(small_case(6, 6, (0, 0, 0, 0)), 1)
    with pytest.raises(ValueError):  # no solver 4
        hc.adjoint_step(U, U, small_case(6, 6, (0, 0, 0, 0)), 1e-3, 1.0, 4, 1, U)
    with pytest.raises(ValueError):  # no boundary type 4
        hc.adjoint_step(U, U, small_case(6, 6, (4, 0, 0, 0)), 1e-3, 1.0, 2, 1, U)


def test_dirichlet_ghost_adjoint_reaches_the_previous_grid():
    """Through Dirichlet sides the input's ghost rows and columns feed the next state
    (their values are copied into U* and on into U_next): their adjoint is not zero,
    and a perturbation of one of them moves the tangent as the adjoint says."""
    case = small_case(9, 7, (3, 3, 3, 3))
    U = perturbed_ot(case, 4)
    rc, dt, ch = oracle.compute_dt(U, case, case.cfl, 0.0)
    Utmp, _ = oracle.advance(U, case, dt, ch, 2, 1)
    lam = np.random.default_rng(2).normal(0, 1, U.shape)
    A = hc.adjoint_step(U, Utmp, case, dt, ch, 2, 1, lam)
    assert np.any(A[:2] != 0) and np.any(A[-2:] != 0) and np.any(A[:, :2] != 0)
    for (j, i, f) in [(0, 5, 0), (1, 1, 7), (10, 3, 4), (5, 11, 2), (0, 0, 8)]:
        e = np.zeros_like(U)
        e[j, i, f] = 1.0
        _, dU1 = oracle.tangent_run(U, e, case, [dt], [ch], 2, 1)
        close(np.sum(lam * dU1), A[j, i, f], np.sum(np.abs(lam * dU1)) + abs(A[j, i, f]))


def test_frozen_branch_differences_match_tangent():
    """The oracle's frozen-branch central differences: the unperturbed copy is the
    double run bit for bit, and on the unperturbed Orszag-Tang state (uniform
    pressure with rounding noise, slopes exactly zero along the axis each IC
    component is constant on) they converge to the tangent model: the error
    shrinks like eps^2 down to rounding, with no eps-independent gap, which a
    branch or derivative convention not shared by the two would leave."""
    case = small_case(24, 24, (0, 0, 0, 0))
    U0 = oracle.init_state(4, case)
    r = oracle.run_loop(U0, case, 2, 1, 6, t_end=1.0)
    probes = [(5, 7, 0), (12, 3, 1), (20, 20, 4), (9, 15, 7), (14, 2, 8), (3, 25, 5)]
    w = np.random.default_rng(4).normal(0, 1, U0.shape[:2])
    tan = []
    for j, i, f in probes:
        e = np.zeros_like(U0)
        e[j, i, f] = 1.0
        _, dU = oracle.tangent_run(U0, e, case, r["dts"], r["chs"], 2, 1)
        tan.append(np.sum(w * dU[..., 0]))
    best = np.full(len(probes), np.inf)
    for eps in (1e-5, 1e-6, 1e-7, 1e-8):
        rho0, rho_pm = oracle.frozen_fd(U0, probes, eps, case, r["dts"], r["chs"], 2, 1)
        assert np.array_equal(rho0, r["U"][..., 0])
        for q in range(len(probes)):
            fd = np.sum(w * (rho_pm[q, 0] - rho_pm[q, 1])) / (2 * eps)
            best[q] = min(best[q], abs(fd - tan[q]) / abs(tan[q]))
    assert np.all(best <= 1e-6), best


@pytest.mark.parametrize("solver,limiter", [(2, 1), (0, 0), (1, 1), (3, 0)])
def test_step_adjoint_dot_uniform_pressure(solver, limiter):
    """The unperturbed Orszag-Tang state: uniform density and pressure, the latter with
    rounding noise from the conserved variables, so the rounding-level slope rule of
    the derivative is exercised everywhere; the adjoint and the tangent model share it."""
    case = small_case(20, 16, (0, 0, 0, 0))
    U = oracle.init_state(4, case)
    rc, dt, ch = oracle.compute_dt(U, case, case.cfl, 0.0)
    Utmp, Unew = oracle.advance(U, case, dt, ch, solver, limiter)
    rng = np.random.default_rng(6)
    for _ in range(3):
        v = rng.normal(0, 1, U.shape)
        lam = rng.normal(0, 1, U.shape)
        _, dU1 = oracle.tangent_run(U, v, case, [dt], [ch], solver, limiter)
        A = hc.adjoint_step(U, Utmp, case, dt, ch, solver, limiter, lam)
        close(np.sum(lam * dU1), np.sum(A * v), np.sum(np.abs(lam * dU1)) + np.sum(np.abs(A * v)))


def test_adjoint_workload_fixture_consistent():
    """tests/golden/adjoint_ot512.json (CPU references of the 512^2, 200-step workload):
    the CPU adjoint agrees with the independent tangent model at every probe and on
    the whole-grid projection, and on the generic state the frozen-branch central
    differences converge to the tangent (error ~ eps^2 down to rounding)."""
    import json
    import os
    ref = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "adjoint_ot512.json")))
    ex, ge = ref["exact"], ref["generic"]
    tan = np.array(ex["tangent"]["probes"])
    cpu = np.array(ex["cpu_adjoint"]["probes"])
    assert np.max(np.abs(cpu - tan)) <= 1e-12 * np.max(np.abs(tan))
    assert abs(ex["cpu_adjoint"]["direction_dot"] - ex["tangent"]["direction_dot"]) <= 1e-12 * abs(
        ex["tangent"]["direction_dot"])
    gt = np.array(ge["tangent"]["probes"])
    e6 = np.linalg.norm(np.array(ge["frozen_fd"]["1e-06"]) - gt) / np.linalg.norm(gt)
    e7 = np.linalg.norm(np.array(ge["frozen_fd"]["1e-07"]) - gt) / np.linalg.norm(gt)
    assert e6 <= 3e-6 and e7 <= 3e-7 and e7 < e6 / 5, (e6, e7)
    if "2e-06" in ge["frozen_fd"]:
        # D(2 eps) - D(eps) ~ 3 c eps^2: the two-point error is the truncation term,
        # which the fourth-order combination removes
        d1, d2 = np.array(ge["frozen_fd"]["1e-06"]), np.array(ge["frozen_fd"]["2e-06"])
        e2 = np.linalg.norm(d2 - gt) / np.linalg.norm(gt)
        e4 = np.linalg.norm((4.0 * d1 - d2) / 3.0 - gt) / np.linalg.norm(gt)
        assert 3.0 < e2 / e6 < 5.0, (e2, e6)
        assert e4 <= 1e-7, e4
