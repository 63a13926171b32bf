"""compute_dt's per-cell candidate as the kernels evaluate it (host build of
mhd_gpu_acc_amd/csrc/physics.hpp).  A cell whose signal speeds a division-free test
proves to be at most c_h returns c_h without evaluating the speeds, which is what
the exact evaluation (solver_cpu.cpp:291-300, physics.hpp:175-210) returns for it.
The shortcut must give the exact candidate bit for bit on every cell: states whose
speed lies within rounding of c_h on either side, the degenerate fast speed where
the discriminant vanishes (a^2 = b_n^2 = b^2: largest rounding of c_f), zero and
extreme magnitudes, and non-physical cells."""
from __future__ import annotations

import numpy as np

import oracle
from tests import hostcheck as hc

GAMMA = 5.0 / 3.0
OT_UNIT = 4  # case id of the unit-box Orszag-Tang vortex


def cons(V, gamma=GAMMA):
    U = np.zeros(9)
    oracle.orc().orc_prim_to_cons(oracle.ptr(np.ascontiguousarray(V, dtype=np.float64)), gamma, oracle.ptr(U))
    return U


def speed(U, gamma=GAMMA):
    """max(|u|+c_fx, |v|+c_fy) as the exact evaluation rounds it (candidate at c_h = 0)."""
    _, exact, _ = hc.cfl_candidates(U, gamma, 0.0)
    return exact


def check(U, gamma, ch):
    fast, exact, hit = hc.cfl_candidates(U, gamma, ch)
    same = (fast.view(np.uint64) == exact.view(np.uint64))
    assert same.all(), (np.flatnonzero(~same)[:5], fast[~same][:5], exact[~same][:5], ch)
    return hit


def random_prims(rng, n):
    V = np.empty((n, 9))
    V[:, 0] = rng.uniform(0.05, 5.0, n)
    V[:, 1:4] = rng.normal(0, 1.5, (n, 3))
    V[:, 4:7] = rng.normal(0, 1.0, (n, 3))
    V[:, 7] = rng.uniform(0.01, 5.0, n)
    V[:, 8] = rng.normal(0, 0.1, n)
    return V


def test_candidate_equals_exact_near_ch():
    rng = np.random.default_rng(11)
    V = random_prims(rng, 4000)
    U = np.array([cons(v) for v in V])
    s = speed(U)
    hits = 0
    # c_h just below, at and above each cell's own speed, through the 1e-7 margin
    for d in (-1e-6, -1e-9, -1e-15, 0.0, 1e-15, 1e-9, 5e-8, 9e-8, 1e-7, 1.1e-7, 2e-7, 1e-6, 1e-3, 0.5):
        for k in range(0, len(U), 500):
            sl = slice(k, k + 500)
            for ch in np.unique(s[sl] * (1.0 + d))[:40]:
                hits += check(U[sl], GAMMA, ch).sum()
    # and one c_h for all cells around the median speed: a mix of both outcomes
    for q in (0.1, 0.5, 0.9, 1.0):
        ch = float(np.quantile(s, q))
        h = check(U, GAMMA, ch)
        hits += h.sum()
        assert not h[s > ch].any()
    assert hits > 10000  # the shortcut was taken, not only the exact path


def test_degenerate_discriminant():
    """B along the normal with gamma p = B_n^2 (a = b_n, no tangential field): the
    discriminant (a^2+b^2)^2 - 4 a^2 b_n^2 cancels to rounding noise, the case whose
    rounded c_f deviates most from the exact one."""
    rng = np.random.default_rng(12)
    rows = []
    for _ in range(3000):
        rho = rng.uniform(0.1, 4.0)
        bn = rng.uniform(0.2, 3.0) * rng.choice([-1.0, 1.0])
        p = bn * bn / GAMMA * (1.0 + rng.choice([0.0, 1e-16, -1e-16, 3e-16, 1e-12]))
        # flow along the field at about the fast speed, so that this direction, not
        # the transverse one (c_f = sqrt(a^2 + b^2) there), sets the candidate
        un = abs(bn) / np.sqrt(rho) * rng.uniform(0.5, 1.5) * rng.choice([-1.0, 1.0])
        ut = rng.normal(0, 0.1)
        bt = rng.choice([0.0, 0.0, 1e-9])
        rows.append(cons([rho, un, ut, 0.0, bn, bt, 0.0, p, 0.0]))  # x-normal field
        rows.append(cons([rho, ut, un, 0.0, bt, bn, 0.0, p, 0.0]))  # y-normal field
    U = np.array(rows)
    s = speed(U)
    hits = 0
    # the rounded c_f exceeds the exact one by up to ~1e-8 here: c_h within that of
    # the rounded speed must still give the exact candidate (the test clears these
    # cells only from about 1e-6 above their speed)
    for d in (-6e-9, -4e-9, -3e-9, -2e-9, -1e-9, 0.0, 1e-9, 1e-8, 3e-8, 1e-7, 3e-7, 1e-6, 2e-6, 1e-5):
        for k in range(len(U)):  # each cell at c_h relative to its own speed
            hits += check(U[k:k + 1], GAMMA, s[k] * (1.0 + d)).sum()
    assert hits > 1000


def test_extreme_magnitudes_and_nonphysical():
    rng = np.random.default_rng(13)
    V = random_prims(rng, 600)
    rows = []
    for scale_r, scale_v, scale_b in ((1e-60, 1.0, 1e-30), (1e60, 1e-30, 1e30), (1.0, 1e50, 1e50),
                                      (1.0, 1e-50, 1e-50), (1e-150, 1e70, 1e-5), (1e120, 1.0, 1e60),
                                      (1.0, 1.0, 1e-170), (1.0, 1.0, 0.0), (1e-300, 1.0, 1.0)):
        W = V.copy()
        W[:, 0] *= scale_r
        W[:, 1:4] *= scale_v
        W[:, 4:7] *= scale_b
        W[:, 7] *= scale_r * scale_v * scale_v + scale_b * scale_b
        rows += [cons(w) for w in W]
    bad = V[:8].copy()
    bad[0, 0] = 0.0
    bad[1, 7] = -1.0
    bad[2, 0] = np.nan
    bad[3, 7] = np.inf
    bad[4, 1] = np.inf
    bad[5, 4] = np.nan
    rows += [cons(w) for w in bad]
    U = np.array(rows)
    with np.errstate(all="ignore"):
        s = speed(U)
    chs = [0.0, 1e-300, 1e-150, 1e-30, 1.0, 1e30, 1e60, 1e150, 1e300, np.inf]
    chs += list(np.quantile(s[np.isfinite(s) & (s > 0)], [0.05, 0.5, 0.95]))
    for ch in chs:
        for d in (0.0, 1e-7, 1e-3):
            check(U, GAMMA, ch * (1.0 + d))


def test_orszag_tang_state_matches_oracle_compute_dt():
    """The kernels' candidates on an evolved Orszag-Tang state reduce to the oracle's
    compute_dt c_h, and the shortcut covers most cells there."""
    case = oracle.case_config(OT_UNIT, 1)
    case.nx = case.ny = 64
    U0 = oracle.init_state(OT_UNIT, case)
    ref = oracle.run_loop(U0, case, 2, 1, 30)
    U, ch = ref["U"], ref["ch"]
    interior = U[2:-2, 2:-2].reshape(-1, 9)
    hit = check(interior, case.gamma, ch)
    fast, _, _ = hc.cfl_candidates(interior, case.gamma, ch)
    rc, _, ch_next = oracle.compute_dt(U, case, case.cfl, ch)
    assert rc == 0
    assert fast.max() == ch_next
    assert hit.mean() > 0.5, hit.mean()
