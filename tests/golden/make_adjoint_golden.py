"""Writes adjoint_ot512.json: CPU reference values for the adjoint workload.

Workload: unit-box Orszag-Tang, 512^2, periodic, gamma 5/3, CFL 0.8, c_r 0.18,
HLLD + van Leer, 200 steps.  Target density: the same run from the initial state
plus a seeded Gaussian density bump (amplitude 0.02, sigma 0.05, p held, E
recomputed).  Loss L = 1/2 sum over interior cells (rho_N - rho_target)^2.  All
references freeze dt / c_h at the recorded sequence.

Two initial states:

* "exact": the Orszag-Tang state itself.  References: the oracle's tangent-linear
  model (its forward code on forward-mode dual numbers) for 32 seeded (cell, field)
  probes and for one seeded Gaussian direction v over the whole grid; and the CPU
  adjoint -- the host build of the hand-written adjoint (libmhd_hostcheck.so)
  chained over the oracle's stored trajectory -- as the SHA-256 of the whole
  gradient (the device adjoint equals it bit for bit), its probe values and its
  projection on v.  This state sits on exact ties of the scheme (slopes exactly
  zero along the axis a field is constant on, B_n = 0 lines, equal wave speeds),
  so no finite difference is well posed there.
* "generic": the same state with seeded noise of relative size 1e-3
  (mhd_gpu_acc_amd.adjoint.generic_state(U0, 7)).  References: the tangent model at
  the probes, and frozen-branch central differences D(eps) at eps = 1e-6, 1e-7 and
  2e-6 (the oracle's forward code on a scalar that carries the perturbed copies
  through every branch of the unperturbed run: finite differences with dt / c_h and
  every branch code frozen).  D(1e-6) and D(2e-6) make the fourth-order central
  difference at eps = 1e-6, (4 D(eps) - D(2 eps)) / 3.

About 75 min on 8 cores.  Parts already present in an existing adjoint_ot512.json
are kept unless --force is given.

    python tests/golden/make_adjoint_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from mhd_gpu_acc_amd.adjoint import bump_params, generic_state, probe_pairs  # noqa: E402
from tests import hostcheck  # noqa: E402

SEED = 20261203
GENERIC_SEED = 7
N, STEPS, SOLVER, LIMITER = 512, 200, 2, 1
OUT = os.path.join(ROOT, "tests", "golden", "adjoint_ot512.json")
T0 = time.time()


def say(msg):
    print(f"{msg}  [{time.time() - T0:.0f} s]", flush=True)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    prev = json.load(open(OUT)) if os.path.exists(OUT) and "--force" not in sys.argv else {}
    case = oracle.Case(N, N, 2, 0.0, 1.0, 0.0, 1.0, 0.8, 1.0, 5.0 / 3.0, (0, 0, 0, 0))
    amp, sigma, cx, cy = bump_params(SEED)
    p = oracle.ICParams()
    p.bump_amp, p.bump_sigma, p.bump_cx, p.bump_cy = amp, sigma, cx, cy
    tgt = oracle.run_loop(oracle.init_state(8, case, p), case, SOLVER, LIMITER, STEPS, t_end=1.0)
    assert tgt["nsteps"] == STEPS
    rho_t = tgt["U"][2:-2, 2:-2, 0]
    pairs = probe_pairs(SEED, N, N)
    U0 = oracle.init_state(4, case)
    out = dict(
        workload=dict(nx=N, ny=N, steps=STEPS, solver="hlld", limiter="vanleer", cfl=0.8, cr=0.18,
                      gamma=5.0 / 3.0, seed=SEED, generic_seed=GENERIC_SEED,
                      bump=dict(amp=amp, sigma=sigma, cx=cx, cy=cy)),
        pairs=[list(map(int, q)) for q in pairs],
        direction="numpy default_rng(seed).standard_normal(state shape (ny+4, nx+4, 9))",
    )
    out["exact"] = part(prev.get("exact", {}), U0, case, rho_t, pairs, cpu_adj=True, fd=False)
    out["generic"] = part(prev.get("generic", {}), generic_state(U0, GENERIC_SEED), case, rho_t, pairs,
                          cpu_adj=False, fd=True)
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=1)
    say("written")


def part(prev, U0, case, rho_t, pairs, cpu_adj, fd):
    run = oracle.run_loop(U0, case, SOLVER, LIMITER, STEPS, t_end=1.0)
    assert run["nsteps"] == STEPS
    resid = run["U"][2:-2, 2:-2, 0] - rho_t
    loss = 0.5 * float(np.sum(resid * resid))
    say(f"forward: L = {loss!r}")
    res = dict(loss=loss, dts_sha256=sha(run["dts"]), chs_sha256=sha(run["chs"]))
    if prev:
        assert prev["loss"] == loss and prev["dts_sha256"] == res["dts_sha256"]
    v = np.random.default_rng(SEED).standard_normal(U0.shape)

    def dloss(dU0):
        _, dU = oracle.tangent_run(U0, dU0, case, run["dts"], run["chs"], SOLVER, LIMITER)
        return float(np.sum(resid * dU[2:-2, 2:-2, 0]))

    if "tangent" in prev:
        res["tangent"] = prev["tangent"]
    else:
        grads = []
        for k, (j, i, f) in enumerate(pairs):
            e = np.zeros_like(U0)
            e[j, i, f] = 1.0
            grads.append(dloss(e))
            say(f"tangent {k}: ({j}, {i}, {f}) {grads[-1]!r}")
        res["tangent"] = dict(probes=grads, direction_dot=dloss(v))
    if cpu_adj:
        res["cpu_adjoint"] = prev.get("cpu_adjoint") or cpu_adjoint(U0, case, run, resid, pairs, v)
    if fd:
        res["frozen_fd"] = dict(prev.get("frozen_fd", {}))
        for eps in (1e-6, 1e-7, 2e-6):
            key = f"{eps:.0e}"
            if key in res["frozen_fd"]:
                continue
            rho0, rho_pm = oracle.frozen_fd(U0, pairs, eps, case, run["dts"], run["chs"], SOLVER, LIMITER)
            assert np.array_equal(rho0, run["U"][..., 0])
            vals = []
            for q in range(len(pairs)):
                rp = rho_pm[q, 0, 2:-2, 2:-2] - rho_t
                rm = rho_pm[q, 1, 2:-2, 2:-2] - rho_t
                vals.append(float(np.sum((rp - rm) * (rp + rm)) * 0.5 / (2 * eps)))
            res["frozen_fd"][key] = vals
            say(f"frozen FD eps={key}")
    return res


def cpu_adjoint(U0, case, run, resid, pairs, v):
    """The host build of the adjoint chained over the stored trajectory."""
    states, tmps, U = [U0], [], U0
    for s in range(STEPS):
        Utmp, U = oracle.advance(U, case, run["dts"][s], run["chs"][s], SOLVER, LIMITER)
        states.append(U)
        tmps.append(Utmp)
    assert np.array_equal(U, run["U"])
    lam = np.zeros_like(U0)
    lam[2:-2, 2:-2, 0] = resid
    for s in range(STEPS - 1, -1, -1):
        lam = hostcheck.adjoint_step(states[s], tmps[s], case, run["dts"][s], run["chs"][s], SOLVER, LIMITER, lam)
    say("CPU adjoint")
    return dict(sha256=sha(lam), layout="AoS float64 (ny+4, nx+4, 9), C order",
                probes=[float(lam[j, i, f]) for j, i, f in pairs], direction_dot=float(np.sum(lam * v)))


if __name__ == "__main__":
    main()
