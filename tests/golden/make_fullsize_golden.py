"""Writes fullsize.json: dt sequences and final states of full-size runs of the
measured configurations, for the GPU parity tests (tests/test_fullsize_gpu.py).

Each run starts from the case's initial state (the shared host IC code) and steps with
the driver loop of scripts/cpu/main_cpu.cpp:196-222 without output.  Where the
reference can run the configuration (minmod slopes, its only limiter), the reference
CPU solver compiled from its own sources (oracle/_ref/libmhdref.so) steps it; van Leer
runs are stepped by the oracle, which is pinned bitwise to the reference elsewhere
(tests/test_oracle.py).  Recorded per run: the step count, the final time, SHA-256 of
the dt and c_h sequences (little-endian float64), the first and last dt as hex, SHA-256
of the final state (host AoS, ghosts included, float64) and its per-field L1 norms.

    make -C oracle && python tests/golden/make_fullsize_golden.py      (about 80 minutes on 8 cores; the 7800^2 run needs about 40 GB of host memory)
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
OUT = os.path.join(ROOT, "tests", "golden", "fullsize.json")

import oracle  # noqa: E402

OT_UNIT, SHOCK_CLOUD, BRIO_WU_T = 4, 5, 6
SEED = 20261203
SOLVERS = {"hll": 0, "hllc": 1, "hlld": 2, "force": 3}

# name: (case id, nx, ny, solver, limiter, max steps, t_end or None for the case's)
RUNS = {
    "brio_wu_800x4_hlld_minmod_cfl0.8": (BRIO_WU_T, 800, 4, "hlld", 0, 100000, None),
    "ot4096_hlld_vanleer_100": (OT_UNIT, 4096, 4096, "hlld", 1, 100, None),
    "ot4096_hlld_minmod_100": (OT_UNIT, 4096, 4096, "hlld", 0, 100, None),
    "shock_cloud_2048x1024_hlld_vanleer_200": (SHOCK_CLOUD, 2048, 1024, "hlld", 1, 200, None),
    "shock_cloud_2048x1024_hlld_minmod_200": (SHOCK_CLOUD, 2048, 1024, "hlld", 0, 200, None),
}
for s in SOLVERS:
    for lim, lname in ((1, "vanleer"), (0, "minmod")):
        RUNS[f"ot1024_{s}_{lname}_t0.5"] = (OT_UNIT, 1024, 1024, s, lim, 100000, 0.5)
# The headline size with the other three Riemann solvers, both limiters.
for s in ("hll", "hllc", "force"):
    for lim, lname in ((1, "vanleer"), (0, "minmod")):
        RUNS[f"ot4096_{s}_{lname}_100"] = (OT_UNIT, 4096, 4096, s, lim, 100, None)
# A grid whose nine planes exceed 4 GiB (7804^2 cells with ghosts, 4.38 GB): the sweeps
# address it with 64-bit flat addresses instead of a buffer resource.
RUNS["ot7800_hll_minmod_3"] = (OT_UNIT, 7800, 7800, "hll", 0, 3, None)
# At CFL 0.8 the HLLC and HLLD runs stop at the blow-up guard before t = 0.5 (so does the
# reference's own solver); at the reference's own Orszag-Tang CFL of 0.4 HLLD finishes.
for lim, lname in ((1, "vanleer"), (0, "minmod")):
    RUNS[f"ot1024_hlld_{lname}_cfl0.4_t0.5"] = (OT_UNIT, 1024, 1024, "hlld", lim, 100000, 0.5, 0.4)


def case_of(case_id, nx, ny, cfl=None):
    c = oracle.case_config(case_id, 1)
    c.nx, c.ny = nx, ny
    if case_id == BRIO_WU_T:
        c.cfl = 0.8
    if cfl is not None:
        c.cfl = cfl
    return c


def initial(case_id, c):
    p = None
    if case_id == SHOCK_CLOUD:
        p = oracle.ICParams()
        p.seed = SEED
    return oracle.init_state(case_id, c, p)


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def main(names):
    data = {}
    if os.path.exists(OUT):
        data = json.load(open(OUT))
    for name in names:
        case_id, nx, ny, sname, lim, steps, t_end = RUNS[name][:7]
        c = case_of(case_id, nx, ny, RUNS[name][7] if len(RUNS[name]) > 7 else None)
        if t_end is not None:
            c.t_end = t_end
        U0 = initial(case_id, c)
        t0 = time.time()
        if lim == 0 and oracle.ref_available():
            r, _ = oracle.ref_run_loop_timed(U0, c, SOLVERS[sname], steps)
            src = "reference"
        else:
            r = oracle.run_loop(U0, c, SOLVERS[sname], lim, steps)
            src = "oracle"
        U = r["U"]
        data[name] = {
            "case": case_id, "nx": nx, "ny": ny, "solver": SOLVERS[sname], "limiter": lim, "cfl": c.cfl,
            "t_end": c.t_end, "max_steps": steps, "seed": SEED if case_id == SHOCK_CLOUD else None,
            "source": src, "code": int(r["code"]), "nsteps": int(r["nsteps"]), "t": float(r["t"]).hex(),
            "dts_sha256": sha(r["dts"]), "chs_sha256": sha(r["chs"]),
            "dt_first": [float(x).hex() for x in r["dts"][:3]], "dt_last": [float(x).hex() for x in r["dts"][-3:]],
            "state_sha256": sha(U),
            "field_l1": [float(x) for x in np.abs(U[2:-2, 2:-2]).reshape(-1, 9).sum(axis=0)],
        }
        print(f"{name}: {src}, {r['nsteps']} steps, code {r['code']}, {time.time() - t0:.0f} s", flush=True)
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:] or list(RUNS))
