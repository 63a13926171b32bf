"""Writes driver_reference.json: what the reference CPU driver produces for a few cases.

The driver is the reference's own scripts/cpu/main_cpu.cpp compiled from its sources by
oracle/Makefile (oracle/_ref/main_cpu).  For each case it records, from one run with
--solver hlld (minmod, the reference's only limiter):

* the SHA-256 and size of every snapshot CSV it writes (initial state, the case's
  snapshot schedule or the final state);
* its progress table lines (step, t, dt, max|B|) and [snap] lines, verbatim;
* the step count of its [TIMING] line.

    make -C oracle ref && python tests/golden/make_driver_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
BIN = os.path.join(ROOT, "oracle", "_ref", "main_cpu")
OUT = os.path.join(ROOT, "tests", "golden", "driver_reference.json")
CASES = ("brio_wu", "orszag_tang", "shock_bubble", "kelvin_helmholtz")
PROGRESS = re.compile(r"^\s+\d+\s+\d+\.\d{5}\s+\S+\s+\S+$")


def run(case, tmp):
    env = dict(os.environ, OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS", "8"))
    out_dir = os.path.join(tmp, case)
    p = subprocess.run([BIN, case, "--solver", "hlld", "--out", out_dir], capture_output=True, text=True, env=env,
                       check=True)
    lines = p.stdout.splitlines()
    files = {}
    for name in sorted(os.listdir(out_dir)):
        data = open(os.path.join(out_dir, name), "rb").read()
        files[name] = dict(sha256=hashlib.sha256(data).hexdigest(), bytes=len(data))
    steps = int(re.search(r"\[TIMING\].* steps=(\d+)", p.stdout).group(1))
    return dict(args=["--solver", "hlld"], files=files, steps=steps,
                progress=[ln for ln in lines if PROGRESS.match(ln)],
                snaps=[ln for ln in lines if ln.startswith("  [snap]")])


def main():
    if not os.path.exists(BIN):
        sys.exit(f"{BIN} not built: make -C oracle ref")
    res = {}
    with tempfile.TemporaryDirectory() as tmp:
        for case in CASES:
            res[case] = run(case, tmp)
            print(case, res[case]["steps"], "steps,", len(res[case]["files"]), "files", flush=True)
    with open(OUT, "w") as fh:
        json.dump(res, fh, indent=1)


if __name__ == "__main__":
    main()
