"""Compare the outputs of two run directories: the reference's scripts/compare_runs.py
(same arguments, table, PASS rule and exit codes), plus a rel-L1 column with an
optional bound and binary state dumps (*.bin, compared field by field).

    python -m mhd_gpu_acc_amd.compare dir_a dir_b [--tol 1e-10] [--rtol 1e-6]
                                                  [--rel-l1 1e-12] [--verbose] [--fields rho,p]

Per matched file (compare_runs.py:53-67): max|a-b|, mean|a-b|, max |a-b|/max(|a|,|b|)
(0 where both are 0), L2 = sqrt(mean |a-b|^2); here also rel-L1 = sum|a-b| / sum|b|
(b = dir_b, the reference side).  A file passes if max|a-b| <= tol or
max rel <= rtol (compare_runs.py:176), and, with --rel-l1, rel-L1 <= that bound.
Exit 0 if every pair passes, 1 otherwise, 2 on a usage / IO error.
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path

import numpy as np

BIN_FIELDS = ("rho", "mx", "my", "mz", "Bx", "By", "Bz", "E", "psi")


def load_csv(path: Path) -> np.ndarray:
    """A field CSV (rows = j from the top, cols = i) as a float64 array."""
    try:
        import pandas as pd
        return pd.read_csv(path, header=None, dtype=np.float64, engine="c").to_numpy()
    except ImportError:  # pragma: no cover - pandas is optional
        return np.loadtxt(path, delimiter=",", dtype=np.float64, ndmin=2)


def load_bin_fields(path: Path) -> dict[str, np.ndarray]:
    """The interior cells of each field of a state dump."""
    from .driver import load_bin
    head, U = load_bin(str(path))
    g = head["ng"]
    inner = U[g:-g, g:-g, :]
    return {name: inner[..., k] for k, name in enumerate(BIN_FIELDS)}


def find_files(root: Path) -> dict[str, Path]:
    return {str(p.relative_to(root)): p for suffix in ("*.csv", "*.bin") for p in sorted(root.rglob(suffix))}


def field_name_from_path(rel: str) -> str:
    """'brio_wu_cpu_t10_rho.csv' -> 'rho' (compare_runs.py:46-49); 'x.bin:Bx' -> 'Bx'."""
    if ":" in rel:
        return rel.rsplit(":", 1)[-1]
    return Path(rel).stem.rsplit("_", 1)[-1]


def diff_stats(a: np.ndarray, b: np.ndarray):
    if a.shape != b.shape:
        return None
    d = np.abs(a - b)
    ref = np.maximum(np.abs(a), np.abs(b))
    with np.errstate(invalid="ignore", divide="ignore"):
        rel = np.where(ref > 0, d / ref, 0.0)
    sb = float(np.abs(b).sum())
    sd = float(d.sum())
    return {"max_abs": float(d.max()) if d.size else 0.0, "mean_abs": float(d.mean()) if d.size else 0.0,
            "max_rel": float(rel.max()) if rel.size else 0.0,
            "l2_abs": float(np.sqrt((d ** 2).mean())) if d.size else 0.0,
            "rel_l1": (sd / sb) if sb > 0 else (0.0 if sd == 0 else float("inf")),
            "shape_a": a.shape, "shape_b": b.shape}


def passes(stats, tol: float, rtol: float, rel_l1: float | None) -> bool:
    ok = (stats["max_abs"] <= tol) or (stats["max_rel"] <= rtol)
    if rel_l1 is not None:
        ok = ok and stats["rel_l1"] <= rel_l1
    return bool(ok)


def _pairs(files_a, files_b, common):
    """(label, loader_a, loader_b) for every compared array."""
    for rel in common:
        if rel.endswith(".bin"):
            fa, fb = load_bin_fields(files_a[rel]), load_bin_fields(files_b[rel])
            for name in BIN_FIELDS:
                yield f"{rel}:{name}", fa[name], fb[name]
        else:
            yield rel, files_a[rel], files_b[rel]


def compare_dirs(dir_a, dir_b, tol=1e-10, rtol=1e-6, rel_l1=None, fields="", verbose=False, out=sys.stdout):
    """Runs the comparison, printing the report to ``out``; returns the exit code."""
    dir_a, dir_b = Path(dir_a), Path(dir_b)
    p = lambda *a: print(*a, file=out)  # noqa: E731
    for d in (dir_a, dir_b):
        if not d.is_dir():
            print(f"[ERROR] directory not found: {d}", file=sys.stderr)
            return 2
    field_filter = {f.strip() for f in fields.split(",") if f.strip()}
    files_a, files_b = find_files(dir_a), find_files(dir_b)
    only_a = sorted(set(files_a) - set(files_b))
    only_b = sorted(set(files_b) - set(files_a))
    common = sorted(set(files_a) & set(files_b))
    p(f"\n{'=' * 70}")
    p("  MHD run comparison")
    p(f"  A : {dir_a}")
    p(f"  B : {dir_b}")
    p(f"{'=' * 70}")
    p(f"  Files in A only : {len(only_a)}")
    p(f"  Files in B only : {len(only_b)}")
    p(f"  Matched pairs   : {len(common)}")
    for title, lst in (("[ONLY IN A]", only_a), ("[ONLY IN B]", only_b)):
        if lst:
            p(f"\n  {title}")
            for f in lst:
                p(f"    {f}")
    if not common:
        p("\n  No matching files to compare.")
        return 0
    w = [42, 10, 10, 10, 10, 10, 8]
    p(f"\n  {'File':<{w[0]}}  {'max|Δ|':>{w[1]}}  {'mean|Δ|':>{w[2]}}  {'max|Δ|/ref':>{w[3]}}  {'L2(Δ)':>{w[4]}}  "
      f"{'relL1':>{w[5]}}  {'result':>{w[6]}}")
    sep = "  " + "-" * (sum(w) + 2 * (len(w) - 1))
    p(sep)
    n_pass = n_fail = n_shape = 0
    for label, a, b in _pairs(files_a, files_b, common):
        if field_filter and field_name_from_path(label) not in field_filter:
            continue
        try:
            if isinstance(a, Path):
                a, b = load_csv(a), load_csv(b)
        except Exception as e:  # noqa: BLE001 - report and count, as the reference does
            p(f"  [ERROR] could not load '{label}': {e}")
            n_fail += 1
            continue
        st = diff_stats(a, b)
        if st is None:
            n_shape += 1
            p(f"  {label:<{w[0]}}  shape A={a.shape} vs B={b.shape}  {'SHAPE!':>{w[6]}}")
            continue
        ok = passes(st, tol, rtol, rel_l1)
        n_pass += ok
        n_fail += not ok
        if verbose or not ok:
            p(f"  {Path(label).name:<{w[0]}}  {st['max_abs']:>{w[1]}.3e}  {st['mean_abs']:>{w[2]}.3e}  "
              f"{st['max_rel']:>{w[3]}.3e}  {st['l2_abs']:>{w[4]}.3e}  {st['rel_l1']:>{w[5]}.3e}  "
              f"{'PASS' if ok else 'FAIL':>{w[6]}}")
    p(sep)
    p(f"\n  Compared : {n_pass + n_fail + n_shape} file pairs")
    p(f"  PASS     : {n_pass}")
    p(f"  FAIL     : {n_fail}")
    if n_shape:
        p(f"  SHAPE    : {n_shape}  (grids differ)")
    p(f"\n  Absolute tolerance : {tol:.2e}")
    p(f"  Relative tolerance : {rtol:.2e}")
    if rel_l1 is not None:
        p(f"  rel-L1 bound       : {rel_l1:.2e}")
    bad = n_fail > 0 or n_shape > 0
    p(f"\n  Overall result     : {'FAIL' if bad else 'PASS'}\n")
    return 1 if bad else 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Compare MHD field outputs from two run directories.")
    ap.add_argument("dir_a")
    ap.add_argument("dir_b")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--rtol", type=float, default=1e-6)
    ap.add_argument("--rel-l1", type=float, default=None, help="also require sum|a-b|/sum|b| <= this")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--fields", default="")
    a = ap.parse_args(argv)
    return compare_dirs(a.dir_a, a.dir_b, a.tol, a.rtol, a.rel_l1, a.fields, a.verbose)


if __name__ == "__main__":
    sys.exit(main())
