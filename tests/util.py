"""Shared helpers for the parity tests."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_point():
    return dict(np.load(os.path.join(GOLDEN, "ref_point.npz")))


def load_grid():
    return dict(np.load(os.path.join(GOLDEN, "ref_grid.npz")))


def load_meta():
    with open(os.path.join(GOLDEN, "ref_meta.json")) as f:
        return json.load(f)


def state_hash(U: np.ndarray) -> str:
    """SHA-256 of the float64 bytes with every NaN mapped to one pattern."""
    a = np.ascontiguousarray(U, dtype=np.float64).copy()
    a[np.isnan(a)] = np.nan
    return hashlib.sha256(a.tobytes()).hexdigest()


def bitwise_mismatch(a, b) -> int:
    """Number of entries whose float64 bit patterns differ (any NaN equals any NaN)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    assert a.shape == b.shape, (a.shape, b.shape)
    na, nb = np.isnan(a), np.isnan(b)
    same = (a.view(np.uint64) == b.view(np.uint64)) | (na & nb)
    return int((~same).sum())


def assert_bitwise(a, b, what=""):
    n = bitwise_mismatch(a, b)
    if n:
        a = np.asarray(a, dtype=np.float64)
        b = np.asarray(b, dtype=np.float64)
        idx = np.argwhere(~((a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))))[:5]
        raise AssertionError(f"{what}: {n} entries differ bitwise; first at {idx.tolist()}: "
                             f"{[a[tuple(i)] for i in idx]} vs {[b[tuple(i)] for i in idx]}")


def rel_l1(a, b, interior=None) -> float:
    """sum |a-b| / sum |b| over the interior (ng=2) of AoS states."""
    if interior is None:
        a = a[2:-2, 2:-2]
        b = b[2:-2, 2:-2]
    den = np.abs(b).sum()
    return float(np.abs(a - b).sum() / den) if den > 0 else float(np.abs(a - b).sum())


def parse_float(v):
    if isinstance(v, str):
        return float(v)
    return v
