"""sqrt_cr (mhd_gpu_acc_amd/csrc/physics.hpp), the square root of the sweeps, against
the device's sqrt() bit for bit: the core rsq + Newton-Raphson sequence that
sqrt_cr runs when every active lane's argument lies in [2^-767, DBL_MAX] must give
the bits of the full lowering there, and waves with any other argument (zero,
denormal, below 2^-767, infinite, NaN, negative) must fall back to sqrt()."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SO = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mhd_gpu_acc_amd",
                  "libmhd_devcheck.so")


def _lib():
    import torch  # noqa: F401  (bind to torch's HIP runtime first, as the product library does)
    if not os.path.exists(SO):
        raise FileNotFoundError(f"{SO} is not built (make -C mhd_gpu_acc_amd/csrc)")
    L = C.CDLL(SO)
    L.dc_sqrt_pair.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong]
    return L


def _run(x: np.ndarray):
    import torch
    xd = torch.from_numpy(x).cuda()
    a = torch.empty_like(xd)
    b = torch.empty_like(xd)
    torch.cuda.synchronize()
    assert _lib().dc_sqrt_pair(xd.data_ptr(), a.data_ptr(), b.data_ptr(), x.size) == 0
    return a.cpu().numpy(), b.cpu().numpy()


def _same(a, b):
    return (a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))


def test_sqrt_cr_core_range_bitwise():
    """Every wave holds core-range arguments only: random exponents over the whole
    range, mantissas random, plus the range ends and their neighbours."""
    rng = np.random.default_rng(5)
    n = 1 << 22
    e = rng.integers(-767, 1024, n)
    m = rng.random(n) + 1.0
    x = np.ldexp(m, e)
    x = x[(x >= 2.0 ** -767) & np.isfinite(x)]
    edge = np.array([2.0 ** -767, np.nextafter(2.0 ** -767, np.inf), np.finfo(np.float64).max,
                     np.nextafter(np.finfo(np.float64).max, 0.0), 1.0, np.nextafter(1.0, 0.0),
                     np.nextafter(1.0, 2.0), 2.0, 4.0, 0.25])
    x = np.concatenate([edge, x[: (x.size // 64) * 64 - edge.size]])
    x = np.ascontiguousarray(x)
    assert x.size % 64 == 0 and (x >= 2.0 ** -767).all()
    a, b = _run(x)
    same = _same(a, b)
    assert same.all(), (x[~same][:5], a[~same][:5], b[~same][:5])
    assert _same(a, np.sqrt(x)).all()  # and correctly rounded, as the host's sqrt


def test_sqrt_cr_mixed_waves_fall_back():
    """Waves mixing core arguments with special and out-of-range ones."""
    rng = np.random.default_rng(6)
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, -1.0, 5e-324, 2.0 ** -1022,
                        np.nextafter(2.0 ** -767, 0.0), 2.0 ** -800, 1e-300, -2.0 ** -767])
    n = 1 << 18
    x = np.ldexp(rng.random(n) + 1.0, rng.integers(-1074, 1024, n))
    # one special value in every other wave, at a varying lane
    for w in range(0, n // 64, 2):
        x[64 * w + (w * 7) % 64] = special[w % special.size]
    with np.errstate(invalid="ignore"):
        a, b = _run(np.ascontiguousarray(x))
    same = _same(a, b)
    assert same.all(), (x[~same][:5], a[~same][:5], b[~same][:5])
    with np.errstate(invalid="ignore"):
        host = np.sqrt(x)
    assert _same(b, host).all()  # the device's sqrt is the host's (IEEE) on every argument
