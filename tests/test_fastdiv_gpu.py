"""FastDiv (mhd_gpu_acc_amd/csrc/physics.hpp), the shared-reciprocal quotients of the
Riemann solvers, against the device's division bit for bit: the division sequence
without its two div_scale steps, with div_fixup kept, must give the bits of a / b
whenever every operand's frexp exponent lies in [-380, 380] or the operand is zero,
infinite or NaN, and FastDiv's range record must say so exactly then.  Two numerators
share each denominator's reciprocal, as in the solvers."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SO = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mhd_gpu_acc_amd",
                  "libmhd_devcheck.so")
DIV_EXP = 380


def _lib():
    import torch  # noqa: F401  (bind to torch's HIP runtime first, as the product library does)
    if not os.path.exists(SO):
        raise FileNotFoundError(f"{SO} is not built (make -C mhd_gpu_acc_amd/csrc)")
    L = C.CDLL(SO)
    L.dc_div_pair.argtypes = [C.c_void_p] * 6 + [C.c_longlong]
    return L


def _run(a1, a2, b):
    import torch
    n = b.size
    d = [torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in (a1, a2, b)]
    q = torch.empty(2 * n, dtype=torch.float64, device="cuda")
    e = torch.empty_like(q)
    ok = torch.empty(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    assert _lib().dc_div_pair(d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), q.data_ptr(), e.data_ptr(),
                              ok.data_ptr(), n) == 0
    return q.cpu().numpy().reshape(n, 2), e.cpu().numpy().reshape(n, 2), ok.cpu().numpy().astype(bool)


def _same(a, b):
    return (a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))


def _in_range(x):
    e = np.frexp(x)[1]  # C frexp: 0 for zero, infinite and NaN
    return (e >= -DIV_EXP) & (e <= DIV_EXP)


def _operands(rng, n, lo, hi):
    """Signed values with frexp exponents drawn from [lo, hi) (mantissas in [0.5, 1))."""
    x = np.ldexp(0.5 + 0.5 * rng.random(n), rng.integers(lo, hi, n))
    return np.where(rng.random(n) < 0.5, x, -x)


SPECIAL = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, -5e-324, 2.0 ** -1022, 2.0 ** -1023,
                    np.finfo(np.float64).max, 2.0 ** -381, 2.0 ** -382, 2.0 ** 379, 2.0 ** 380, 1.0, -3.0])


def test_fastdiv_in_range_bitwise():
    """Operands over the whole range FastDiv accepts, ends included: every quotient, both
    numerators of a shared reciprocal, has the bits of the division (and the host's)."""
    rng = np.random.default_rng(11)
    n = 1 << 21
    b = _operands(rng, n, -DIV_EXP, DIV_EXP + 1)
    a1 = _operands(rng, n, -DIV_EXP, DIV_EXP + 1)
    a2 = _operands(rng, n, -40, 41)
    ends = np.array([2.0 ** -381, np.nextafter(2.0 ** 380, 0.0), -2.0 ** -381, 1.0, np.nextafter(1.0, 2.0), 0.75])
    k = np.arange(ends.size)
    b[: ends.size], a1[: ends.size], a2[: ends.size] = ends, ends[::-1], ends[(k + 2) % ends.size]
    q, e, ok = _run(a1, a2, b)
    assert ok.all()
    same = _same(q, e)
    assert same.all(), (np.argwhere(~same)[:5], q[~same][:5], e[~same][:5])
    with np.errstate(all="ignore"):
        assert _same(e, np.stack([a1 / b, a2 / b], axis=1)).all()  # the device's division is IEEE


def test_fastdiv_special_and_out_of_range():
    """Zero, infinite and NaN operands are in range and give the division's bits; the
    range record is exactly the frexp test on the three operands, and every pair it
    accepts gives the division's bits (the ones it refuses are divided again by the
    solvers)."""
    rng = np.random.default_rng(12)
    n = 1 << 20
    def draw():  # three quarters near the range's ends, a quarter over every exponent
        return np.where(rng.random(n) < 0.75, _operands(rng, n, -400, 401), _operands(rng, n, -1074, 1025))

    b, a1, a2 = draw(), draw(), draw()
    m = SPECIAL.size
    i, j = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    b[: m * m], a1[: m * m], a2[: m * m] = SPECIAL[j.ravel()], SPECIAL[i.ravel()], SPECIAL[(i.ravel() + 3) % m]
    for s in range(3):  # specials at random places
        idx = rng.integers(m * m, n, n // 16)
        (b, a1, a2)[s][idx] = SPECIAL[rng.integers(0, m, idx.size)]
    q, e, ok = _run(a1, a2, b)
    assert (ok == (_in_range(a1) & _in_range(a2) & _in_range(b))).all()
    assert ok.sum() > n // 5 and (~ok).sum() > n // 5
    same = _same(q, e)
    bad = ok[:, None] & ~same
    assert not bad.any(), (np.argwhere(bad)[:5], q[bad][:5], e[bad][:5])
    with np.errstate(all="ignore"):
        assert _same(e, np.stack([a1 / b, a2 / b], axis=1)).all()
