"""Discrete adjoint of a fixed number of steps: ``Tape`` over the mhd_tape_* C ABI and
``MHDSteps``, its ``torch.autograd.Function``.

``MHDSteps.apply(U0, run, nsteps, K)`` maps a device state U0 -- a float64 CUDA
tensor of shape (9, ny+4, nx+4), the run's SoA layout with ghost cells -- to the
state after ``nsteps`` steps (compute_dt + advance_second_order each, the c_h
ratchet included), keeping a device checkpoint every K steps.  Its backward runs
the hand-written adjoint kernels: each K-step segment is recomputed from its
checkpoint, then reversed.  dt and c_h are frozen at the values the forward
recorded, and every branch (Riemann regions and fallbacks, limiter selections,
physical-state guards) at the one the forward took; the gradient is with respect
to every value of U0, ghost cells included.  Every Riemann solver (HLL, HLLC,
HLLD, FORCE) and limiter, and every boundary type (periodic, transmissive,
reflecting, Dirichlet).

Nothing here computes on the host: both directions are kernels of libmhd_hip.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib
from .solver import Run

__all__ = ["Tape", "MHDSteps", "run_steps", "replay", "to_planes", "from_planes", "bump_params"]

_dp = C.POINTER(C.c_double)


def _bind():
    L = lib()
    if getattr(L, "_tape_bound", False):
        return L
    L.mhd_tape_create.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
    L.mhd_tape_destroy.argtypes = [C.c_void_p]
    L.mhd_tape_forward.argtypes = [C.c_void_p, _dp, _dp]
    L.mhd_tape_backward.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    L.mhd_adjoint_step_grids.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_void_p,
                                         C.c_void_p]
    L._tape_bound = True
    return L


class Tape:
    """nsteps steps of ``run`` with a checkpoint every K steps (mhd_tape_create)."""

    def __init__(self, run: Run, nsteps: int, K: int = 20):
        import weakref
        L = _bind()
        h = C.c_void_p()
        check(L.mhd_tape_create(run.handle, int(nsteps), int(K), C.byref(h)))
        self._h = h
        self.run = run
        self.owner = None  # token of the autograd graph whose forward this tape recorded
        if not hasattr(run, "_tapes"):
            run._tapes = weakref.WeakSet()
        run._tapes.add(self)  # closed with the run
        self.nsteps = int(nsteps)
        self.K = int(K)
        self.dts = None
        self.chs = None

    def close(self):
        if getattr(self, "_h", None):
            lib().mhd_tape_destroy(self._h)
            self._h = None
            self.run._tapes.discard(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def forward(self):
        """Advance the run nsteps steps from its current state; returns (dts, chs)."""
        dts, chs = np.zeros(self.nsteps), np.zeros(self.nsteps)
        check(lib().mhd_tape_forward(self._h, dts.ctypes.data_as(_dp), chs.ctypes.data_as(_dp)))
        self.dts, self.chs = dts, chs
        return dts, chs

    def backward(self, lam_final: np.ndarray) -> np.ndarray:
        """dL/dU0 (AoS host array, the state shape) from dL/dU_N (same)."""
        lam_final = np.ascontiguousarray(lam_final, dtype=np.float64)
        if lam_final.shape != self.run.cfg.shape:
            raise ValueError(f"adjoint shape {lam_final.shape} != {self.run.cfg.shape}")
        out = np.empty_like(lam_final)
        check(lib().mhd_tape_backward(self._h, lam_final.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p),
                                      0))
        return out

    def backward_device(self, lam_final_ptr: int, lam0_ptr: int):
        """As backward, on dense device SoA buffers (9 x plane_stride doubles)."""
        check(lib().mhd_tape_backward(self._h, C.c_void_p(lam_final_ptr), C.c_void_p(lam0_ptr), 2))


def adjoint_step_grids(run: Run, Uold_ptr: int, Utmp_ptr: int, dt: float, ch: float, lam_ptr: int,
                       lam_tmp_ptr: int):
    """One reverse step over caller device grids (mhd_adjoint_step_grids)."""
    check(_bind().mhd_adjoint_step_grids(run.handle, C.c_void_p(Uold_ptr), C.c_void_p(Utmp_ptr), float(dt),
                                         float(ch), C.c_void_p(lam_ptr), C.c_void_p(lam_tmp_ptr)))


def to_planes(U: np.ndarray, device="cuda"):
    """AoS host state (ny+4, nx+4, 9) -> SoA float64 tensor (9, ny+4, nx+4) on ``device``."""
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.moveaxis(U, 2, 0))).to(device)


def from_planes(T) -> np.ndarray:
    return np.ascontiguousarray(np.moveaxis(T.detach().cpu().numpy(), 0, 2))


def _check_tensor(U, run: Run):
    import torch
    shape = (9,) + run.cfg.shape[:2]
    if not isinstance(U, torch.Tensor) or U.dtype != torch.float64 or not U.is_cuda:
        raise TypeError("expected a float64 CUDA tensor")
    if tuple(U.shape) != shape:
        raise ValueError(f"state shape {tuple(U.shape)} != {shape}")


class _on_current_stream:
    """Queue the run's work on torch's current stream for the duration (stream order,
    no host synchronisation of torch's streams); the run's own stream afterwards."""

    def __init__(self, run: Run):
        self.run = run

    def __enter__(self):
        import torch
        self.prev = self.run.stream
        self.run.set_stream(torch.cuda.current_stream().cuda_stream)

    def __exit__(self, *exc):
        self.run.set_stream(self.prev)


def _function():
    import torch

    class MHDSteps(torch.autograd.Function):
        """U_N = step^nsteps(U0) with the hand-written discrete adjoint as backward."""

        @staticmethod
        def forward(ctx, U0, run: Run, nsteps: int, K: int = 20):
            _check_tensor(U0, run)
            U0 = U0.contiguous()
            # an idle tape of the same shape (one whose backward has run) is reused
            pool = run.__dict__.setdefault("_idle_tapes", {})
            tape = pool.pop((nsteps, K), None)
            if tape is None or tape._h is None:
                tape = Tape(run, nsteps, K)
            tape.owner = None  # a recording interrupted below belongs to no graph
            out = torch.empty_like(U0)
            with _on_current_stream(run):
                run.set_state_device(U0.data_ptr())
                tape.forward()
                run.get_state_device(out.data_ptr())
            ctx.tape = tape
            ctx.token = tape.owner = object()  # this graph's claim on the tape's recording
            return out

        @staticmethod
        def backward(ctx, grad):
            tape = ctx.tape
            if tape._h is None or tape.owner is not ctx.token:
                raise RuntimeError("MHDSteps: this graph's tape was closed or reused by a later forward "
                                   "after its first backward; run the forward again to differentiate again")
            grad = grad.contiguous()
            lam0 = torch.empty_like(grad)
            with _on_current_stream(tape.run):
                tape.backward_device(grad.data_ptr(), lam0.data_ptr())
            # The recording stays valid (a second backward, with retain_graph, recomputes
            # from the same checkpoints) until a later forward takes the tape from the
            # pool.  The pool keeps one idle tape per run: a tape it displaces is closed
            # once no graph refers to it any more (Tape.__del__).
            tape.run.__dict__["_idle_tapes"] = {(tape.nsteps, tape.K): tape}
            return lam0, None, None, None

    return MHDSteps


_MHDSteps = None


def __getattr__(name):
    global _MHDSteps
    if name == "MHDSteps":
        if _MHDSteps is None:
            _MHDSteps = _function()
        return _MHDSteps
    raise AttributeError(name)


def run_steps(U0, run: Run, nsteps: int, K: int = 20):
    """Differentiable nsteps-step map of ``run`` (see MHDSteps)."""
    return __getattr__("MHDSteps").apply(U0, run, nsteps, K)


def replay(run: Run, U0, dts, chs):
    """The same steps with a given dt and c_h sequence (no ratchet), from the device
    state U0 (tensor); returns the final state as a new tensor.  With the sequence
    a tape recorded this reproduces its forward bit for bit (finite differences
    with frozen dt / c_h use it)."""
    import torch
    _check_tensor(U0, run)
    U0 = U0.contiguous()
    out = torch.empty_like(U0)
    with _on_current_stream(run):
        run.set_state_device(U0.data_ptr())
        for dt, ch in zip(dts, chs):
            run.ch = float(ch)
            run.advance(float(dt))
        run.get_state_device(out.data_ptr())
    return out


MASK = (1 << 64) - 1


def _splitmix(state: int):
    state = (state + 0x9E3779B97F4A7C15) & MASK
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return state, z ^ (z >> 31)


def bump_params(seed: int, amp: float = 0.02, sigma: float = 0.05):
    """(amp, sigma, cx, cy) of the seeded Gaussian density bump on [0,1]^2: the centre
    is uniform in [0.2, 0.8]^2 from two SplitMix64 draws (u = (z >> 11) 2^-53)."""
    st, z1 = _splitmix(seed)
    st, z2 = _splitmix(st)
    u1, u2 = (z1 >> 11) * 2.0 ** -53, (z2 >> 11) * 2.0 ** -53
    return amp, sigma, 0.2 + 0.6 * u1, 0.2 + 0.6 * u2


def probe_pairs(seed: int, nx: int, ny: int, n: int = 32):
    """n seeded (row, column, field) probes over the interior cells (ghosts offset 2)."""
    st, out = seed, []
    for _ in range(n):
        st, a = _splitmix(st)
        st, b = _splitmix(st)
        st, c = _splitmix(st)
        out.append((2 + a % ny, 2 + b % nx, c % 9))
    return out


def adjoint_workload(n: int = 512, seed: int = 20261203):
    """(cfg, U0, U_bump) of the adjoint workload: the unit-box Orszag-Tang case of the
    case table (periodic, CFL 0.8, gamma 5/3) at n^2, its initial state, and the same
    state plus the seeded Gaussian density bump (bump_params(seed))."""
    from dataclasses import replace

    from .solver import CaseId, ICParams, get_case_config, initial_state
    cfg = replace(get_case_config(CaseId.OrszagTangUnit), nx=n, ny=n)
    p = ICParams()
    p.bump_amp, p.bump_sigma, p.bump_cx, p.bump_cy = bump_params(seed)
    return cfg, initial_state(CaseId.OrszagTangUnit, cfg), initial_state(CaseId.OrszagTangBump, cfg, p)


def generic_state(U0, seed: int, amp: float = 1e-3):
    """U0 with seeded relative noise of size amp on every conserved field (rho scaled
    by 1 + amp u, the others offset by amp n times their field's RMS; u uniform in
    [-1, 1), n standard normal from numpy's default_rng(seed)).  The unperturbed
    Orszag-Tang state sits on exact ties of the scheme (slopes exactly zero along the
    axis a field is constant on, B_n = 0 lines, equal wave speeds), where the loss is
    not differentiable at any scale a finite difference can resolve; this state is
    generic, so central differences are well posed there."""
    rng = np.random.default_rng(seed)
    U = np.array(U0, dtype=np.float64, copy=True)
    U[..., 0] *= 1.0 + amp * rng.uniform(-1.0, 1.0, U.shape[:2])
    for k in range(1, 9):
        rms = float(np.sqrt(np.mean(U0[..., k] ** 2))) or 1.0
        U[..., k] += amp * rms * rng.standard_normal(U.shape[:2])
    return U


def main(argv=None):
    """One tape forward and backward of the adjoint workload (HLLD, van Leer) on the
    GPU, with wall times; a small driver for profilers:

        python -m mhd_gpu_acc_amd.adjoint --n 512 --steps 4 --K 2
    """
    import argparse
    import json
    import time

    import torch

    from .solver import Limiter, RiemannSolver

    ap = argparse.ArgumentParser(description=main.__doc__.splitlines()[0])
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--K", type=int, default=20)
    a = ap.parse_args(argv)
    cfg, U0, _ = adjoint_workload(a.n)
    lam = to_planes(np.random.default_rng(1).standard_normal(U0.shape))
    out = torch.empty_like(lam)
    with Run(cfg, RiemannSolver.HLLD, Limiter.VAN_LEER) as run:
        run.set_state(U0)
        run.ch = 0.0
        with Tape(run, a.steps, a.K) as tape:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            tape.forward()
            t1 = time.perf_counter()
            tape.backward_device(lam.data_ptr(), out.data_ptr())
            torch.cuda.synchronize()
            t2 = time.perf_counter()
    print(json.dumps(dict(n=a.n, steps=a.steps, K=a.K, forward_ms=1e3 * (t1 - t0), backward_ms=1e3 * (t2 - t1),
                          grad_abs_sum=float(out.abs().sum()))), flush=True)


if __name__ == "__main__":
    main()
