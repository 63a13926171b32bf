"""Writes the golden vectors of the reference CPU path.

Runs in the build container only (it needs ``oracle/_ref/libmhdref.so``, the
reference compiled from its own sources by ``oracle/Makefile``).  The output
files are plain data -- inputs and the reference's outputs -- committed next to
this script:

* ``ref_point.npz``  point known-answer tests of head/physics.hpp,
  head/riemann.hpp and the solver_cpu.cpp helpers (minmod, is_physical,
  MUSCL-Hancock reconstruct);
* ``ref_grid.npz``   boundary-condition KATs (head/cpu/boundary_cpu.hpp), the
  case table (src/test_cases.cpp), initial-state hashes (src/init.cpp) and
  multi-step trajectories (dt and c_h per step, SHA-256 of the final state) of
  the driver loop (scripts/cpu/main_cpu.cpp:196-222).

Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402
from oracle import ptr, ibuf  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
GAMMAS = (5.0 / 3.0, 1.4, 2.0)


def state_hash(U: np.ndarray) -> str:
    """SHA-256 of the float64 bytes with every NaN mapped to one pattern."""
    a = np.ascontiguousarray(U, dtype=np.float64).copy()
    a[np.isnan(a)] = np.nan
    return hashlib.sha256(a.tobytes()).hexdigest()


# ---------------------------------------------------------------------------------------
# Point states
# ---------------------------------------------------------------------------------------
def prim_to_cons_np(V, gamma):
    U = V.copy()
    rho, u, v, w, bx, by, bz, p = (V[:, k] for k in range(8))
    U[:, 1], U[:, 2], U[:, 3] = rho * u, rho * v, rho * w
    U[:, 7] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * (bx * bx + by * by + bz * bz)
    return U


def random_prims(rng, n):
    V = np.empty((n, 9))
    V[:, 0] = np.exp(rng.uniform(np.log(0.05), np.log(20.0), n))
    V[:, 1:4] = rng.uniform(-3, 3, (n, 3))
    V[:, 4:7] = rng.uniform(-3, 3, (n, 3))
    V[:, 7] = np.exp(rng.uniform(np.log(0.01), np.log(100.0), n))
    V[:, 8] = rng.uniform(-0.5, 0.5, n)
    return V


def special_prims():
    """Degenerate primitive states: vacuum-like, zero fields, unphysical, NaN."""
    rows = [
        [1.0, 0, 0, 0, 0, 0, 0, 1.0, 0],            # at rest, B = 0
        [1.0, 0, 0, 0, 0.75, 1.0, 0, 1.0, 0],       # Brio-Wu left
        [0.125, 0, 0, 0, 0.75, -1.0, 0, 0.1, 0],    # Brio-Wu right
        [1e-13, 0.1, 0, 0, 0.1, 0, 0, 1e-13, 0],    # below the 1e-12 floors
        [1e-6, 5.0, -5.0, 0, 1e-3, 1e-3, 0, 1e-8, 0],  # near vacuum, fast
        [1.0, 0, 0, 0, 2.0, 0, 0, 1.0, 0.3],        # B purely normal (x)
        [1.0, 0, 0, 0, 0, 2.0, 0, 1.0, -0.3],       # B purely normal (y)
        [1.0, 1.0, 1.0, 1.0, 0, 0, 3.0, 1.0, 0],    # B_z only
        [1e4, 0, 0, 0, 0, 0, 0, 1e4, 0],            # strong jump partner
        [1.0, 0, 0, 0, 0, 0, 0, -1.0, 0],           # negative pressure
        [-1.0, 0, 0, 0, 0, 0, 0, 1.0, 0],           # negative density
        [np.nan, 0, 0, 0, 0, 0, 0, 1.0, 0],         # NaN density
        [1.0, np.nan, 0, 0, 0, 0, 0, 1.0, 0],       # NaN velocity
        [1.0, 0, 0, 0, 0, 0, 0, np.inf, 0],         # infinite pressure
        [1.0, 0, 0, 0, 0, 0, 0, 1.0, np.nan],       # NaN psi
        [0.0, 0, 0, 0, 0, 0, 0, 0.0, 0],            # all zero
    ]
    return np.array(rows, dtype=np.float64)


def make_point(rng):
    out = {}
    n_rand = 320
    Vr = random_prims(rng, n_rand)
    Vs = special_prims()
    V = np.concatenate([Vr, Vs])
    n = len(V)
    gam = np.array([GAMMAS[k % 3] for k in range(n)])
    U = np.stack([prim_to_cons_np(V[k:k + 1], gam[k])[0] for k in range(n)])
    out["U"], out["V"], out["gamma"] = U, V, gam
    lib = oracle.ref()

    c2p = np.zeros_like(U)
    p2c = np.zeros_like(U)
    fl = np.zeros((2, 2, n, 9))
    fs = np.zeros((2, n))
    ms = np.zeros((2, 2, n))
    phys = np.zeros(n, dtype=np.int32)
    chs = (0.0, 1.7)
    for k in range(n):
        lib.ref_set_gamma(gam[k])
        lib.ref_cons_to_prim(ptr(U[k]), ptr(c2p[k]))
        lib.ref_prim_to_cons(ptr(V[k]), ptr(p2c[k]))
        phys[k] = lib.ref_is_physical(ptr(V[k]))
        for d in range(2):
            fs[d, k] = lib.ref_fast_speed(ptr(V[k]), d)
            for c, ch in enumerate(chs):
                f = np.zeros(9)
                lib.ref_flux(ptr(U[k]), d, ch, ptr(f))
                fl[d, c, k] = f
                ms[d, c, k] = lib.ref_max_signal_speed(ptr(V[k]), d, ch)
    out.update(c2p=c2p, p2c=p2c, flux=fl, fast=fs, maxsig=ms, is_physical=phys, chs=np.array(chs))

    # Riemann pairs: random-random, random-special, near-equal, GLM B_n jumps.
    m = 384
    il = rng.integers(0, n, m)
    ir = rng.integers(0, n, m)
    il[:n_rand // 2] = np.arange(n_rand // 2)
    ir[:n_rand // 2] = np.arange(n_rand // 2, n_rand)
    UL = U[il].copy()
    UR = U[ir].copy()
    pg = gam[il].copy()
    # Left and right states with the same gamma so the pair is meaningful.
    for k in range(m):
        UR[k] = prim_to_cons_np(V[ir[k]:ir[k] + 1], pg[k])[0]
        UL[k] = prim_to_cons_np(V[il[k]:il[k] + 1], pg[k])[0]
    # Identical pairs, tiny perturbations and pure B_n / psi jumps.
    for k in range(300, 330):
        UR[k] = UL[k]
    for k in range(330, 360):
        UR[k] = UL[k] * (1.0 + 1e-9 * rng.standard_normal(9))
    for k in range(360, m):
        UR[k] = UL[k].copy()
        UR[k, 4] += 0.3
        UR[k, 5] -= 0.2
        UR[k, 8] += 0.1
    out.update(UL=UL, UR=UR, pair_gamma=pg)
    raw = np.zeros((4, 2, 2, m, 9))
    rf = np.zeros((4, 2, 2, m, 9))
    glm = np.zeros((2, 2, m, 9))
    for k in range(m):
        lib.ref_set_gamma(pg[k])
        for d in range(2):
            for c, ch in enumerate(chs):
                for s in range(4):
                    f = np.zeros(9)
                    lib.ref_solver_flux(s, ptr(UL[k]), ptr(UR[k]), d, ch, ptr(f))
                    raw[s, d, c, k] = f
                    f2 = np.zeros(9)
                    lib.ref_riemann_flux(ptr(UL[k]), ptr(UR[k]), d, s, ch, ptr(f2))
                    rf[s, d, c, k] = f2
                g = raw[0, d, c, k].copy()
                lib.ref_apply_glm_flux(ptr(g), ptr(UL[k]), ptr(UR[k]), d, ch)
                glm[d, c, k] = g
    out.update(solver_raw=raw, riemann=rf, glm=glm)

    # minmod
    base = np.array([0.0, -0.0, 1.0, -1.0, 0.5, -0.5, 2.0, 1e-300, -1e-300, np.inf, -np.inf, np.nan])
    a = np.repeat(base, len(base))
    b = np.tile(base, len(base))
    ra = rng.uniform(-2, 2, 400)
    rb = rng.uniform(-2, 2, 400)
    ma = np.concatenate([a, ra])
    mb = np.concatenate([b, rb])
    mm = np.array([lib.ref_minmod(x, y) for x, y in zip(ma, mb)])
    out.update(mm_a=ma, mm_b=mb, minmod=mm)

    # reconstruct: triples of primitive states, two directions, two c_h values.
    t = 256
    Wm = random_prims(rng, t)
    Wc = random_prims(rng, t)
    Wp = random_prims(rng, t)
    # smooth triples (slopes of one sign), flat triples, and floor-crossing triples
    Wc[:64] = Wm[:64] * 1.01
    Wp[:64] = Wm[:64] * 1.02
    Wc[64:80] = Wm[64:80]
    Wp[64:80] = Wm[64:80]
    Wm[80:96, 7] = 1e-12
    Wc[80:96, 7] = 2e-12
    Wm[96:112, 0] = 1e-13
    rg = np.array([GAMMAS[k % 3] for k in range(t)])
    dtd = rng.uniform(0.01, 0.4, t)
    dtd[112:128] = 50.0  # large half steps trip the conserved-state guard
    rec = np.zeros((2, 2, t, 2, 9))
    for k in range(t):
        lib.ref_set_gamma(rg[k])
        for d in range(2):
            for c, ch in enumerate(chs):
                lib.ref_set_ch(ch)
                l = np.zeros(9)
                r = np.zeros(9)
                lib.ref_reconstruct(ptr(Wm[k]), ptr(Wc[k]), ptr(Wp[k]), dtd[k], d, ptr(l), ptr(r))
                rec[d, c, k, 0] = l
                rec[d, c, k, 1] = r
    lib.ref_set_ch(0.0)
    out.update(rec_Wm=Wm, rec_Wc=Wc, rec_Wp=Wp, rec_gamma=rg, rec_dtd=dtd, reconstruct=rec)
    return out


# ---------------------------------------------------------------------------------------
# Grids and trajectories
# ---------------------------------------------------------------------------------------
def ref_grid(case_id, nx, ny, x0, x1, y0, y1, gamma, ng=2):
    """Reference grid filled from the reference's initial_state_at at cell centres."""
    lib = oracle.ref()
    lib.ref_set_gamma(gamma)
    g = lib.ref_grid_new(nx, ny, ng, x0, x1, y0, y1)
    dx = lib.ref_grid_dx(g)
    dy = lib.ref_grid_dy(g)
    U = np.zeros((ny + 2 * ng, nx + 2 * ng, 9))
    for j in range(ny + 2 * ng):
        y = y0 + (float(j - ng) + 0.5) * dy
        for i in range(nx + 2 * ng):
            x = x0 + (float(i - ng) + 0.5) * dx
            lib.ref_initial_state_at(case_id, x, y, ptr(U[j, i]))
    lib.ref_grid_set(g, ptr(U))
    return g, U


def grid_get(g, shape):
    U = np.zeros(shape)
    oracle.ref().ref_grid_get(g, ptr(U))
    return U


# Trajectory specs: name, case id (IC), geometry, gamma, cfl, t_end, bc, solver, limiter,
# max steps, a list of (i, j, field, value) perturbations applied to the IC.
PI = 3.14159265358979323846


def traj_specs():
    specs = []
    bw_ref = dict(case=2, nx=800, ny=4, x=(0.0, 1.0), y=(0.0, 4.0 / 800.0), gamma=2.0)
    for s, name in ((2, "hlld"), (0, "hll"), (3, "force"), (1, "hllc")):
        specs.append(dict(name=f"brio_wu_ref_{name}", cfl=0.4, t_end=0.1, bc=[1, 1, 0, 0], solver=s,
                          steps=2000, **bw_ref))
        specs.append(dict(name=f"brio_wu_cfl08_{name}", cfl=0.8, t_end=0.1, bc=[1, 1, 1, 1], solver=s,
                          steps=2000, **bw_ref))
    ot = dict(case=3, x=(0.0, 2 * PI), y=(0.0, 2 * PI), gamma=5.0 / 3.0, bc=[0, 0, 0, 0])
    for s, name in ((2, "hlld"), (0, "hll"), (3, "force"), (1, "hllc")):
        specs.append(dict(name=f"ot_ref64_{name}", nx=64, ny=64, cfl=0.4, t_end=10.0, solver=s, steps=40, **ot))
    specs.append(dict(name="ot_ref128_hlld", nx=128, ny=128, cfl=0.4, t_end=10.0, solver=2, steps=20, **ot))
    specs.append(dict(name="ot_ref_rect_hlld", nx=40, ny=24, cfl=0.8, t_end=10.0, solver=2, steps=60, **ot))
    refl = dict(case=3, x=(0.0, 2 * PI), y=(0.0, 2 * PI), gamma=5.0 / 3.0, nx=48, ny=40)
    specs.append(dict(name="ot_reflect_hlld", cfl=0.4, t_end=10.0, bc=[2, 2, 2, 2], solver=2, steps=30, **refl))
    specs.append(dict(name="ot_reflect_hllc", cfl=0.4, t_end=10.0, bc=[2, 2, 2, 2], solver=1, steps=30, **refl))
    specs.append(dict(name="ot_mixed_force", cfl=0.4, t_end=10.0, bc=[2, 1, 0, 0], solver=3, steps=30, **refl))
    specs.append(dict(name="kh_dirichlet_hlld", case=0, nx=32, ny=48, x=(0.0, 1.0), y=(-1.0, 1.0), gamma=1.4,
                      cfl=0.3, t_end=10.0, bc=[3, 1, 2, 3], solver=2, steps=30))
    specs.append(dict(name="kh_periodic_hll", case=0, nx=32, ny=48, x=(0.0, 1.0), y=(-1.0, 1.0), gamma=1.4,
                      cfl=0.3, t_end=10.0, bc=[0, 0, 0, 0], solver=0, steps=30))
    specs.append(dict(name="shock_bubble_hllc", case=1, nx=100, ny=40, x=(0.0, 0.225), y=(0.0, 0.089),
                      gamma=1.4, cfl=0.4, t_end=10.0, bc=[1, 1, 1, 1], solver=1, steps=40))
    specs.append(dict(name="ot_fault_hlld", nx=32, ny=32, cfl=0.4, t_end=10.0, solver=2, steps=10,
                      faults=[(10, 12, 0, float("nan")), (20, 7, 7, -5.0), (5, 25, 1, float("inf"))], **ot))
    specs.append(dict(name="ot_fault_force", nx=32, ny=32, cfl=0.4, t_end=10.0, solver=3, steps=10,
                      faults=[(10, 12, 0, float("nan")), (20, 7, 7, -5.0)], **ot))
    specs.append(dict(name="ot_all_bad", nx=8, ny=8, cfl=0.4, t_end=10.0, solver=2, steps=5,
                      faults=[("all", 7, -1.0)], **ot))
    return specs


def apply_faults(U, faults, ng=2):
    for f in faults:
        if f[0] == "all":
            U[ng:-ng, ng:-ng, f[1]] = f[2]
        else:
            i, j, k, v = f
            U[j + ng, i + ng, k] = v


def run_traj(spec):
    lib = oracle.ref()
    g, U0 = ref_grid(spec["case"], spec["nx"], spec["ny"], *spec["x"], *spec["y"], spec["gamma"])
    shape = U0.shape
    if spec.get("faults"):
        apply_faults(U0, spec["faults"])
        lib.ref_grid_set(g, ptr(U0))
    lib.ref_set_gamma(spec["gamma"])
    lib.ref_set_ch(0.0)
    lib.ref_set_cr(0.18)
    tmp = lib.ref_grid_new(spec["nx"], spec["ny"], 2, *spec["x"], *spec["y"])
    new = lib.ref_grid_new(spec["nx"], spec["ny"], 2, *spec["x"], *spec["y"])
    lib.ref_grid_copy(g, tmp)
    lib.ref_grid_copy(g, new)
    n = spec["steps"]
    dts = np.zeros(n)
    chs = np.zeros(n)
    ns = C.c_int(0)
    t = C.c_double(0.0)
    code = lib.ref_run_loop(g, tmp, new, spec["cfl"], spec["t_end"], spec["solver"], ibuf(spec["bc"]), n,
                            ptr(dts), ptr(chs), C.byref(ns), C.byref(t))
    U = grid_get(g, shape)
    for h in (g, tmp, new):
        lib.ref_grid_free(h)
    k = ns.value
    return dict(code=code, nsteps=k, t=t.value, dts=dts[:k], chs=chs[:k], ic_hash=state_hash(U0),
                final_hash=state_hash(U), rho_mid=U[U.shape[0] // 2, :, 0].copy())


def make_grid(rng):
    out = {}
    meta = {}
    lib = oracle.ref()
    # Case table (ids 0..3 of src/test_cases.cpp) at n = 1, 2.
    tab = np.zeros((4, 2, 10))
    bcs = np.zeros((4, 2, 4), dtype=np.int32)
    for cid in range(4):
        for n in (1, 2):
            cfg = np.zeros(10)
            bc = (C.c_int * 4)()
            snaps = np.zeros(16)
            lib.ref_case_config(cid, n, ptr(cfg), bc, ptr(snaps))
            tab[cid, n - 1] = cfg
            bcs[cid, n - 1] = list(bc)
    out["case_table"], out["case_bc"] = tab, bcs

    # Analytic ICs at every cell (ghosts included) of each reference case, n = 1.
    ic_hashes = {}
    for cid in range(4):
        g = lib.ref_make_n_grid(cid, 1)
        cfg = tab[cid, 0]
        shape = (int(cfg[1]) + 4, int(cfg[0]) + 4, 9)
        ic_hashes[str(cid)] = state_hash(grid_get(g, shape))
        lib.ref_grid_free(g)
    meta["ic_hashes"] = ic_hashes

    # Boundary KATs on a random 7 x 6 interior grid.
    nx, ny = 7, 6
    G = rng.uniform(-2, 2, (ny + 4, nx + 4, 9))
    S = rng.uniform(-2, 2, (ny + 4, nx + 4, 9))
    combos = [[t] * 4 for t in range(4)] + [list(rng.integers(0, 4, 4)) for _ in range(12)]
    bc_out = np.zeros((len(combos), ny + 4, nx + 4, 9))
    for c, combo in enumerate(combos):
        h = lib.ref_grid_new(nx, ny, 2, 0.0, 1.0, 0.0, 1.0)
        lib.ref_grid_set(h, ptr(G))
        lib.ref_apply_boundary(h, ibuf([int(v) for v in combo]))
        bc_out[c] = grid_get(h, G.shape)
        lib.ref_grid_free(h)
    hs = lib.ref_grid_new(nx, ny, 2, 0.0, 1.0, 0.0, 1.0)
    hd = lib.ref_grid_new(nx, ny, 2, 0.0, 1.0, 0.0, 1.0)
    lib.ref_grid_set(hs, ptr(S))
    lib.ref_grid_set(hd, ptr(G))
    lib.ref_copy_ghost_cells(hs, hd)
    out["copy_ghost"] = grid_get(hd, G.shape)
    lib.ref_grid_free(hs)
    lib.ref_grid_free(hd)
    out.update(bc_G=G, bc_S=S, bc_combos=np.array(combos, dtype=np.int32), bc_out=bc_out)

    # One-step KAT with a non-trivial c_h on a random smooth state (exercises
    # compute_dt + advance + damping together).
    traj = {}
    for spec in traj_specs():
        r = run_traj(spec)
        print(f"  {spec['name']}: code={r['code']} steps={r['nsteps']} t={r['t']:.6g}", flush=True)
        key = spec["name"]
        out[f"{key}__dts"] = r["dts"]
        out[f"{key}__chs"] = r["chs"]
        out[f"{key}__rho_mid"] = r["rho_mid"]
        traj[key] = dict(spec={k: v for k, v in spec.items()}, code=r["code"], nsteps=r["nsteps"],
                         t=r["t"].hex(), ic_hash=r["ic_hash"], final_hash=r["final_hash"])
    meta["trajectories"] = traj
    return out, meta


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref/libmhdref.so is not built: run `make -C oracle ref` where the reference is mounted")
    rng = np.random.default_rng(20261203)
    point = make_point(rng)
    np.savez_compressed(os.path.join(OUT, "ref_point.npz"), **point)
    grid, meta = make_grid(rng)
    np.savez_compressed(os.path.join(OUT, "ref_grid.npz"), **grid)

    def fix(o):
        if isinstance(o, float) and o != o:
            return "nan"
        if isinstance(o, float) and o in (float("inf"), float("-inf")):
            return str(o)
        if isinstance(o, (list, tuple)):
            return [fix(v) for v in o]
        if isinstance(o, dict):
            return {k: fix(v) for k, v in o.items()}
        if isinstance(o, np.integer):
            return int(o)
        return o

    with open(os.path.join(OUT, "ref_meta.json"), "w") as f:
        json.dump(fix(meta), f, indent=1, sort_keys=True)
    print("wrote", sorted(os.listdir(OUT)))


if __name__ == "__main__":
    main()
