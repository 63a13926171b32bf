"""CPU-side checks of the drop-in boundary (no GPU needed).

* libmhd_hip.so loads and exports every function include/mhd_hip.h declares;
* the case table and the host-side analytic initial states of the product are
  bitwise those of the oracle (which is pinned to the reference);
* argument validation and error reporting of the C ABI.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest

import mhd_gpu_acc_amd as m
import oracle
from tests.util import assert_bitwise

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(mhd_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 19, names
    L = m.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


@pytest.mark.parametrize("case_id", range(7))
@pytest.mark.parametrize("n", [1, 2])
def test_case_table_matches_oracle(case_id, n):
    c = m.get_n_case_config(case_id, n)
    o = oracle.case_config(case_id, n)
    got = np.array([c.nx, c.ny, c.ng, c.x_min, c.x_max, c.y_min, c.y_max, c.cfl, c.t_end, c.gamma])
    exp = np.array([o.nx, o.ny, o.ng, o.x_min, o.x_max, o.y_min, o.y_max, o.cfl, o.t_end, o.gamma])
    assert_bitwise(got, exp, f"case {case_id}")
    assert c.bc.as_list() == list(o.bc)


def _params():
    p = m.ICParams()
    po = oracle.ICParams()
    vals = dict(gamma=1.83, b_scale=0.77, seed=987654321, n_modes=4, pert_amp=0.05, bump_amp=0.02,
                bump_sigma=0.05, bump_cx=0.41, bump_cy=0.63)
    for k, v in vals.items():
        setattr(p, k, v)
        setattr(po, k, v)
    modes = ((0.9, 1, 2, 0.1), (-0.4, -3, 1, 2.2), (0.3, 0, 4, 5.0), (0.7, 2, -2, 1.3))
    for q, (a, kx, ky, ph) in enumerate(modes):
        for obj in (p, po):
            obj.mode_amp[q], obj.mode_kx[q], obj.mode_ky[q], obj.mode_phase[q] = a, kx, ky, ph
    return p, po


@pytest.mark.parametrize("case_id", range(9))
def test_initial_states_match_oracle(case_id):
    base = case_id if case_id < 7 else 4
    c = m.get_n_case_config(base, 1)
    # keep the grids small: same domain and BCs, fewer cells
    c.nx, c.ny = (96, 40) if base != 2 and base != 6 else (800, 4)
    o = oracle.Case(c.nx, c.ny, 2, c.x_min, c.x_max, c.y_min, c.y_max, c.cfl, c.t_end, c.gamma,
                    tuple(c.bc.as_list()))
    p, po = _params()
    U = m.initial_state(case_id, c, p)
    Uo = oracle.init_state(case_id, o, po)
    assert_bitwise(U, Uo, f"IC {case_id}")


def test_errors_are_reported():
    cfg = m.get_n_case_config("orszag_tang_unit", 1)
    cfg.ng = 3
    with pytest.raises(m.MHDError) as e:
        m.Run(cfg)
    assert e.value.code == 1 and "ng" in str(e.value)
    with pytest.raises(ValueError):
        m.parse_case_id("no_such_case")
    with pytest.raises(ValueError):
        m.get_n_case_config(42, 1)
    with pytest.raises(ValueError):
        m.get_n_case_config(2, 0)


def test_no_device_is_an_error_not_a_fallback():
    n = C.c_int()
    m.lib().mhd_device_count(C.byref(n))
    if n.value > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(m.MHDError) as e:
        m.Run(m.get_n_case_config("brio_wu", 1))
    assert e.value.code == 6


def test_cpp_front_compiles_links_and_reports(tmp_path):
    """include/mhd_hip.hpp (the head/-shaped C++ front) compiles with g++, links against
    the library, and runs: without a GPU the run constructor throws the library's
    no-device error (no CPU fallback); with one it steps Brio-Wu from its device initial
    state and drives it to t_end in the reference driver's 978 steps."""
    import subprocess
    src = tmp_path / "front.cpp"
    src.write_text(r'''
#include <cstdio>
#include "mhd_hip.hpp"
int main() {
    try {
        mhd_hip::BoundaryConfig bc;  // Brio-Wu: transmissive in x, periodic in y (test_cases.cpp)
        bc.bottom = bc.top = mhd_hip::BoundaryType::Periodic;
        mhd_hip::Run run(800, 4, 0.0, 1.0, 0.0, 4.0 / 800.0, 2.0, 0.4, mhd_hip::RiemannSolver::HLLD, bc);
        run.init_state_device(MHD_CASE_BRIO_WU);
        double t = 0.0;
        for (int s = 0; s < 5; ++s) { const double dt = run.compute_dt(0.4); run.advance_second_order(dt); t += dt; }
        mhd_hip::Run bw(800, 4, 0.0, 1.0, 0.0, 4.0 / 800.0, 2.0, 0.4, mhd_hip::RiemannSolver::HLLD, bc);
        bw.init_state_device(MHD_CASE_BRIO_WU);
        const mhd_drive_result r = bw.drive(MHD_CASE_BRIO_WU, "brio_wu", 0.1, "", 1, 0.1, false);
        std::printf("ran %.17g %lld\n", t, r.steps);
    } catch (const std::runtime_error& e) {
        std::printf("error: %s\n", e.what());
    }
    return 0;
}
''')
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    libdir = os.path.join(root, "mhd_gpu_acc_amd")
    exe = tmp_path / "front"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(root, "include"),
                    str(src), "-o", str(exe), "-L" + libdir, "-lmhd_hip", "-Wl,-rpath," + libdir,
                    "-Wl,-rpath,/opt/rocm/lib"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120).stdout
    assert out.startswith("ran ") or "no HIP device" in out, out
    if out.startswith("ran "):
        assert out.split()[2] == "978", out  # the reference driver's Brio-Wu step count
