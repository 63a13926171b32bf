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
    # the bandwidth probe refuses bad arguments before touching a device: null, overlapping
    # or short buffers, a size that is not a multiple of 16 bytes, reps < 1
    L, t = m.lib(), (C.c_double * 3)()
    base, mib = 1 << 40, 1 << 20
    for a, b, nbytes, reps in ((0, base, mib, 1), (base, base + mib // 2, mib, 1), (base, base + 4 * mib, 1000, 1),
                               (base, base + 4 * mib, mib + 8, 1), (base, base + 4 * mib, mib, 0)):
        assert L.mhd_probe_bandwidth(a, b, nbytes, reps, t) == 1, (a, b, nbytes, reps)


def test_no_device_is_an_error_not_a_fallback():
    n = C.c_int()
    m.lib().mhd_device_count(C.byref(n))
    if n.value > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(m.MHDError) as e:
        m.Run(m.get_n_case_config("brio_wu", 1))
    assert e.value.code == 6


_FRONT_SRC = r'''
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>
#include "mhd_hip.hpp"
// argv[1]: directory for the driver's CSV files
int main(int argc, char** argv) {
    try {
        mhd_hip::BoundaryConfig bc;  // Brio-Wu: transmissive in x, periodic in y (test_cases.cpp)
        bc.bottom = bc.top = mhd_hip::BoundaryType::Periodic;
        const double t_end = 0.1;
        // the reference driver's loop (main_cpu.cpp:196-222) written against the front
        mhd_hip::Run run(800, 4, 0.0, 1.0, 0.0, 4.0 / 800.0, 2.0, 0.4, mhd_hip::RiemannSolver::HLLD, bc);
        run.init_state_device(MHD_CASE_BRIO_WU);
        double t = 0.0;
        long long steps = 0;
        while (t < t_end) {
            double dt = run.compute_dt(0.4);
            if (t_end - t < dt) dt = t_end - t;
            run.advance_second_order(dt);
            t += dt;
            ++steps;
        }
        bool rejected = false;
        try { run.compute_dt(0.5); } catch (const std::invalid_argument&) { rejected = true; }
        // the library's driver on a second run, writing the reference's CSV files
        mhd_hip::Run bw(800, 4, 0.0, 1.0, 0.0, 4.0 / 800.0, 2.0, 0.4, mhd_hip::RiemannSolver::HLLD, bc);
        bw.init_state_device(MHD_CASE_BRIO_WU);
        const mhd_drive_result r = bw.drive(MHD_CASE_BRIO_WU, "brio_wu", t_end, argc > 1 ? argv[1] : "", 1, 0.1,
                                            false);
        std::vector<double> a(804 * 8 * 9), b(804 * 8 * 9);
        run.download_to_aos(a.data());
        bw.download_to_aos(b.data());
        const bool same = std::memcmp(a.data(), b.data(), a.size() * sizeof(double)) == 0;
        std::printf("ran %lld %lld %d %d %.17g\n", steps, r.steps, (int)same, (int)rejected, t);
    } catch (const std::runtime_error& e) {
        std::printf("error: %s\n", e.what());
    }
    return 0;
}
'''


@pytest.fixture(scope="module")
def cpp_front(tmp_path_factory):
    """A program written against include/mhd_hip.hpp (the head/-shaped C++ front),
    compiled with g++ -Wall -Wextra -Werror and linked against the library."""
    import subprocess
    d = tmp_path_factory.mktemp("front")
    src = d / "front.cpp"
    src.write_text(_FRONT_SRC)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    libdir = os.path.join(root, "mhd_gpu_acc_amd")
    exe = d / "front"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(root, "include"),
                    str(src), "-o", str(exe), "-L" + libdir, "-lmhd_hip", "-Wl,-rpath," + libdir,
                    "-Wl,-rpath,/opt/rocm/lib"], check=True)
    return exe


def test_cpp_front_compiles_links_and_reports(cpp_front, tmp_path):
    """Without a GPU the run constructor throws the library's no-device error (no CPU
    fallback); the device path is the gpu-marked test below."""
    import subprocess
    n = C.c_int()
    m.lib().mhd_device_count(C.byref(n))
    if n.value > 0:
        pytest.skip("a HIP device is present: test_cpp_front_steps_brio_wu_on_the_gpu covers it")
    out = subprocess.run([str(cpp_front), str(tmp_path)], capture_output=True, text=True, timeout=120).stdout
    assert "no HIP device" in out, out


@pytest.mark.gpu
def test_cpp_front_steps_brio_wu_on_the_gpu(cpp_front, tmp_path):
    """Run in a fresh process: the C++ front's compute_dt / advance_second_order loop
    (the reference driver's, main_cpu.cpp:196-222) takes Brio-Wu to t = 0.1 in the
    reference driver's 978 steps and ends in the state the library's driver reaches;
    the driver's CSV files are byte for byte the reference driver's
    (tests/golden/driver_reference.json); compute_dt with a CFL other than the run's
    throws std::invalid_argument."""
    import hashlib
    import json
    import subprocess
    print("[gpu] C++ front: brio_wu hlld minmod to t=0.1", flush=True)
    out = subprocess.run([str(cpp_front), str(tmp_path)], capture_output=True, text=True, timeout=120).stdout
    assert out.startswith("ran "), out
    f = out.split()
    ref = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "driver_reference.json")))["brio_wu"]
    assert int(f[1]) == ref["steps"] == 978 and int(f[2]) == 978, out
    assert f[3] == "1", "the front's loop and the library's driver end in different states"
    assert f[4] == "1", "compute_dt with a different CFL was not rejected"
    for name, meta in ref["files"].items():
        ours = tmp_path / name.replace("_cpu_", "_gpu_")
        assert hashlib.sha256(ours.read_bytes()).hexdigest() == meta["sha256"], name
