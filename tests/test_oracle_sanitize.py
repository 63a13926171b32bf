"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer (host only):
oracle/_san/sanitize drives every oracle entry point the tests use over small grids,
including the failure paths (a NaN cell, a grid with no signal speed, HLLC Brio-Wu
running into the dt floor), and aborts on the first report."""
from __future__ import annotations

import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_san", "sanitize")


@pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_san/sanitize not built (make -C oracle sanitize)")
def test_oracle_is_clean_under_asan_and_ubsan():
    # verify_asan_link_order=0: the sanitizer runtime need not be the first library
    # loaded when the environment preloads another one
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="4")
    p = subprocess.run([BIN], capture_output=True, text=True, env=env, timeout=600)
    assert p.returncode == 0, p.stderr[-4000:]
    assert "sanitize: ok" in p.stdout
    assert "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr, p.stderr[-4000:]
