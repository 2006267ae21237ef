"""The engine's host numerics (FP64 operator tables of every order and
convention, Yukawa and Helmholtz operators, offset tables, the FP16 operator
split) built with g++ under AddressSanitizer and UndefinedBehaviorSanitizer
(tests/asan/Makefile) and run: any memory error or undefined behaviour ends
the run with a non-zero status and a sanitizer report."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

ASAN = os.path.join(ROOT, "tests", "asan")


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None, reason="needs g++ and make")
def test_host_numerics_under_asan_ubsan():
    b = subprocess.run(["make", "-s", "-C", ASAN], capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(ASAN, "_build", "host_checks")], capture_output=True, text=True,
                       timeout=600, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "0 failures" in r.stdout
