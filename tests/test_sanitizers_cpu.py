"""CPU-only sanitizer builds (tests/native/Makefile): the oracle under
AddressSanitizer + UndefinedBehaviorSanitizer (with its OpenMP loops), and
the host code that starts threads of its own -- the stager's copy pool
(fortkernels-_amd/csrc/gpp_stage.cpp), the input generator
(gpp_inputs.cpp) and the FF planning's worker pool (host_pool.h) -- under
ThreadSanitizer (and ASan for the pools).  Each
driver must exit 0 with no sanitizer report.  The reference has real races
(gppKer.cpp:67,97) and an out-of-bounds read (:288) that the oracle's
restatement does not repeat."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

NATIVE = ROOT / "tests" / "native"
DRIVERS = ["oracle_asan", "copypool_tsan", "copypool_asan", "inputs_tsan",
           "hostpool_tsan", "hostpool_asan"]


@pytest.fixture(scope="module")
def built():
    if not (shutil.which("gcc") and shutil.which("g++") and shutil.which("make")):
        pytest.skip("no host compilers")
    subprocess.run(["make", "-s", "-j4", "-C", str(NATIVE)], check=True,
                   capture_output=True, text=True)
    return NATIVE / "_build"


@pytest.mark.parametrize("name", DRIVERS)
def test_sanitized_driver_is_clean(built, name):
    env = dict(os.environ, OMP_NUM_THREADS="4",
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([str(built / name)], env=env, capture_output=True,
                       text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    for marker in ("ERROR: AddressSanitizer", "WARNING: ThreadSanitizer",
                   "runtime error:", "ERROR: LeakSanitizer"):
        assert marker not in out, out[-4000:]
    assert "sanitizer run ok" in out
