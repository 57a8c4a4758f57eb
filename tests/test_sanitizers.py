"""Sanitizer runs of the library's host C++ (no GPU): tests/sanitize/driver.cpp
exercises the CPU winding sweep behind mli_winding (four threads on disjoint
query spans, the reference's fan-out, python/igl/_lib.py:72-95), the
stripifier, the boundary-edge counter and the threaded bounding-box and
fast-winding tree builds and queries, compiled by g++ with
AddressSanitizer + UndefinedBehaviorSanitizer and, separately, with
ThreadSanitizer. A sanitizer report fails the run (non-zero exit)."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None,
                                reason="needs g++ and make")


@pytest.fixture(scope="module")
def drivers(tmp_path_factory):
    out = tmp_path_factory.mktemp("sanitize")
    proc = subprocess.run(["make", "-C", SAN, f"OUT={out}", "asan", "tsan", "-j2"],
                          capture_output=True, text=True)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    return out


@pytest.mark.parametrize("kind", ["asan", "tsan"])
def test_host_code_is_sanitizer_clean(drivers, kind):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["TSAN_OPTIONS"] = "halt_on_error=1:second_deadlock_stack=1"
    proc = subprocess.run([os.path.join(drivers, f"driver_{kind}")], capture_output=True,
                          text=True, env=env, timeout=300)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "all checks passed" in proc.stdout
    assert "runtime error" not in proc.stderr  # UBSan's report text
