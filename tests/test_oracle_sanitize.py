"""The oracle's C code (plife_oracle.c, including its threaded phases that
restate backend/LoadDistributor.java:37-66) runs clean under
AddressSanitizer + UndefinedBehaviorSanitizer and under ThreadSanitizer, on
a driver that steps small 2D/3D worlds with 1 and 4 threads and checks that
both give the same bits and the brute-force neighbour sets (host code only)."""
from __future__ import annotations

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    return os.path.join(ORACLE, "_san")


@pytest.mark.parametrize("name", ["check_asan_ubsan", "check_tsan"])
def test_oracle_under_sanitizer(built, name):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0"
    env["TSAN_OPTIONS"] = "halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    r = subprocess.run([os.path.join(built, name)], capture_output=True, text=True, env=env,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "all ok" in r.stdout
    assert "WARNING: ThreadSanitizer" not in r.stderr
    assert "runtime error" not in r.stderr
