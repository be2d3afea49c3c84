"""The CPU oracle built with AddressSanitizer and UndefinedBehaviorSanitizer (oracle/Makefile target
`sanitize`, host code only) runs every DP and NEP oracle routine on small frames without a finding:
a small box with every image enumerated, a two-type box under the nearest-image rule, and two NEP
frames, one replicated by expand_box (oracle/san_driver.c)."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


def test_oracle_clean_under_asan_ubsan():
    if shutil.which(os.environ.get("CC", "gcc")) is None:
        pytest.skip("no C compiler")
    b = subprocess.run(["make", "-s", "-B", "-C", ORACLE, "sanitize"], capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(ORACLE, "_san", "san_driver")], capture_output=True, text=True, timeout=300,
                       env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-6000:]
    assert "san_driver: ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
