"""The CPU oracle and the header-only C++ interface (FKDPoint.h, FKDTree.h)
under AddressSanitizer and UndefinedBehaviorSanitizer: `make -C oracle
sanitize` builds oracle/san_check.cpp with fkd_oracle.cpp, and its run must
end cleanly (no sanitizer report, every result check passed)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_san", "san_check")


@pytest.fixture(scope="module")
def san_bin():
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("sanitizer build failed:\n" + r.stdout + r.stderr)
    return BIN


def test_binary_is_instrumented(san_bin):
    libs = subprocess.run(["ldd", san_bin], capture_output=True, text=True).stdout
    assert "libasan" in libs and "libubsan" in libs, libs


def test_oracle_and_interface_clean_under_sanitizers(san_bin):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([san_bin], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
    assert "all checks passed" in r.stdout
