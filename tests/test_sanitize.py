"""The host C++ of libvba and the CPU oracle under AddressSanitizer,
UndefinedBehaviorSanitizer and ThreadSanitizer (tests/sanitize/).

The driver (tests/sanitize/host_check.cpp) runs the seeded generators, the
model build and support check, the oracle's dense, sparse and matrix-free LM
on one and three threads, the point elimination, the COLMAP text/binary and
problem-dump round trips, and truncated or corrupted COLMAP binaries and dumps
(each must be an error, not a crash). Device code is not built: these are host
compiler builds of the host translation units only.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")
FIXTURE = os.path.join(HERE, "golden", "colmap_text")


@pytest.fixture(scope="module")
def built():
    if shutil.which("g++") is None or shutil.which("make") is None:
        pytest.skip("g++ and make are needed for the sanitizer builds")
    r = subprocess.run(["make", "-C", SAN, "-j8"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    return os.path.join(SAN, "_build")


@pytest.mark.parametrize("binary", ["host_check", "host_check_tsan"])
def test_host_code_is_clean_under_sanitizers(built, binary, tmp_path):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["TSAN_OPTIONS"] = "halt_on_error=1:second_deadlock_stack=1"
    r = subprocess.run([os.path.join(built, binary), FIXTURE, str(tmp_path)], capture_output=True, text=True,
                       timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-6000:]
    assert "host_check: 0 failed checks" in r.stdout
    for marker in ("runtime error", "ERROR: AddressSanitizer", "ERROR: LeakSanitizer", "WARNING: ThreadSanitizer"):
        assert marker not in out, out[-6000:]
