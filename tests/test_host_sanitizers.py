"""The device-free host layer (configuration parser, component factory
and small components, response helpers, ROI builders, frame-file
archive) built with g++ under AddressSanitizer + UndefinedBehavior-
Sanitizer and under ThreadSanitizer (tests/sanitize/Makefile), running
tests/sanitize/host_checks.cxx: serial checks under ASan/UBSan, eight
threads at once under TSan.  Any sanitizer report or failed check fails
the test."""
import os
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-j2", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        if "sanitize" in r.stderr and ("unrecognized" in r.stderr or "cannot find" in r.stderr):
            pytest.skip("this g++ has no sanitizer runtime")
        raise AssertionError(r.stderr[-4000:])
    return HERE


@pytest.mark.parametrize("binary,mode", [("host_asan", "serial"), ("host_tsan", "threads")])
def test_host_layer_under_sanitizers(built, tmp_path, binary, mode):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               TSAN_OPTIONS="halt_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(built, "_build", binary), mode, str(tmp_path)], capture_output=True,
                       text=True, env=env, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    for word in ("AddressSanitizer", "ThreadSanitizer", "runtime error", "LeakSanitizer"):
        assert word not in out, out[-4000:]
    assert f"{mode}: 0 failed checks" in out
