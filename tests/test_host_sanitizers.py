"""Host code (the engine's host sources and the oracle) built with g++ under
AddressSanitizer and UndefinedBehaviorSanitizer and run on ordinary, edge
and error inputs (tests/cpp/host_sanitize.cpp).  CPU only: no device code is
built with the sanitizers."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_code_under_asan_and_ubsan():
    subprocess.run(["make", "-C", CPP, "host_sanitize"], check=True, stdout=subprocess.DEVNULL)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="4")
    out = subprocess.run([os.path.join(CPP, "host_sanitize")], capture_output=True, text=True, timeout=600, env=env)
    assert out.returncode == 0, out.stdout[-4000:] + out.stderr[-4000:]
    assert "host_sanitize: ok" in out.stdout
    assert "runtime error" not in out.stderr and "AddressSanitizer" not in out.stderr
