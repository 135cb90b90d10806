"""Host code under sanitizers (CPU only): the oracle's thread pool, the
arena's allocator logic with concurrent callers and the expression analysis,
built with g++ -fsanitize=address,undefined and -fsanitize=thread from
tests/cpp/Makefile.sanitize and run by tests/cpp/sanitize_host.cpp."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")


@pytest.fixture(scope="module")
def built():
    if not shutil.which("g++") or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("g++ or the ROCm headers are absent")
    r = subprocess.run(["make", "-C", CPP, "-f", "Makefile.sanitize"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]


@pytest.mark.parametrize("binary,marker", [("sanitize_asan", "Sanitizer"), ("sanitize_tsan", "ThreadSanitizer")])
def test_host_code_clean_under_sanitizer(built, binary, marker):
    r = subprocess.run([os.path.join(CPP, binary)], capture_output=True, text=True, timeout=300)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert marker not in out, out[-4000:]
    assert out.count("PASS") == 3, out
