"""The C++ drop-in (include/shtensor_amd.hpp) exercised by a compiled host
program: the reference's test_kernel.cpp / test_tensor.cpp scenarios with
std::vector data, a device-pointer call and a reserve -> contract flow, each
checked bitwise against the oracle (tests/cpp/test_dropin.cpp, built by
__graft_entry__.build() / make -C tests/cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_dropin")


@pytest.mark.gpu
def test_cpp_dropin_program():
    assert os.path.exists(BIN), "tests/cpp/test_dropin is not built: run make -C tests/cpp"
    print("running", BIN, flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=100)
    print(r.stdout[-4000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    assert r.stdout.count("PASS") >= 19


def test_cpp_header_compiles(tmp_path):
    """The header is self-contained C++17 over sht.h (no HIP headers)."""
    src = tmp_path / "t.cpp"
    src.write_text('#include "shtensor_amd.hpp"\n'
                   'int main() { Shtensor::VArray<2> v; (void)v; return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        "-I" + os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
