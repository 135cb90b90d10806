"""bench.py's command line and its kernel-symbol table, on the CPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
import shtensor_amd as sht  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)


def test_every_kernel_class_has_a_profiler_symbol():
    """The traffic pass finds the dominant kernel in rocprofv3's output by
    this table: every class name the library reports must be in it."""
    lib = sht.lib()
    names = set()
    for k in list(range(0, 64)) + list(range(100, 164)):
        n = lib.sht_kclass_name(k).decode()
        if n != "unknown":
            names.add(n)
    assert names, "no kernel class names"
    missing = sorted(n for n in names if n not in bench.KERNEL_SYMBOL)
    assert not missing, missing
