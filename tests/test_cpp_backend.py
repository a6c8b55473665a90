"""The C++ `HIP` backend tag: compiles on the CPU; its self-test runs on the GPU."""
import os
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp")


def test_cpp_backend_compiles():
    subprocess.run(["make", "-s", "-B"], cwd=HERE, check=True)
    assert os.path.exists(os.path.join(HERE, "hip_backend_selftest"))


@pytest.mark.gpu
def test_cpp_backend_selftest():
    exe = os.path.join(HERE, "hip_backend_selftest")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s"], cwd=HERE, check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=180)
    print(r.stderr)
    assert r.returncode == 0, r.stderr
    assert "ALL PASSED" in r.stderr
