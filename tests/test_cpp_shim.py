"""The C++ host layer with the reference's function names (include/cfs_plume.hpp),
exercised by tests/cpp/test_shim.cpp against the FP64 oracle."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "build", "test_shim")


def _binary():
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp")], check=True)
    return BIN


def test_cpp_shim_host_only():
    r = subprocess.run([_binary(), "--no-gpu"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_shim_against_oracle(gpu):
    print("[case] C++ host layer: 21x14x10 r=5 encode, load, project, unproject, decode, submatrix", flush=True)
    r = subprocess.run([_binary()], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "7 of 7 tests passed" in r.stdout
