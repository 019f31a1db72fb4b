"""The C++ host API (cumat_amd/cuMat) through its own test binaries
(tests/cpp/test_*.cpp, built by build()): on the CPU that every binary links
against the C ABI library and registers its tests; on the GPU that every
test passes."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "bin")
NAMES = ["test_core", "test_sparse", "test_cg", "test_eigen_interop"]
# built only where Eigen's headers are found (tests/cpp/Makefile)
OPTIONAL = {"test_eigen_interop"}


def _bin(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        if name in OPTIONAL:
            pytest.skip(f"{name} is built only where Eigen's headers are available")
        pytest.fail(f"{path} is not built (run __graft_entry__.build())")
    return path


@pytest.mark.parametrize("name", NAMES)
def test_cpp_binary_lists_tests(name):
    out = subprocess.run([_bin(name), "--list"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert len(out.stdout.strip().splitlines()) >= 5


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_cpp_binary_passes(name, gpu):
    print("running", name, flush=True)
    out = subprocess.run([_bin(name)], capture_output=True, text=True, timeout=300)
    print(out.stdout[-4000:])
    assert out.returncode == 0, out.stdout[-4000:] + out.stderr[-2000:]
    assert " 0 failed" in out.stdout
