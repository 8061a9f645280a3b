"""Runs the C++ plugin-adapter test (tests/cpp/test_kernel_precision.cpp): the reference's
TestKernelPrecisionCUDA cases against HIPExactKernel<double, D> through the C ABI, plus
Clone, error and multi-threaded-subject behaviour."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_kernel_precision")


@pytest.mark.parametrize("precision", ["fp64", "fp32c", "fp32x"])
def test_cpp_adapter(precision):
    if not os.path.exists(EXE):
        pytest.fail("C++ adapter test is not built; run __graft_entry__.build()")
    r = subprocess.run([EXE, precision], capture_output=True, text=True, timeout=120)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
