"""The C++ drop-in API (include/awk/*.h over the C-ABI) on the GPU: KernelFactory and
the reference's DiffOperRoutineTests GPU round trip, host/device vector-field
copies, the apply/splat adjoint identity of cudaHField3DUtilsRoutineTests, a
GreedyWarp-shaped template over GreedyIteratorGPU reproducing the reference's
GreedyWarp regression output, an LDMMWarp-shaped template over LDMMIteratorGPU,
and LDMMAtlasGPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cxx_api_program():
    exe = os.path.join(ROOT, "tests", "cxx", "test_cxx_api")
    assert os.path.exists(exe), "build() compiles tests/cxx/test_cxx_api"
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden")], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "CXX API TEST OK" in r.stdout
