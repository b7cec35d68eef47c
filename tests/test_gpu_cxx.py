"""The C++ drop-in API (include/awk/*.h over the C-ABI) on the GPU: the reference's
DiffOperRoutineTests GPU round trip, the apply/splat adjoint identity of
cudaHField3DUtilsRoutineTests, GreedyIteratorGPU iterations, and LDMMIteratorGPU / LDMMAtlasGPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cxx_api_program():
    exe = os.path.join(ROOT, "tests", "cxx", "test_cxx_api")
    assert os.path.exists(exe), "build() compiles tests/cxx/test_cxx_api"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "CXX API TEST OK" in r.stdout
