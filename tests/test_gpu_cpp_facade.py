"""The reference's self-tests (UnitTests.cu) restated on the C++ facade
(alias-table-gpu_amd/cpp/atg.hpp) and run against libatg.so on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "alias-table-gpu_amd", "cpp", "atg_unit_tests")


def test_reference_unit_tests_on_cpp_facade(gpu):
    assert os.path.exists(EXE), "build the library first (__graft_entry__.build())"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=900)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "OK (0 failures)" in r.stdout
