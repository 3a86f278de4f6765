"""The C++ host API (mckl_amd/include/mckl_hip.hpp) driven by a compiled
test program, as the reference's example programs drive its headers."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "host_api_test")


def test_host_api_program(dev):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/host_api_test is not built (python -m mckl_amd.build)")
    kat = os.path.join(ROOT, "tests", "golden", "random_philox4x32.data")
    r = subprocess.run([EXE, kat], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
