"""The C++ host API end to end: carpetx_amd/host/prolongate_check, built
by build() against libcxp.so, runs the one-box configuration (VC order-3
prolongation of a tensor cubic, injection round trip, work list and graph
replay) and TestProlongate-style polynomial exactness of every operator
without limiter branches, and exits 0 when all checks pass."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "carpetx_amd", "host", "prolongate_check")


def test_cpp_host_check_program():
    assert os.path.exists(BIN), "build() did not produce " + BIN
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
