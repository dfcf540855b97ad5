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


def test_interpolater_adapter_against_amrex_shim():
    """The amrex::Interpolater adapter of carpetx_amd/host/cxp_interpolater.hxx,
    compiled against tests/amrex_shim (AMReX's names and signatures), on the
    GPU: interp / interp_face / CoarseBox equal the C ABI's results for six
    operators (prolongate_per_group honoured), a cubic is reproduced, the
    reference's aborts fire, and the group-tag overrides of get_interpolator
    select the right operator."""
    b = os.path.join(ROOT, "tests", "amrex_shim", "interpolater_check")
    assert os.path.exists(b), "build() did not produce " + b
    r = subprocess.run([b], capture_output=True, text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "interpolater check ok" in r.stdout
