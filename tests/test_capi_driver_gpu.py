"""The C++ caller (tests/capi/capi_driver.cpp, built by terraneo_amd/build.py)
drives apply -> GCA -> GMG -> Stokes FGMRES through the header alone, in its
own process; it checks operator symmetry and the solve's true residual."""
import os
import subprocess

import pytest

from conftest import ROOT

DRIVER = os.path.join(ROOT, "tests", "capi", "capi_driver")


@pytest.mark.gpu
def test_cpp_caller_through_header(cuda):
    assert os.path.exists(DRIVER), "C++ driver not built (python -m terraneo_amd.build)"
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=150)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert "capi_driver OK" in r.stdout


def test_cpp_caller_is_built():
    # the build step compiles it on the CPU (no GPU needed)
    assert os.path.exists(DRIVER)
