"""The StokesSolver drop-in header (include/lfdem_amd_stokes.hpp) driven by
a C++ program on the GPU: reset with drag blocks, columns whose rows arrive
in descending order, fill, RHS per particle, solve; R x = b holds to 1e-11
in FP64 on the host, and an undersized solution vector throws."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "stokes_hpp_check")


def test_header_solves_on_the_gpu(capi):
    if not os.path.exists(EXE):
        pytest.fail("tests/native/stokes_hpp_check is not built (run __graft_entry__.build())")
    print("stokes header check", flush=True)
    r = subprocess.run([EXE, "solve"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.startswith("ok residual")
