"""include/lfdem_amd_stokes.hpp compiles against stand-ins for LF_DEM's
vec3d / DBlock / ODBlock (their member names, MatrixBlocks.h:5-84, vec3d.h)
and links to the C ABI; without a GPU, StokesSolver::init throws
std::runtime_error with the library's message, as the reference's error
paths throw (StokesSolver.cpp:683-687).  The same program solves a system
through the header on the GPU (tests/test_gpu_stokes_header.py)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(ROOT, "lf_dem_amd", "_lib")
SRC = os.path.join(HERE, "native", "stokes_hpp_check.cpp")


def compile_check(out):
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-o", out, SRC,
                    "-L" + LIBDIR, "-llfdem_amd", "-Wl,-rpath," + LIBDIR], check=True)


def test_header_compiles_and_throws_without_a_device(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "liblfdem_amd.so")):
        pytest.fail("liblfdem_amd.so is not built (run __graft_entry__.build())")
    exe = str(tmp_path / "stokes_hpp_check")
    compile_check(exe)
    from lf_dem_amd import capi
    if capi.lib().lfd_device_count() > 0:
        pytest.skip("a GPU is visible: the device path is tests/test_gpu_stokes_header.py")
    r = subprocess.run([exe, "nodevice"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok init threw: StokesSolver::init: lfd_stokes_create")
