"""The drop-in C++ surface compiles from the reference's own header names: the command-line driver
(sbp_amd/host/main.cpp, the reference's src/main.cpp single-rank flow) and tests/dropin/api_check.cpp (the Delta,
entropy and Blockmodel calls as the reference's finetune and block-merge code makes them) build with g++ against
include/sbp/ and link to libsbp_amd.so.  No device is touched."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "sbp_amd", "libsbp_amd.so")


def gxx(src, out):
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include", "sbp"), src, "-o", out,
           "-L", os.path.join(ROOT, "sbp_amd"), "-lsbp_amd", "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib",
           "-Wl,-rpath," + os.path.join(ROOT, "sbp_amd")]
    return subprocess.run(cmd, capture_output=True, text=True, timeout=300)


@pytest.mark.skipif(not os.path.exists(LIB), reason="libsbp_amd.so is not built")
@pytest.mark.parametrize("src", ["sbp_amd/host/main.cpp", "tests/dropin/api_check.cpp"])
def test_compiles_against_reference_header_names(src, tmp_path):
    r = gxx(os.path.join(ROOT, src), str(tmp_path / "prog"))
    assert r.returncode == 0, r.stderr[-3000:]
    assert os.path.exists(tmp_path / "prog")
