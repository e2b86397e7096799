"""The near-field preconditioner's fp8 encoder (OCP E4M3FN, round to nearest
even, subnormals, saturation at 448) against brute-force nearest rounding:
tests/fp8_encode_check.cpp is built with hipcc as a host program around the
same header the encoding kernel includes, and run on the CPU."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "specklesim_amd", "csrc")


def test_e4m3_encoder_nearest_even(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not found")
    exe = str(tmp_path / "fp8_check")
    subprocess.run([hipcc, "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17",
                    "-I", CSRC, os.path.join(HERE, "fp8_encode_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "mismatches 0" in out.stdout, out.stdout
