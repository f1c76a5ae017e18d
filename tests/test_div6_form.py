"""The float division by 6 of the device QUICK scheme (the 1/6 of src/Flow.jl:5) is a reciprocal
multiply with one FMA correction, kept when the quotient is a normal number.  This CPU test
checks that form against IEEE x/6 for all 2^32 floats (tests/c/div6_scan.c); the device build is
checked the same way by wl_self_test in the GPU tests."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_div6_form_equals_ieee_for_all_floats(tmp_path):
    exe = str(tmp_path / "div6_scan")
    subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", "-fopenmp", "-o", exe,
                    os.path.join(HERE, "c", "div6_scan.c"), "-lm"], check=True)
    env = dict(os.environ, OMP_NUM_THREADS=str(min(8, os.cpu_count() or 1)))
    out = subprocess.run([exe], check=True, capture_output=True, text=True, env=env, timeout=600)
    assert out.stdout.strip() == "0"
