"""FP64 log and pow of calc_p_rho_phi (`csrc/k_f64math.h`).

The kernel rounds log(pfd / pfu) and temp^(cp/cv) to FP32 from an FP64
evaluation written from the series (atanh series for log, Taylor series for
exp).  This test compiles the header itself for the host and compares, for
every FP32 input of the ranges the kernel sees (pfd / pfu in (1, 2); temp in
[1/16, 2) with the exponent 1.4f), the FP32-rounded result with the C
library's FP64 log and pow rounded the same way: no mismatch allowed, and
the FP64 values within a few ulp.  The device build of the same header is
pinned by the GPU parity tests of the calc_p_rho_phi stage and the step.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "wrf-jax_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_f64_log_pow_match_c_library(tmp_path):
    exe = str(tmp_path / "f64math_check")
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-I", CSRC, "-o", exe, os.path.join(HERE, "cf64", "f64math_check.cpp"), "-lm"],
                   check=True)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    lines = res.stdout.strip().splitlines()
    assert len(lines) == 2, res.stdout
    limits = {"log": (16_000_000, 2), "pow": (41_000_000, 8)}
    for line in lines:
        name, *rest = line.split()
        f = dict(x.split("=") for x in rest)
        n_min, ulp_max = limits[name]
        assert int(f["checked"]) >= n_min, line
        assert int(f["mismatches"]) == 0, line
        assert int(f["max_ulp"]) <= ulp_max, line
