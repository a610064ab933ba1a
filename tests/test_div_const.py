"""The exact division by 12 and 60 used by every 5th/3rd-order flux.

The flux formulas of the reference divide by 12 and 60
(`dyn_em/module_advect_em_fast.py:165-254`).  The kernels replace the IEEE
division with a multiply by RN(1/c), one FMA correction and a rare path for
subnormal quotients (`div_const`, `wrfdc_common.h`).  This test compiles a
host restatement of those steps and compares it bit for bit with the CPU's
own IEEE division over every float with a subnormal or near-subnormal
quotient and a strided sample of all the others (both signs, zeros
included).  The device function itself is pinned by the bit-exact GPU parity
tests of the advection kernels.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_div_const_matches_ieee_division(tmp_path):
    exe = str(tmp_path / "div_const_check")
    subprocess.run(["gcc", "-O2", "-mfma", "-fopenmp", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "cdiv", "div_const_check.c"), "-lm"], check=True)
    env = dict(os.environ, OMP_NUM_THREADS=str(min(8, os.cpu_count() or 1)))
    res = subprocess.run([exe, "251"], capture_output=True, text=True, env=env, timeout=600)
    lines = res.stdout.strip().splitlines()
    assert len(lines) == 2, res.stdout + res.stderr
    for line in lines:
        fields = dict(f.split("=") for f in line.split())
        assert int(fields["checked"]) > 280_000_000, line
        assert int(fields["mismatches"]) == 0, line
    assert res.returncode == 0
