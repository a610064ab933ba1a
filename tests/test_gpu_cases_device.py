"""The GPU-built ideal case (wrfdc.cases_device) against the host one.

* every deterministic field (terrain, base state, qv, the moist hydrostatic
  perturbation pressure, the boundary arrays formed from them) equals
  make_case's to FP32 rounding (the two evaluate the same FP64 formulas,
  torch's and NumPy's libm may differ in the last FP64 bit);
* the noise fields have the stated mean and spread;
* the lateral relaxation of the initial state toward its boundary arrays is
  zero (the boundary arrays hold the coupled edge state), and a step stays
  finite and parity-close to the oracle stepping the same arrays.
"""

import numpy as np
import pytest

from parity import REL_L2_STEP, STEP_DIAGNOSTICS, compare, rel_l2

pytestmark = pytest.mark.gpu

DETERMINISTIC = ("ht", "mub", "pb", "alb", "phb", "t_2", "t_1", "p", "al", "ph_2", "ph_1", "moist", "znw", "rdnw", "fnm", "c3h",
                 "t_bxs", "t_bye", "ph_bxe", "ph_bys", "qv_bxs", "qv_bye", "mu_bxs")


def test_device_case_matches_host_case(gpu):
    from wrfdc.cases import make_case
    from wrfdc.cases_device import device_case

    print("[cases_device] 40x36x12 on the GPU vs make_case", flush=True)
    kw = dict(dx=2000.0, dt=12.0, nmoist=1, nscalar=0, hill_height=400.0)
    host = make_case(40, 36, 12, **kw)
    dom = device_case(40, 36, 12, device=gpu, rows_per_chunk=7, **kw)
    got = dom.download()
    for name in DETERMINISTIC:
        e = rel_l2(got[name], host.arrays[name])
        assert e < 1e-6, (name, e)
    u = got["u_2"][:36, :12, :41]
    v = got["v_2"][:37, :12, :40]
    assert abs(float(u.mean()) - 10.0) < 0.05 and abs(float(u.std()) - 1.0) < 0.05
    assert abs(float(v.mean())) < 0.05 and abs(float(v.std()) - 1.0) < 0.05
    assert np.array_equal(got["u_1"], got["u_2"]) and np.array_equal(got["v_1"], got["v_2"])
    dom.close()


def test_device_case_relaxes_to_itself_and_steps(gpu):
    from oracle import Oracle
    from wrfdc.cases_device import device_case

    print("[cases_device] relaxation of the initial state, one step vs the oracle", flush=True)
    dom = device_case(40, 36, 12, device=gpu, dx=2000.0, dt=12.0, nmoist=1, nscalar=0, hill_height=400.0)
    dom.stage("rk_step_prep", 1, 0)
    dom.stage("relax_bdy_dry", 1, 0)
    got = dom.download(["ru_tendf", "rv_tendf", "t_tendf", "ph_tendf", "mu_tend"])
    for name, a in got.items():
        assert float(np.abs(a).max()) == 0.0, name
    dom.close()
    dom = device_case(40, 36, 12, device=gpu, dx=2000.0, dt=12.0, nmoist=1, nscalar=0, hill_height=400.0)
    start = dom.download()
    o = Oracle(dom.config, start)
    o.step(1, 0.0)
    dom.step(1, 0.0)
    assert dom.checksums()[3] == 0
    bad = compare(dom.download(), o.arrays, tol=REL_L2_STEP, skip=STEP_DIAGNOSTICS)
    o.close()
    dom.close()
    assert not bad, bad
