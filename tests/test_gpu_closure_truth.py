"""The learned closure against a high-resolution truth run (wrfdc.closure_truth),
at a reduced size (truth 180x180x20 at 1 km, model 60x60x20 at 3 km): the
training loss falls, the network leaves less of the held-out sub-grid stress
unexplained than Smagorinsky, and the model runs with the network and with
Smagorinsky stay finite with errors against the coarse-grained truth of the
same order.  The workload's size (3000x3000x60 truth) runs with
``python -m wrfdc.closure_truth``."""

import math

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(600)
def test_closure_against_truth_run(gpu):
    from wrfdc.closure_truth import run

    print("[closure_truth] truth 180x180x20 at 1 km, model 60x60x20 at 3 km", flush=True)
    res = run(gpu, nxc=60, nz=20, spinup_steps=20, heldout_steps=20, compare_steps=5, epochs=8, batch=1024,
              log=lambda m: print("[closure_truth]", m, flush=True))
    loss = res["training_loss"]
    assert loss[-1] < loss[0]
    fe = res["heldout_flux_error"]
    assert all(math.isfinite(v) for v in fe.values())
    assert fe["network"] < fe["smagorinsky"], fe
    assert res["nonfinite"] == 0
    for n, e in res["truth_rel_l2"].items():
        assert math.isfinite(e["network"]) and math.isfinite(e["smagorinsky"]), n
        assert e["network"] < 10.0 * e["smagorinsky"] + 1e-6, (n, e)
