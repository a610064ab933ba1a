"""The tridiagonal w solve in two GPU layouts (csrc/k_colsolve.h), measured
against each other on the same systems.

* one column per lane (64 columns per wave, serial sweeps in k: the layout
  of the column kernel) and one column per wavefront (lanes = levels, each
  sweep an affine scan in 6 shuffle steps) both match an FP64 Thomas solve
  of the same systems (relative error 1e-5 per column), and each other;
* their launch times at 400x400x60 (HIP events, 20 launches each) are
  printed with the bandwidth of 5 field passes per launch.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _thomas64(a, alpha, gamma, r):
    """FP64 sweeps of module_small_step_em.py:2949-2999 for one column (index 0 = level 1)."""
    kde = len(r)
    w = np.array(r, dtype=np.float64)
    for k in range(1, kde):
        w[k] = (w[k] - a[k] * w[k - 1]) * alpha[k]
    for k in range(kde - 2, 0, -1):
        w[k] = w[k] - gamma[k] * w[k + 1]
    return w


@pytest.mark.timeout(600)
def test_w_solve_layouts(gpu):
    import torch

    from wrfdc.cases import make_case
    from wrfdc.device import Domain

    print("[wsolve] 400x400x60: one column per lane vs one column per wavefront", flush=True)
    case = make_case(400, 400, 60, nmoist=1, nscalar=0)
    dom = Domain(case.config, gpu)
    dom.upload(case.arrays)
    dom.step(1, 0.0)  # realistic factors a, alpha, gamma (last acoustic substep) and w
    st = dom.download(["a", "alpha", "gamma", "w_2"])
    w0 = torch.tensor(st["w_2"], device=gpu)
    out = {}
    for name in ("w_solve:columns", "w_solve:levels"):
        dom.buffers["w_2"].copy_(w0)
        dom.stage(name, 3, 0)
        out[name] = dom.download(["w_2"])["w_2"]
    wc, wl = out["w_solve:columns"], out["w_solve:levels"]
    rng = np.random.default_rng(5)
    for _ in range(200):
        i, j = int(rng.integers(2, 400)), int(rng.integers(2, 400))  # 1-based interior column
        col = (j - 1, slice(None), i - 1)
        ref = _thomas64(st["a"][col].astype(np.float64), st["alpha"][col].astype(np.float64),
                        st["gamma"][col].astype(np.float64), st["w_2"][col].astype(np.float64))
        scale = max(np.abs(ref).max(), 1e-30)
        assert np.abs(wc[col] - ref).max() <= 1e-5 * scale, ("columns", i, j)
        assert np.abs(wl[col] - ref).max() <= 1e-5 * scale, ("levels", i, j)
    inner = (slice(1, 399), slice(None), slice(1, 399))
    d = np.linalg.norm((wc - wl)[inner]) / np.linalg.norm(wc[inner])
    print(f"[wsolve] columns vs levels: relative L2 {d:.2e}", flush=True)
    assert d < 1e-6
    # timing
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    cfg = case.config
    npts = (cfg.nx - 2) * (cfg.ny - 2) * (cfg.nz + 1)
    for name in ("w_solve:columns", "w_solve:levels"):
        dom.buffers["w_2"].copy_(w0)
        dom.stage(name, 3, 0)
        torch.cuda.synchronize()
        ev0.record(dom.stream)
        for _ in range(20):
            dom.stage(name, 3, 0)
        ev1.record(dom.stream)
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / 20
        print(f"[wsolve] {name}: {ms:.3f} ms per launch, {5 * 4 * npts / (ms * 1e-3) / 1e9:.0f} GB/s (5 field passes)", flush=True)
    dom.close()
