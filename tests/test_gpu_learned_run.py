"""The learned-closure workload run to its length: 1000x1000x60 at dx 3 km,
200 RK3 steps (one simulated hour), once with the trained network as km and
once with Smagorinsky, from the same seeded state built on the GPU.

Checks on both runs: no value turns non-finite; dry-air mass and total
energy stay within 1e-4 of their start (the boundary values are the initial
edge state with zero tendency, so only the relaxation zone and the
diffusion change them); winds stay below 60 m/s.  The distance between the
two runs is printed (relative L2 over the interior).  It must not be zero
(the network's km is in use).  The mean flow (u, theta) must stay within
2 % of the Smagorinsky run.  v and w start as zero-mean noise that the two
closures damp differently, so they only have to stay within 50 %; the first
run measured u 0.4 %, theta 0.05 %, v 14 % and w 19 %.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_learned_closure_200_steps_1000x1000x60(gpu):
    import torch

    from wrfdc.cases_device import device_case
    from wrfdc.closure import Network, train_on_truth

    print("[learned run] training the network on a coarse-grained 1 km run", flush=True)
    net, hist = train_on_truth(gpu, epochs=30)
    nx = ny = 1000
    nz, steps, block = 60, 200, 50
    runs = {}
    for name in ("network", "smagorinsky"):
        print(f"[learned run] {name}: {nx}x{ny}x{nz}, {steps} steps", flush=True)
        dom = device_case(nx, ny, nz, device=gpu, config_index=4, member=0, dx=3000.0, dt=18.0, nmoist=1)
        if name == "network":
            net.attach(dom)
        ck0 = dom.checksums()
        print(f"[learned run] {name} at start: dry mass {ck0[4]:.9e} kg, energy {ck0[5]:.9e} J", flush=True)
        for b in range(steps // block):
            dom.step(block, b * block * dom.config.dt)
            ck = dom.checksums()  # raises FloatingPointError on NaN/Inf
            print(f"[learned run] {name} after {(b + 1) * block} steps: dry mass {ck[4]:.9e} kg, energy {ck[5]:.9e} J",
                  flush=True)
        got = dom.download(["u_2", "v_2", "w_2", "t_2"])
        if name == "network":
            Network.detach(dom)
        dom.close()
        torch.cuda.empty_cache()
        assert ck[3] == 0, name
        assert abs(ck[4] - ck0[4]) / ck0[4] < 1e-4, name
        assert abs(ck[5] - ck0[5]) / ck0[5] < 1e-4, name
        for n in ("u_2", "v_2", "w_2"):
            assert np.abs(got[n][:ny, :nz, :nx]).max() < 60.0, (name, n)
        runs[name] = {n: a[10:ny - 10, :nz, 10:nx - 10] for n, a in got.items()}
    dist = {n: _rel(runs["network"][n], runs["smagorinsky"][n]) for n in ("u_2", "v_2", "w_2", "t_2")}
    print(f"[learned run] network vs Smagorinsky after {steps} steps (relative L2, interior): {dist}", flush=True)
    for n, d in dist.items():
        assert 0.0 < d < (0.02 if n in ("u_2", "t_2") else 0.5), (n, d)
