"""The learned eddy viscosity against a high-resolution truth run.

``run`` follows the learned-closure workload: the truth is this dycore at
dx / ratio on (nxc * ratio)^2 columns (3000x3000x60 at 1 km for nxc = 1000:
about 196 GB of HBM with one species, built on the GPU by
``cases_device.device_case``); the model grid is nxc^2 columns at dx.

1. spin the truth up for ``spinup_steps`` of its own (dt / ratio) steps;
2. coarse-grain it onto the model grid, form the training rows (features of
   the coarse state, the truth's sub-grid stresses) and train the network on
   the GPU (Adam, ``epochs`` passes);
3. advance the truth ``heldout_steps`` further and measure, on the held-out
   rows, the fraction of the sub-grid stress each diffusivity leaves
   unexplained (``closure.flux_error``: 0 perfect, 1 = no model) for the
   network, Smagorinsky and no model;
4. the plain high-resolution run as the reference: two model-grid domains start
   from the coarse-grained truth, one with the network's km, one with
   Smagorinsky; after ``compare_steps`` model steps (the truth advancing
   ``ratio`` times as many) each is compared with the coarse-grained truth
   (relative L2 of u, v, w, theta over the interior, and of their change
   over the window).

``python -m wrfdc.closure_truth --nxc 1000`` runs it at the workload's size
and prints one JSON object.
"""

from __future__ import annotations

import argparse
import json
import time

import numpy as np


def _rel(a, b) -> float:
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run(device: str = "cuda:0", nxc: int = 100, nz: int = 30, ratio: int = 3, dx: float = 3000.0, dt: float = 18.0,
        spinup_steps: int = 30, heldout_steps: int = 30, compare_steps: int = 10, epochs: int = 2, batch: int = 4096,
        lr: float = 2e-3, border: int = 10, log=None) -> dict:
    import torch

    from .cases_device import _fill_boundary_arrays, device_case
    from .closure import Network, coarse_grain, fit_norm, flux_error, training_rows

    def say(msg):
        if log:
            log(msg)

    out = {"nx_truth": nxc * ratio, "nx_model": nxc, "nz": nz, "ratio": ratio, "dx_truth": dx / ratio, "dx_model": dx}
    t0 = time.perf_counter()
    fine = device_case(nxc * ratio, nxc * ratio, nz, device=device, dx=dx / ratio, dt=dt / ratio, nmoist=1, nscalar=0, config_index=4)
    out["build_truth_s"] = time.perf_counter() - t0
    say(f"truth {nxc * ratio}x{nxc * ratio}x{nz} built in {out['build_truth_s']:.1f} s")
    t0 = time.perf_counter()
    fine.step(spinup_steps, 0.0)
    fine.synchronize()
    out["truth_ms_per_step"] = 1e3 * (time.perf_counter() - t0) / max(spinup_steps, 1)
    say(f"truth spun up: {spinup_steps} steps, {out['truth_ms_per_step']:.0f} ms per step")
    coarse = device_case(nxc, nxc, nz, device=device, dx=dx, dt=dt, nmoist=1, nscalar=0, config_index=4, member=1)

    def to_model_grid(dom):
        coarse_grain(fine, dom, ratio)
        _fill_boundary_arrays(dom.buffers, dom.dims)
        torch.cuda.synchronize()

    # training
    to_model_grid(coarse)
    norm = fit_norm(coarse)
    X, T = training_rows(coarse, norm, fine, ratio)
    out["training_rows"] = int(X.shape[0])
    net = Network(norm, device, seed=0)
    t0 = time.perf_counter()
    hist = net.fit(X, T, epochs=epochs, batch=batch, lr=lr, log=(lambda m: say("train " + m)))
    out["train_s"] = time.perf_counter() - t0
    out["training_loss"] = [float(h) for h in hist]
    del X, T
    torch.cuda.empty_cache()

    # held-out stress
    fine.step(heldout_steps, spinup_steps * fine.config.dt)
    to_model_grid(coarse)
    Xb, Tb = training_rows(coarse, norm, fine, ratio)
    tb = Tb.cpu().numpy()
    out["heldout_flux_error"] = {"network": flux_error(net.km(Xb, tb), tb), "smagorinsky": flux_error(tb[:, 5], tb),
                                 "no_model": flux_error(np.zeros(len(tb)), tb)}
    say(f"held-out flux error {out['heldout_flux_error']}")
    del Xb, Tb, tb
    torch.cuda.empty_cache()

    # two model runs compared with the truth
    smag = device_case(nxc, nxc, nz, device=device, dx=dx, dt=dt, nmoist=1, nscalar=0, config_index=4, member=2)
    to_model_grid(coarse)
    to_model_grid(smag)
    names = ["u_2", "v_2", "w_2", "t_2"]
    inner = (slice(border, nxc - border), slice(0, nz), slice(border, nxc - border))
    start = {n: a[inner] for n, a in coarse.download(names).items()}
    net.attach(coarse)
    dtbc0 = (spinup_steps + heldout_steps) * fine.config.dt
    coarse.step(compare_steps, dtbc0)
    smag.step(compare_steps, dtbc0)
    got_net = {n: a[inner] for n, a in coarse.download(names).items()}
    got_smag = {n: a[inner] for n, a in smag.download(names).items()}
    ck_net, ck_smag = coarse.checksums(), smag.checksums()
    fine.step(compare_steps * ratio, dtbc0)
    to_model_grid(smag)
    truth = {n: a[inner] for n, a in smag.download(names).items()}
    vs_truth = {}
    for n in names:
        vs_truth[n] = {
            "network": _rel(got_net[n], truth[n]), "smagorinsky": _rel(got_smag[n], truth[n]),
            "network_change": _rel(got_net[n] - start[n], truth[n] - start[n]),
            "smagorinsky_change": _rel(got_smag[n] - start[n], truth[n] - start[n]),
        }
    out["compare_steps"] = compare_steps
    out["truth_rel_l2"] = vs_truth
    out["nonfinite"] = int(ck_net[3] + ck_smag[3])
    say(f"against the truth after {compare_steps} model steps: {json.dumps(vs_truth)}")
    Network.detach(coarse)
    for d in (coarse, smag, fine):
        d.close()
    torch.cuda.empty_cache()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--nxc", type=int, default=1000)
    ap.add_argument("--nz", type=int, default=60)
    ap.add_argument("--spinup-steps", type=int, default=30)
    ap.add_argument("--heldout-steps", type=int, default=30)
    ap.add_argument("--compare-steps", type=int, default=10)
    ap.add_argument("--epochs", type=int, default=2)
    args = ap.parse_args(argv)
    res = run(nxc=args.nxc, nz=args.nz, spinup_steps=args.spinup_steps, heldout_steps=args.heldout_steps,
              compare_steps=args.compare_steps, epochs=args.epochs, log=lambda m: print("[closure_truth]", m, flush=True))
    print(json.dumps(res))


if __name__ == "__main__":
    main()
