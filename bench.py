"""Benchmark of the MI355X RK3 dycore: Mpoint-steps/s and % of HBM roofline.

Workload (default): the 400x400x60 plain dycore with 8 advected species
(qv, qc, qr, qi, qs, qg, Ni, Nr), positive-definite advection on, dt 18 s,
4 acoustic substeps, FP32, seeded synthetic ideal case (500 m hill,
u = 10 m/s + noise).  The state is resident in HBM before the timed region;
a step is one full RK3 step (3 stages, all acoustic substeps, diffusion,
species, lateral boundaries).

Prints one JSON line.  ``roofline`` covers the dominant kernel (algorithmic
bytes per launch / mean launch time from HIP events on the dycore's stream);
``step_roofline`` the whole step against the minimal fused traffic model.
``cpu_baseline`` times the CPU oracle's OpenMP build on every host core this
process may use (OMP_NUM_THREADS caps it), over tiles that cover the same
workload (a bounded sample: a few steps instead of the timed run).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "wrf-jax_amd"))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c1", choices=["c1", "c0"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--traffic-calibration", action="store_true",
                    help="after the run, one 256 MiB torch copy (known HBM traffic) for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only launch the dominant kernel --kernel-reps times after one step, then a 256 MiB "
                         "calibration copy (for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs)")
    return ap.parse_args()


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.lower().startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _tile_grid(n: int) -> tuple[int, int]:
    """px x py = n tiles, as square as n allows."""
    px = int(n ** 0.5)
    while n % px:
        px -= 1
    return n // px, px


def cpu_baseline(workload: str, nsteps: int = 4) -> dict:
    """The CPU oracle (FP32, OpenMP build) on all the host cores this process
    may use: the workload's domain cut into one tile per thread, each tile a
    domain of its own with the workload's spacing, species and vertical grid
    (specified boundaries, no halo exchange, so the CPU is spared the
    communication a tiled run would pay), all tiles stepped concurrently."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    from oracle import Oracle

    from wrfdc.cases import CONFIGS, make_case

    params = dict(CONFIGS[workload])
    nx, ny, nz = params.pop("nx"), params.pop("ny"), params.pop("nz")
    allowed = len(os.sched_getaffinity(0))
    threads = max(1, min(allowed, int(os.environ.get("OMP_NUM_THREADS", allowed))))
    px, py = _tile_grid(threads)
    tiles = [make_case(nx // px, ny // py, nz, member=q, **params) for q in range(px * py)]
    oracles = [Oracle(t.config, t.arrays, "f32_omp") for t in tiles]
    t0 = time.perf_counter()
    Oracle.step_many(oracles, nsteps, 0.0, threads)
    sec = time.perf_counter() - t0
    for o in oracles:
        o.close()
    npts = sum(t.config.nx * t.config.ny * t.config.nz for t in tiles)
    return {"value": npts * nsteps / sec / 1e6, "unit": "Mpoint-steps/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "sample": f"{nsteps} RK3 steps of the FP32 oracle (OpenMP, {threads} threads) on {px}x{py} tiles of "
                      f"{nx // px}x{ny // py}x{nz} covering the {nx}x{ny}x{nz} workload, {sec:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design, not an implementation)", file=sys.stderr)
        sys.exit(2)
    import torch

    from wrfdc import bytes_model as roofline
    from wrfdc.cases import named_case
    from wrfdc.device import Domain

    t0 = time.perf_counter()
    case = named_case(args.workload)
    cfg = case.config
    print(f"[bench] case {args.workload} {cfg.nx}x{cfg.ny}x{cfg.nz} built in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    dom = Domain(cfg, "cuda:0")
    dom.upload(case.arrays)
    n_mass = cfg.nx * cfg.ny * cfg.nz
    n_species = cfg.nmoist + cfg.nscalar

    if args.traffic_probe:
        dom.step(1, 0.0)
        kern = roofline.dominant_kernel(cfg)
        for _ in range(args.kernel_reps):
            dom.stage(kern.stage, kern.rk_step, kern.iteration)
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] traffic probe: {args.kernel_reps} x {kern.name}, then a copy of {src.numel() * 4} bytes",
              file=sys.stderr, flush=True)
        dom.close()
        return

    for w in range(args.warmup):
        dom.step(1, w * cfg.dt)
    dom.synchronize()
    torch.cuda.synchronize()
    print("[bench] warmup done", file=sys.stderr, flush=True)

    stream = dom.stream
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ev0.record(stream)
    dom.step(args.steps, args.warmup * cfg.dt)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    gpu_ms = ev0.elapsed_time(ev1)
    ms_per_step = gpu_ms / args.steps
    ck = dom.checksums()  # raises FloatingPointError on a non-finite state
    print(f"[bench] {args.steps} steps: {gpu_ms:.1f} ms (wall {wall * 1e3:.1f} ms)", file=sys.stderr, flush=True)

    # dominant kernel, timed live with HIP events on the dycore's stream
    kern = roofline.dominant_kernel(cfg)
    dom.stage(kern.stage, kern.rk_step, kern.iteration)  # warm
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.kernel_reps):
        dom.stage(kern.stage, kern.rk_step, kern.iteration)
    ev1.record(stream)
    torch.cuda.synchronize()
    k_ms = ev0.elapsed_time(ev1) / args.kernel_reps
    k_bytes = kern.bytes(cfg)
    achieved = k_bytes / (k_ms * 1e-3) / 1e9
    traffic = roofline.measured_traffic(kern.name, args.workload)

    step_bytes = roofline.step_bytes(n_mass, n_species)
    step_gbs = step_bytes / (ms_per_step * 1e-3) / 1e9

    out = {
        "metric": "dycore Mpoint-steps/s + % HBM roofline, plain vs batched/learned",
        "value": n_mass / (ms_per_step * 1e-3) / 1e6,
        "unit": "Mpoint-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain dycore {cfg.nx}x{cfg.ny}x{cfg.nz}, {n_species} advected species, PD on, "
                               f"dt {cfg.dt:g} s, {cfg.num_sound_steps()} acoustic substeps, FP32",
                   "form": "plain"},
        "roofline": {"bound": "hbm", "kernel": kern.name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "algorithmic_bytes": k_bytes,
                     "mean_launch_ms": k_ms},
        "step_roofline": {"algorithmic_bytes": step_bytes, "achieved": step_gbs, "frac": step_gbs / HBM_PEAK_GBS},
        "checksums": ck.tolist(),
    }
    dom.close()
    if args.traffic_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.workload)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
