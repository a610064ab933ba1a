"""Benchmark of the MI355X RK3 dycore: Mpoint-steps/s and % of HBM roofline.

Workload (default): the 400x400x60 plain dycore with 8 advected species
(qv, qc, qr, qi, qs, qg, Ni, Nr), positive-definite advection on, dt 18 s,
4 acoustic substeps, FP32, seeded synthetic ideal case (500 m hill,
u = 10 m/s + noise).  The state is resident in HBM before the timed region;
a step is one full RK3 step (3 stages, all acoustic substeps, diffusion,
species, lateral boundaries).

Forms (``--forms``, all three by default; the headline ``value`` is plain):
``forms.batched`` runs the 32-member ensemble of uneven domains (100^2..300^2
columns, 45..60 levels, dx 1..9 km) with one RK3 step of every member
captured in one HIP graph over the autotuned streams; ``forms.learned`` runs
1000x1000x60 at dx 3 km with the bf16 MLP eddy viscosity (trained first on a
coarse-grained 1 km run of this dycore) and the same domain with
Smagorinsky.  Each reports Mpoint-steps/s and its ratio to plain.

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
    ap.add_argument("--forms", default="plain,batched,learned",
                    help="comma list of plain (the headline value), batched (32-member ensemble, one captured "
                         "graph per step), learned (1000x1000x60 with the trained closure) and shard (one GPU's "
                         "share of the plain workload split 8 ways; not in the default list)")
    ap.add_argument("--form-steps", type=int, default=5, help="timed steps of the batched and learned forms")
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


def cpu_baseline(workload: str, nsteps: int = 12) -> dict:
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


def _log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def batched_form(steps: int, plain_value: float) -> dict:
    """32 seeded members of uneven domains (100^2..300^2 columns, 45..60
    levels, dx 1..9 km, 6 moist species), one RK3 step of all of them
    captured in one HIP graph over concurrent streams, the stream layout
    and the face-block tile shape chosen by the autotuner; value = sum of N_mass x steps / s."""
    import torch

    from wrfdc.device import Domain
    from wrfdc.ensemble import Ensemble, autotune, c2_members, eager_step_time, member_case, n_mass

    t0 = time.perf_counter()
    specs = c2_members()
    doms = []
    for m, spec in enumerate(specs):
        case = member_case(spec, m)
        d = Domain(case.config, "cuda:0")
        d.upload(case.arrays)
        doms.append(d)
        del case
    nm = n_mass(doms)
    _log(f"batched: {len(doms)} members, {nm / 1e6:.1f} M mass points, built in {time.perf_counter() - t0:.1f} s")
    eager = eager_step_time(doms, reps=2)
    nb, assign, table, tiles = autotune(doms, candidates=(1, 2, 4, 8, 16), reps=2,
                                        log=lambda m: _log("batched autotune: " + m))
    face_levels = min(tiles, key=lambda e: e[1])[0]
    ens = Ensemble(doms, assign, nb)
    ens.step(1, 0.0)
    ens.synchronize()
    t1 = time.perf_counter()
    ens.step(steps, float(specs[0].dt))
    ens.synchronize()
    sec = time.perf_counter() - t1
    bad = sum(int(d.checksums()[3]) for d in doms)  # raises on a non-finite member
    ens.close()
    for d in doms:
        d.close()
    torch.cuda.empty_cache()
    value = nm * steps / sec / 1e6
    _log(f"batched: {1e3 * sec / steps:.1f} ms per ensemble step, eager {1e3 * eager:.1f} ms")
    return {"value": value, "unit": "Mpoint-steps/s", "ms_per_step": 1e3 * sec / steps, "steps": steps, "n_mass": nm,
            "members": len(doms), "streams": nb, "eager_ms_per_step": 1e3 * eager, "graph_vs_eager": eager / (sec / steps),
            "vs_plain": value / plain_value, "nonfinite": bad,
            "face_levels": face_levels, "autotune_ms": {str(b): 1e3 * t for b, t in table},
            "autotune_face_levels_ms": {str(lg): 1e3 * t for lg, t in tiles},
            "workload": "32 seeded members, 100^2..300^2 x 45..60, dx 1..9 km, 6 moist species, PD on, FP32, "
                        "one captured RK3 step per replay"}


def shard_form(steps: int, plain_value: float, warmup: int = 2) -> dict:
    """One GPU's share of the 400x400x60 workload if it were split over 8
    GPUs (design only: no exchange is run): the same case on a 1 x 8 row slab
    (400 x 50 owned rows + 3 halo rows each side) and on a 2 x 4 tile
    (200 x 100 owned + 3 each side), each a domain of its own with the
    workload's species and spacing.  value = owned mass points x steps / s;
    ``computed_value`` counts the halo points too."""
    import torch

    from wrfdc.cases import named_case
    from wrfdc.device import Domain

    halo = 3
    out = {}
    for name, (ox, oy) in (("slab_1x8", (400, 50)), ("tile_2x4", (200, 100))):
        nx, ny = ox + (2 * halo if ox < 400 else 0), oy + 2 * halo
        case = named_case("c1", nx=nx, ny=ny)
        dom = Domain(case.config, "cuda:0")
        dom.upload(case.arrays)
        del case
        for w in range(warmup):
            dom.step(1, w * dom.config.dt)
        dom.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(dom.stream)
        dom.step(steps, warmup * dom.config.dt)
        ev1.record(dom.stream)
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / steps
        bad = int(dom.checksums()[3])
        nz = dom.config.nz
        dom.close()
        owned, computed = ox * oy * nz, nx * ny * nz
        out[name] = {"domain": f"{nx}x{ny}x{nz}", "owned": f"{ox}x{oy}x{nz}", "ms_per_step": ms,
                     "value": owned / (ms * 1e-3) / 1e6, "computed_value": computed / (ms * 1e-3) / 1e6,
                     "vs_plain_per_point": computed / (ms * 1e-3) / 1e6 / plain_value, "nonfinite": bad}
        _log(f"shard {name}: {nx}x{ny}x{nz}, {ms:.2f} ms per step")
    torch.cuda.empty_cache()
    out["unit"] = "Mpoint-steps/s"
    out["steps"] = steps
    out["workload"] = "one GPU's share of 400x400x60 split 8 ways (no exchange run), 8 species, PD on, FP32"
    return out


def learned_form(steps: int, plain_value: float, kernel_reps: int) -> dict:
    """The learned closure at 1000x1000x60, dx 3 km (1 moist species): a
    network trained on the coarse-grained sub-grid stresses of a 1 km run of
    this dycore replaces smag_km; the same domain is also stepped with
    Smagorinsky.  value = N_mass x steps / s with the network."""
    import torch

    from wrfdc.cases import make_case
    from wrfdc.closure import FLOP_PER_POINT, Network, row_count, train_on_truth
    from wrfdc.device import Domain

    t0 = time.perf_counter()
    net, hist = train_on_truth("cuda:0", epochs=30)
    _log(f"learned: network trained in {time.perf_counter() - t0:.1f} s (loss {hist[0]:.4f} -> {hist[-1]:.4f})")
    case = make_case(1000, 1000, 60, name="learned", config_index=4, dx=3000.0, dt=18.0, nmoist=1, nscalar=0,
                     hydrometeors=False)
    cfg = case.config
    dom = Domain(cfg, "cuda:0")
    dom.upload(case.arrays)
    del case
    nm = cfg.nx * cfg.ny * cfg.nz
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)

    def timed(n: int, dtbc0: float) -> float:
        torch.cuda.synchronize()
        ev0.record(dom.stream)
        dom.step(n, dtbc0)
        ev1.record(dom.stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / n

    dom.step(1, 0.0)
    smag_ms = timed(steps, cfg.dt)
    net.attach(dom)
    dom.step(1, (steps + 1) * cfg.dt)
    ms = timed(steps, (steps + 2) * cfg.dt)
    ck = dom.checksums()
    # the network's km pass (smag prep + MFMA kernel), timed live
    dom.stage("closure_km", 1, 0)
    torch.cuda.synchronize()
    ev0.record(dom.stream)
    for _ in range(kernel_reps):
        dom.stage("closure_km", 1, 0)
    ev1.record(dom.stream)
    torch.cuda.synchronize()
    k_ms = ev0.elapsed_time(ev1) / kernel_reps
    Network.detach(dom)
    dom.close()
    torch.cuda.empty_cache()
    flops = FLOP_PER_POINT * row_count(cfg)
    tflops = flops / (k_ms * 1e-3) / 1e12
    value = nm / (ms * 1e-3) / 1e6
    _log(f"learned: {ms:.2f} ms per step with the network, {smag_ms:.2f} ms with Smagorinsky; km pass {k_ms:.3f} ms")
    return {"value": value, "unit": "Mpoint-steps/s", "ms_per_step": ms, "steps": steps, "n_mass": nm,
            "smagorinsky_ms_per_step": smag_ms, "vs_smagorinsky": smag_ms / ms, "vs_plain": value / plain_value,
            "training_loss": [hist[0], hist[-1]], "checksums": ck.tolist(),
            "roofline": {"bound": "mfma", "kernel": "k_closure_km (+ its k_smag prep)", "achieved": tflops,
                         "peak": 2500.0, "unit": "TFLOP/s", "frac": tflops / 2500.0, "flop_per_launch": flops,
                         "mean_launch_ms": k_ms, "dtype": "bf16 MFMA, FP32 accumulate",
                         "limited_by": "VALU issue (activations, bias, bf16 packing around the MFMA products)"},
            "workload": "1000x1000x60, dx 3 km, dt 18 s, 1 moist species, PD on, FP32 dycore, bf16 MLP km "
                        "(trained on a 1 km run coarse-grained 3x)"}


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
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    out["forms"] = {}
    if "batched" in forms:
        out["forms"]["batched"] = batched_form(args.form_steps, out["value"])
    if "learned" in forms:
        out["forms"]["learned"] = learned_form(args.form_steps, out["value"], args.kernel_reps)
    if "shard" in forms:
        out["forms"]["shard"] = shard_form(args.steps, out["value"])
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.workload)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
