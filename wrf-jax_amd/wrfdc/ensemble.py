"""Batched ensembles of uneven domains advanced by one captured graph.

An ``Ensemble`` holds member ``Domain``s (each its own grid, dx, dt and
species) and the HIP graph in which one RK3 step of every member was
captured (``wrfdc_ensemble_create``): members on the same stream run one
after another, the streams concurrently, and each replay
(``wrfdc_ensemble_step``) launches the whole ensemble's step at once.
``autotune`` picks how many streams to use, which member goes on which and
the tile shape of the face-block stencil kernels: it times each member's
step on its own, assigns members longest first to the least loaded stream
for each candidate stream count, times the captured ensemble step of each
layout and keeps the fastest; then, on that layout, it times each candidate
level-group size of the face-block tiles (``Domain.set_tile("face_levels",
n)``, 0 = the per-kernel defaults) and leaves the members set to the
fastest.  Neither choice changes any result bit.

``c2_members`` draws the seeded members of the batched workload: 32
domains of 100^2..300^2 columns and 45..60 levels, dx 1..9 km, dt = 6 s per
km of dx, acoustic steps by WRF's formula, u = 10 + U(-5, 5) m/s and
N ~ U(0.008, 0.012) 1/s per member.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass(frozen=True)
class MemberSpec:
    nx: int
    ny: int
    nz: int
    dx: float
    dt: float
    u0: float
    bvf: float


def c2_members(n: int = 32, seed: int = 20261202 + 1000 * 2) -> list[MemberSpec]:
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(n):
        nxy = int(rng.integers(100, 301))
        nz = int(rng.integers(45, 61))
        dxkm = float(rng.uniform(1.0, 9.0))
        u0 = 10.0 + float(rng.uniform(-5.0, 5.0))
        bvf = float(rng.uniform(0.008, 0.012))
        out.append(MemberSpec(nxy, nxy, nz, 1000.0 * dxkm, 6.0 * dxkm, u0, bvf))
    return out


def member_case(spec: MemberSpec, member: int, nmoist: int = 6):
    from .cases import make_case

    return make_case(spec.nx, spec.ny, spec.nz, name=f"c2-{member}", config_index=2, member=member, dx=spec.dx, dt=spec.dt,
                     u0=spec.u0, bvf=spec.bvf, time_step_sound=0, nmoist=nmoist, nscalar=0)


class Ensemble:
    def __init__(self, domains, branch_of=None, nbranches: int = 1):
        self.lib = _lib.load()
        self.domains = list(domains)
        n = len(self.domains)
        for d in self.domains:
            d.synchronize()
        ctxs = (ctypes.c_void_p * n)(*[d.ctx.value for d in self.domains])
        br = None if branch_of is None else (ctypes.c_int * n)(*[int(b) for b in branch_of])
        h = ctypes.c_void_p()
        rc = self.lib.wrfdc_ensemble_create(ctxs, n, br, int(nbranches), ctypes.byref(h))
        _lib.raise_for(rc, self.lib, None, "wrfdc_ensemble_create")
        self.h = h
        self.nbranches = nbranches
        self.branch_of = list(branch_of) if branch_of is not None else [q % nbranches for q in range(n)]

    @property
    def nodes(self) -> int:
        return int(self.lib.wrfdc_ensemble_nodes(self.h))

    def step(self, nsteps: int = 1, dtbc0: float = 0.0) -> None:
        _lib.raise_for(self.lib.wrfdc_ensemble_step(self.h, int(nsteps), float(dtbc0)), self.lib, None, "wrfdc_ensemble_step")

    def synchronize(self) -> None:
        _lib.raise_for(self.lib.wrfdc_ensemble_sync(self.h), self.lib, None, "wrfdc_ensemble_sync")

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.wrfdc_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def n_mass(domains) -> int:
    return sum(d.config.nx * d.config.ny * d.config.nz for d in domains)


def eager_step_time(domains, reps: int = 3) -> float:
    """Seconds per ensemble step when the members are stepped one after
    another, each by wrfdc_step (the plain path)."""
    for d in domains:
        d.step(1, 0.0)
    for d in domains:
        d.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for d in domains:
            d.step(1, 0.0)
    for d in domains:
        d.synchronize()
    return (time.perf_counter() - t0) / reps


def member_times(domains, reps: int = 3) -> list[float]:
    out = []
    for d in domains:
        d.step(1, 0.0)
        d.synchronize()
        t0 = time.perf_counter()
        d.step(reps, 0.0)
        d.synchronize()
        out.append((time.perf_counter() - t0) / reps)
    return out


def assign_longest_first(times, nbranches: int) -> list[int]:
    """Members in decreasing time, each to the stream with the least load."""
    load = [0.0] * nbranches
    out = [0] * len(times)
    for q in sorted(range(len(times)), key=lambda q: -times[q]):
        b = min(range(nbranches), key=lambda b: load[b])
        out[q] = b
        load[b] += times[q]
    return out


def graph_step_time(ens: Ensemble, reps: int = 5) -> float:
    ens.step(1)
    ens.synchronize()
    t0 = time.perf_counter()
    ens.step(reps)
    ens.synchronize()
    return (time.perf_counter() - t0) / reps


def autotune(domains, candidates=(1, 2, 4, 8), reps: int = 3, log=None, face_levels=(0, 2, 8, 16)):
    """(best stream count, member -> stream, table of (streams, s/step),
    table of (face_levels, s/step)); the members are left set to the best
    face_levels, so an Ensemble captured afterwards uses it."""
    times = member_times(domains, reps)
    table = []
    best = None
    for nb in candidates:
        nb = min(nb, len(domains))
        if any(t[0] == nb for t in table):
            continue
        assign = assign_longest_first(times, nb)
        ens = Ensemble(domains, assign, nb)
        t = graph_step_time(ens, reps)
        ens.close()
        table.append((nb, t))
        if log:
            log(f"{nb} stream(s): {1e3 * t:.2f} ms per ensemble step")
        if best is None or t < best[1]:
            best = (nb, t, assign)
    nb, assign = best[0], best[2]
    tiles = []
    for lg in face_levels:
        for d in domains:
            d.set_tile("face_levels", lg)
        ens = Ensemble(domains, assign, nb)
        t = graph_step_time(ens, reps)
        ens.close()
        tiles.append((lg, t))
        if log:
            log(f"{nb} stream(s), face_levels {lg}: {1e3 * t:.2f} ms per ensemble step")
    if tiles:
        lg = min(tiles, key=lambda e: e[1])[0]
        for d in domains:
            d.set_tile("face_levels", lg)
    return nb, assign, table, tiles


__all__ = ["Ensemble", "MemberSpec", "assign_longest_first", "autotune", "c2_members", "eager_step_time", "graph_step_time",
           "member_case", "member_times", "n_mass"]
