"""Time loop over a WRF run directory: the em_real driver around the dycore.

Mirrors the outer loop of src/wrf_jax/trajectory.py (``run_compiled_direct_
trajectory`` :3933-4700 with its helpers ``_reset_dtbc_on_boundary_time``
:1494-1501, ``_refresh_wrfbdy_fields_on_boundary_time`` :1504-1526,
``_next_boundary_reset_step`` :3871-3882, ``_timestamp`` :985-986) and its
output side (``PrognosticSnapshot`` :919-928, ``FieldComparison`` :931-943,
``compare_snapshots_to_wrfout`` :4729-4772, ``write_snapshots_netcdf``
:4783-4888, ``_rmse_arrays`` / ``_nrmse_percent_arrays`` :1405-1420):

* the state stays on the GPU; the loop advances it in chunks of whole steps
  that end at the next boundary-record time or snapshot time;
* after a step that ends on a multiple of ``interval_seconds`` the clock's
  dtbc restarts at 0 and wrfbdy record elapsed // interval (when the file
  has it) replaces the boundary values and tendencies;
* at every multiple of the snapshot interval U, V, T and QVAPOR are read
  back in WRF NetCDF order (bottom_top, south_north, west_east[_stag]).

Physics is off in this path (the dycore only), so the reference's radiation
and physics cadence has no counterpart here.
"""

from __future__ import annotations

import time as _time
from dataclasses import dataclass, field
from datetime import datetime, timedelta
from pathlib import Path
from typing import Any, Callable, Mapping

import numpy as np

from .data_io import CaseData, load_case
from .real_case import boundary_record, dims_of, dycore_config, registry_arrays
from .step_em import TimeState

# units and descriptions written with each snapshot variable
FIELD_METADATA = {
    "U": {"units": "m s-1", "description": "x-wind component", "stagger": "X", "MemoryOrder": "XYZ"},
    "V": {"units": "m s-1", "description": "y-wind component", "stagger": "Y", "MemoryOrder": "XYZ"},
    "T": {"units": "K", "description": "perturbation potential temperature theta-t0", "stagger": "", "MemoryOrder": "XYZ"},
    "QVAPOR": {"units": "kg kg-1", "description": "Water vapor mixing ratio", "stagger": "", "MemoryOrder": "XYZ"},
}


@dataclass(frozen=True)
class PrognosticSnapshot:
    """U/V/T/QVAPOR at one output time, in WRF NetCDF dimension order."""

    elapsed_seconds: int
    timestamp: str
    u: Any
    v: Any
    t: Any
    qvapor: Any
    extra: Mapping[str, Any] | None = None


@dataclass(frozen=True)
class FieldComparison:
    """One field's error against a WRF history frame."""

    hour: int
    elapsed_seconds: int
    timestamp: str
    field: str
    rmse: float
    nrmse_percent: float
    max_abs_error: float
    units: str
    finite: bool


@dataclass
class TrajectoryResult:
    case: CaseData
    config: Any
    snapshots: tuple
    time: TimeState
    metrics: dict = field(default_factory=dict)
    final_arrays: dict | None = None


def timestamp(start: str, elapsed_seconds: int) -> str:
    t = datetime.strptime(start, "%Y-%m-%d_%H:%M:%S") + timedelta(seconds=int(elapsed_seconds))
    return t.strftime("%Y-%m-%d_%H:%M:%S")


def next_boundary_reset_step(step: int, step_count: int, dt_seconds: int, boundary_interval_seconds: int) -> int | None:
    """First step >= ``step`` whose end falls on a boundary-record time, or
    None (no interval, or beyond ``step_count``)."""
    if boundary_interval_seconds <= 0:
        return None
    every = int(boundary_interval_seconds) // int(dt_seconds)
    if every <= 0:
        return None
    nxt = -(-int(step) // every) * every
    if nxt < int(step):
        nxt += every
    return nxt if nxt <= int(step_count) else None


def _snapshot(dom, cfg, start: str, elapsed: int) -> PrognosticSnapshot:
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    got = dom.download(["u_2", "v_2", "t_2", "moist"])
    # registry [j][k][i] -> WRF (bottom_top, south_north, west_east[_stag])
    u = np.ascontiguousarray(got["u_2"][:ny, :nz, : nx + 1].transpose(1, 0, 2))
    v = np.ascontiguousarray(got["v_2"][: ny + 1, :nz, :nx].transpose(1, 0, 2))
    t = np.ascontiguousarray(got["t_2"][:ny, :nz, :nx].transpose(1, 0, 2))
    q = np.ascontiguousarray(got["moist"][0, :ny, :nz, :nx].transpose(1, 0, 2))
    return PrognosticSnapshot(elapsed_seconds=int(elapsed), timestamp=timestamp(start, elapsed), u=u, v=v, t=t, qvapor=q)


def run_trajectory(case_dir: str | Path | CaseData, *, run_seconds: int | None = None, snapshot_interval_seconds: int = 3600,
                   snapshot_callback: Callable[[PrognosticSnapshot], None] | None = None, keep_snapshots: bool = True,
                   device: str = "cuda:0", domain: Any = None, final_fields: tuple = (),
                   check_finite: bool = True) -> TrajectoryResult:
    """Run the case for ``run_seconds`` (default: the namelist's run length).

    ``domain`` replaces the HIP domain by another object with the same
    ``upload`` / ``step`` / ``download`` methods (tests drive the CPU oracle
    through it); by default the state lives in a ``Domain`` on ``device``.
    ``final_fields`` names registry arrays to read back at the end.  With
    ``check_finite`` the FP64 checksums are taken at every snapshot and a
    NaN/Inf raises FloatingPointError."""
    case = case_dir if isinstance(case_dir, CaseData) else load_case(case_dir)
    cfg = dycore_config(case)
    dims = dims_of(cfg)
    own = domain is None
    if own:
        from .device import Domain

        domain = Domain(cfg, device)
    dom = domain
    try:
        dom.upload(registry_arrays(case, cfg))
        dt = float(cfg.dt)
        total = int(run_seconds if run_seconds is not None else case.config.run_seconds)
        steps = int(round(total / dt))
        interval = int(case.config.boundary_interval_seconds)
        records = len(case.forcing.times)
        snap_every = int(snapshot_interval_seconds)
        snap_seconds = set(range(snap_every, steps * int(dt) + 1, snap_every)) if snap_every > 0 else set()
        snap_steps = sorted(int(s // dt) for s in snap_seconds)
        clock = TimeState(itimestep=np.int32(0), elapsed_seconds=np.float32(0.0), dtbc=np.float32(0.0))
        snapshots: list = []
        metrics = {"steps": 0, "boundary_refreshes": 0, "chunks": 0, "seconds": 0.0}
        t0 = _time.perf_counter()
        step = 0
        while step < steps:
            chunk = steps - step
            nb = next_boundary_reset_step(step + 1, steps, int(dt), interval)
            if nb is not None:
                chunk = min(chunk, nb - step)
            ns = next((s for s in snap_steps if s > step), None)
            if ns is not None:
                chunk = min(chunk, ns - step)
            dom.step(chunk, float(clock.dtbc))
            for _ in range(chunk):
                clock = clock.advance(dt, interval)
            step += chunk
            metrics["chunks"] += 1
            elapsed = int(round(step * dt))
            if interval > 0 and elapsed > 0 and elapsed % interval == 0:
                rec = elapsed // interval
                if 0 < rec < records:
                    dom.upload(boundary_record(case.boundary_path, rec, dims))
                    metrics["boundary_refreshes"] += 1
            if elapsed in snap_seconds:
                if check_finite and hasattr(dom, "checksums"):
                    dom.checksums()
                snap = _snapshot(dom, cfg, case.config.start_time, elapsed)
                if snapshot_callback is not None:
                    snapshot_callback(snap)
                if keep_snapshots:
                    snapshots.append(snap)
        if hasattr(dom, "synchronize"):
            dom.synchronize()
        metrics["steps"] = steps
        metrics["seconds"] = _time.perf_counter() - t0
        final = dom.download(list(final_fields)) if final_fields else None
        return TrajectoryResult(case=case, config=cfg, snapshots=tuple(snapshots), time=clock, metrics=metrics,
                                final_arrays=final)
    finally:
        if own:
            dom.close()


# ---------------------------------------------------------------------------
# output and comparison
# ---------------------------------------------------------------------------
def _rmse(actual: np.ndarray, expected: np.ndarray) -> float:
    d = actual - expected
    return float(np.sqrt(np.mean(d * d)))


def _nrmse_percent(actual: np.ndarray, expected: np.ndarray) -> float:
    r = _rmse(actual, expected)
    den = float(np.sqrt(np.mean(expected * expected)))
    if den == 0.0:
        return 0.0 if r == 0.0 else float("inf")
    return 100.0 * r / den


def _times_chars(stamps: list) -> np.ndarray:
    rows = []
    for s in stamps:
        b = s.encode("ascii")
        if len(b) != 19:
            raise ValueError(f"WRF timestamp must be 19 ASCII characters, got {s!r}")
        rows.append(list(b))
    return np.array(rows, dtype=np.uint8).view("S1")


def write_snapshots_netcdf(snapshots: tuple, path: str | Path) -> None:
    """WRF-style history file of the snapshots (NetCDF 64-bit offset): Times,
    XTIME (minutes), elapsed_seconds, U, V, T, QVAPOR."""
    from scipy.io import netcdf_file

    if not snapshots:
        raise ValueError("no snapshots to write")
    nz, ny, nx1 = np.asarray(snapshots[0].u).shape
    nx = nx1 - 1
    Path(path).parent.mkdir(parents=True, exist_ok=True)
    nc = netcdf_file(str(path), "w", version=2)
    try:
        nc.TITLE = "dycore trajectory snapshots"
        nc.createDimension("Time", None)
        for name, n in (("DateStrLen", 19), ("bottom_top", nz), ("south_north", ny), ("south_north_stag", ny + 1),
                        ("west_east", nx), ("west_east_stag", nx + 1)):
            nc.createDimension(name, n)
        v = nc.createVariable("Times", "c", ("Time", "DateStrLen"))
        v[:] = _times_chars([s.timestamp for s in snapshots])
        el = np.array([s.elapsed_seconds for s in snapshots], dtype=np.int32)
        v = nc.createVariable("XTIME", "f", ("Time",))
        v[:] = el.astype(np.float32) / np.float32(60.0)
        v.units = "minutes since simulation start"
        v = nc.createVariable("elapsed_seconds", "i", ("Time",))
        v[:] = el
        v.units = "s"
        for name, attr, dims in (("U", "u", ("Time", "bottom_top", "south_north", "west_east_stag")),
                                 ("V", "v", ("Time", "bottom_top", "south_north_stag", "west_east")),
                                 ("T", "t", ("Time", "bottom_top", "south_north", "west_east")),
                                 ("QVAPOR", "qvapor", ("Time", "bottom_top", "south_north", "west_east"))):
            v = nc.createVariable(name, "f", dims)
            v[:] = np.stack([np.asarray(getattr(s, attr), dtype=np.float32) for s in snapshots])
            for key, value in FIELD_METADATA[name].items():
                setattr(v, key, value)
    finally:
        nc.close()


def compare_snapshots_to_wrfout(snapshots: tuple, wrfout_path: str | Path) -> tuple:
    """RMSE / NRMSE (%) / max |error| of U, V, T, QVAPOR of every snapshot
    against the frame of a WRF history file with the same timestamp."""
    from scipy.io import netcdf_file

    nc = netcdf_file(str(wrfout_path), "r", mmap=False, maskandscale=False)
    try:
        raw = np.asarray(nc.variables["Times"].data)
        times = tuple(bytes(np.asarray(row).tobytes()).decode("ascii").rstrip("\x00") for row in raw)
        index = {t: n for n, t in enumerate(times)}
        out = []
        for snap in snapshots:
            if snap.timestamp not in index:
                raise KeyError(f"{snap.timestamp} not found in {wrfout_path}")
            n = index[snap.timestamp]
            for fld, actual in (("U", snap.u), ("V", snap.v), ("T", snap.t), ("QVAPOR", snap.qvapor)):
                if fld not in nc.variables:
                    continue
                var = nc.variables[fld]
                a = np.asarray(actual, dtype=np.float64)
                e = np.asarray(var.data[n], dtype=np.float64)
                units = getattr(var, "units", b"")
                units = units.decode("ascii", "replace") if isinstance(units, bytes) else str(units)
                out.append(FieldComparison(hour=int(snap.elapsed_seconds // 3600), elapsed_seconds=int(snap.elapsed_seconds),
                                           timestamp=snap.timestamp, field=fld, rmse=_rmse(a, e),
                                           nrmse_percent=_nrmse_percent(a, e), max_abs_error=float(np.max(np.abs(a - e))),
                                           units=units, finite=bool(np.all(np.isfinite(a)))))
        return tuple(out)
    finally:
        nc.close()


__all__ = ["FIELD_METADATA", "FieldComparison", "PrognosticSnapshot", "TrajectoryResult", "compare_snapshots_to_wrfout",
           "next_boundary_reset_step", "run_trajectory", "timestamp", "write_snapshots_netcdf"]
