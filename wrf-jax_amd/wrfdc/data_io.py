"""WRF NetCDF case files: ``wrfinput_d01``, ``wrfbdy_d01``, ``namelist.input``.

Mirrors src/wrf_jax/data_io.py:1-241 (``load_wrfinput``,
``load_wrfinput_fields``, ``load_wrfinput_attrs``, ``load_wrfbdy``,
``load_wrfbdy_fields``, ``load_case``, ``assert_state_finite``): variables come
back under lower-case names with the leading Time axis dropped (record 0 of
wrfinput, the requested record of wrfbdy) and transposed into the
reference's memory order:

* 3-D (bottom_top[_stag], south_north[_stag], west_east[_stag]) -> (i, k, j)
  and soil (soil_layers[_stag], south_north, west_east) -> (i, k, j);
* 2-D (south_north[_stag], west_east[_stag]) -> (i, j);
* boundary 3-D (bdy_width, bottom_top[_stag], edge) -> (edge, k, bdy_width),
  boundary 2-D (bdy_width, edge) -> (edge, 1, bdy_width);
* anything else (1-D columns, scalars) unchanged.

The files are read with ``scipy.io.netcdf_file``, which reads the NetCDF
classic and 64-bit-offset formats (what WRF writes with ``io_form 2`` and
classic output).  A NetCDF-4/HDF5 file needs a reader this environment does
not have and is rejected with ``NotImplementedError`` (converting it with
``nccopy -k 64-bit-offset`` makes it readable).  As in the reference, a
required variable that is missing raises ``KeyError`` and a wrfbdy record
index out of range raises ``IndexError``.
"""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path
from typing import Any

import numpy as np

from .case_config import CaseConfig, load_case_config

STATE_VARIABLES = ("U", "V", "T", "QVAPOR")

_HDF5_MAGIC = b"\x89HDF\r\n\x1a\n"


@dataclass(frozen=True)
class WrfGrid:
    """Grid extents and spacing of one domain (src/wrf_jax/grid.py)."""

    west_east: int
    south_north: int
    bottom_top: int
    west_east_stag: int
    south_north_stag: int
    bottom_top_stag: int
    dx: float
    dy: float
    map_proj: int | None = None


@dataclass(frozen=True)
class WrfState:
    """U, V, T, QVAPOR of the initial state (file order, record 0)."""

    u: Any
    v: Any
    t: Any
    qvapor: Any
    time_index: int = 0
    time_seconds: float = 0.0


@dataclass(frozen=True)
class BoundaryForcing:
    """Every numeric wrfbdy variable with all its records (src/wrf_jax/forcing.py)."""

    times: tuple
    names: tuple
    arrays: tuple

    def record_count(self) -> int:
        return len(self.times) if self.times else (int(self.arrays[0].shape[0]) if self.arrays else 0)


@dataclass(frozen=True)
class CaseData:
    """Loaded runtime inputs of a single-domain case."""

    config: CaseConfig
    grid: WrfGrid
    state: WrfState
    forcing: BoundaryForcing
    fields: dict
    attrs: dict
    input_path: Path
    boundary_path: Path
    namelist_path: Path


# ---------------------------------------------------------------------------
# file access
# ---------------------------------------------------------------------------
class _Dataset:
    """Read-only view of one NetCDF classic file: variables with their
    dimension names, global attributes and dimension sizes."""

    def __init__(self, path: str | Path):
        from scipy.io import netcdf_file

        self.path = Path(path)
        with open(self.path, "rb") as fh:
            head = fh.read(8)
        if head.startswith(_HDF5_MAGIC):
            raise NotImplementedError(f"{self.path} is a NetCDF-4/HDF5 file; only the NetCDF classic and 64-bit-offset formats "
                                      "can be read here (nccopy -k 64-bit-offset converts it)")
        if not head.startswith(b"CDF"):
            raise ValueError(f"{self.path} is not a NetCDF file")
        self._nc = netcdf_file(str(self.path), "r", mmap=False, maskandscale=False)

    def __enter__(self) -> "_Dataset":
        return self

    def __exit__(self, *exc) -> None:
        self._nc.close()

    @property
    def variables(self) -> dict:
        return self._nc.variables

    @property
    def dims(self) -> dict:
        return self._nc.dimensions

    def attrs(self) -> dict:
        out = {}
        # scipy keeps the global attributes in this dict (also exposed as
        # attributes of the file object)
        for name, value in self._nc._attributes.items():
            if isinstance(value, bytes):
                value = value.decode("ascii", "replace")
            elif isinstance(value, np.ndarray) and value.size == 1:
                value = value.reshape(-1)[0].item()
            out[str(name)] = value
        return out

    def size(self, dim: str) -> int:
        n = self._nc.dimensions[dim]
        if n is None:  # the record (Time) dimension: the record count of its variables
            recs = [v.shape[0] for v in self._nc.variables.values() if v.dimensions and v.dimensions[0] == dim]
            return int(max(recs)) if recs else 0
        return int(n)

    def require(self, name: str):
        if name not in self._nc.variables:
            raise KeyError(f"required WRF variable {name!r} is missing from {self.path}")
        return self._nc.variables[name]


def _decode_times(values: Any) -> tuple:
    a = np.asarray(values)
    if a.dtype.kind == "S" and a.ndim == 2:
        return tuple(b"".join(row).decode("ascii") for row in a)
    if a.dtype.kind == "S":
        return tuple(v.decode("ascii") for v in a.reshape(-1).tolist())
    return tuple(str(v) for v in a.reshape(-1).tolist())


def _numeric(var) -> bool:
    return np.dtype(var.data.dtype).kind in "fiu"


def _memory_order(values: np.ndarray, dims: tuple) -> np.ndarray:
    """File dimension order (Time already dropped) -> the reference's memory order."""
    if len(dims) == 3 and dims[1].startswith("south_north") and dims[2].startswith("west_east") and \
            (dims[0].startswith("bottom_top") or dims[0].startswith("soil_layers")):
        return np.ascontiguousarray(np.transpose(values, (2, 0, 1)))
    if len(dims) == 2 and dims[0].startswith("south_north") and dims[1].startswith("west_east"):
        return np.ascontiguousarray(np.transpose(values, (1, 0)))
    return np.array(values)


def _boundary_order(values: np.ndarray, dims: tuple) -> np.ndarray:
    """wrfbdy (bdy_width, level, edge) -> (edge, level, bdy_width); (bdy_width,
    edge) -> (edge, 1, bdy_width)."""
    if len(dims) == 3 and dims[0] == "bdy_width" and dims[1].startswith("bottom_top"):
        return np.ascontiguousarray(np.transpose(values, (2, 1, 0)))
    if len(dims) == 2 and dims[0] == "bdy_width":
        return np.ascontiguousarray(np.transpose(values, (1, 0))[:, None, :])
    return np.array(values)


def _first_record(var) -> tuple:
    dims = tuple(var.dimensions)
    values = np.asarray(var.data)
    if dims and dims[0] == "Time":
        return values[0], dims[1:]
    return values, dims


# ---------------------------------------------------------------------------
# wrfinput
# ---------------------------------------------------------------------------
def _grid(ds: _Dataset, config: CaseConfig | None) -> WrfGrid:
    a = ds.attrs()
    return WrfGrid(
        west_east=ds.size("west_east"), south_north=ds.size("south_north"), bottom_top=ds.size("bottom_top"),
        west_east_stag=ds.size("west_east_stag"), south_north_stag=ds.size("south_north_stag"),
        bottom_top_stag=ds.size("bottom_top_stag"),
        dx=float(a.get("DX", config.dx if config is not None else 0.0)),
        dy=float(a.get("DY", config.dy if config is not None else 0.0)),
        map_proj=int(a["MAP_PROJ"]) if "MAP_PROJ" in a else None,
    )


def load_wrfinput(path: str | Path, config: CaseConfig | None = None) -> tuple:
    """(grid, initial U/V/T/QVAPOR state) of ``wrfinput_d01``."""
    with _Dataset(path) as ds:
        grid = _grid(ds, config)
        vals = [np.array(_first_record(ds.require(n))[0]) for n in STATE_VARIABLES]
    return grid, WrfState(u=vals[0], v=vals[1], t=vals[2], qvapor=vals[3], time_index=0, time_seconds=0.0)


def load_wrfinput_fields(path: str | Path) -> dict:
    """Every numeric variable of ``wrfinput`` (record 0) in memory order, by
    lower-case name."""
    out = {}
    with _Dataset(path) as ds:
        for name, var in ds.variables.items():
            if name == "Times" or not _numeric(var):
                continue
            values, dims = _first_record(var)
            out[name.lower()] = _memory_order(values, dims)
    return out


def load_wrfinput_attrs(path: str | Path) -> dict:
    """Global attributes of ``wrfinput`` (scalars unwrapped, text decoded)."""
    with _Dataset(path) as ds:
        return ds.attrs()


# ---------------------------------------------------------------------------
# wrfbdy
# ---------------------------------------------------------------------------
def load_wrfbdy(path: str | Path) -> BoundaryForcing:
    """Every numeric wrfbdy variable with all its records, in file order."""
    with _Dataset(path) as ds:
        times = _decode_times(ds.variables["Times"].data) if "Times" in ds.variables else ()
        names, arrays = [], []
        for name, var in ds.variables.items():
            if name == "Times" or not _numeric(var):
                continue
            names.append(name)
            arrays.append(np.array(var.data))
    return BoundaryForcing(times=times, names=tuple(names), arrays=tuple(arrays))


def load_wrfbdy_fields(path: str | Path, *, time_index: int = 0) -> dict:
    """One wrfbdy record in the reference's boundary order, by lower-case name."""
    out = {}
    with _Dataset(path) as ds:
        for name, var in ds.variables.items():
            if name == "Times" or not _numeric(var):
                continue
            dims = tuple(var.dimensions)
            values = np.asarray(var.data)
            if dims and dims[0] == "Time":
                count = int(values.shape[0])
                idx = int(time_index)
                if idx < 0 or idx >= count:
                    raise IndexError(f"wrfbdy time_index {idx} out of range for {name}: 0..{count - 1}")
                values, dims = values[idx], dims[1:]
            out[name.lower()] = _boundary_order(values, dims)
    return out


def wrfbdy_record_count(path: str | Path) -> int:
    with _Dataset(path) as ds:
        return ds.size("Time") if "Time" in ds.dims else 0


# ---------------------------------------------------------------------------
# run directory
# ---------------------------------------------------------------------------
def load_case(case_dir: str | Path) -> CaseData:
    """``namelist.input``, ``wrfinput_d01`` and ``wrfbdy_d01`` (record 0) of a
    WRF run directory."""
    root = Path(case_dir)
    namelist_path, input_path, boundary_path = root / "namelist.input", root / "wrfinput_d01", root / "wrfbdy_d01"
    config = load_case_config(namelist_path)
    grid, state = load_wrfinput(input_path, config)
    fields = load_wrfinput_fields(input_path)
    attrs = load_wrfinput_attrs(input_path)
    fields.update(load_wrfbdy_fields(boundary_path))
    forcing = load_wrfbdy(boundary_path)
    return CaseData(config=config, grid=grid, state=state, forcing=forcing, fields=fields, attrs=attrs,
                    input_path=input_path, boundary_path=boundary_path, namelist_path=namelist_path)


def assert_state_finite(state: WrfState) -> None:
    """Raise FloatingPointError if U, V, T or QVAPOR holds NaN or infinity."""
    for name, values in (("U", state.u), ("V", state.v), ("T", state.t), ("QVAPOR", state.qvapor)):
        if not bool(np.all(np.isfinite(np.asarray(values)))):
            raise FloatingPointError(f"{name} contains NaN or infinity")


__all__ = [
    "BoundaryForcing", "CaseData", "STATE_VARIABLES", "WrfGrid", "WrfState", "assert_state_finite", "load_case",
    "load_wrfbdy", "load_wrfbdy_fields", "load_wrfinput", "load_wrfinput_attrs", "load_wrfinput_fields",
    "wrfbdy_record_count",
]
