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
``nccopy -k 64-bit-offset`` makes it readable).  Errors use the reference's
exception types: a required variable that is missing raises ``KeyError``, a
wrfbdy record index out of range ``IndexError``, a non-finite initial state
``FloatingPointError``.  ``load_case`` opens each file once.
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
        var = self._nc.variables.get(name)
        if var is None:
            raise KeyError(f"{self.path.name} has no variable {name} (variables: {', '.join(sorted(self._nc.variables)[:8])}, ...)")
        return var

    def numeric(self):
        """(name, variable) of every numeric variable, in file order."""
        for name, var in self._nc.variables.items():
            if np.dtype(var.data.dtype).kind in "fiu":
                yield name, var

    def times(self) -> tuple:
        """The Times variable as text, one string per record (trailing NUL
        and space padding removed); () without one."""
        var = self._nc.variables.get("Times")
        if var is None:
            return ()
        a = np.asarray(var.data)
        if a.dtype.kind == "S" and a.ndim == 2:  # (Time, DateStrLen) single characters
            a = np.ascontiguousarray(a).view(f"S{a.shape[1]}").reshape(a.shape[0])
        return tuple((v.decode("ascii") if isinstance(v, bytes) else str(v)).rstrip("\x00 ") for v in a.reshape(-1).tolist())


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


def _state(ds: _Dataset) -> WrfState:
    u, v, t, q = (np.array(_first_record(ds.require(n))[0]) for n in STATE_VARIABLES)
    return WrfState(u=u, v=v, t=t, qvapor=q, time_index=0, time_seconds=0.0)


def _input_fields(ds: _Dataset) -> dict:
    return {name.lower(): _memory_order(*_first_record(var)) for name, var in ds.numeric()}


def load_wrfinput(path: str | Path, config: CaseConfig | None = None) -> tuple:
    """(grid, initial U/V/T/QVAPOR state) of ``wrfinput_d01``."""
    with _Dataset(path) as ds:
        return _grid(ds, config), _state(ds)


def load_wrfinput_fields(path: str | Path) -> dict:
    """Every numeric variable of ``wrfinput`` (record 0) in memory order, by
    lower-case name."""
    with _Dataset(path) as ds:
        return _input_fields(ds)


def load_wrfinput_attrs(path: str | Path) -> dict:
    """Global attributes of ``wrfinput`` (scalars unwrapped, text decoded)."""
    with _Dataset(path) as ds:
        return ds.attrs()


# ---------------------------------------------------------------------------
# wrfbdy
# ---------------------------------------------------------------------------
def _forcing(ds: _Dataset) -> BoundaryForcing:
    pairs = [(name, np.array(var.data)) for name, var in ds.numeric()]
    return BoundaryForcing(times=ds.times(), names=tuple(n for n, _ in pairs), arrays=tuple(a for _, a in pairs))


def _boundary_record(ds: _Dataset, record: int) -> dict:
    """Record ``record`` of every numeric variable in boundary order; the
    variables without a Time axis as they are."""
    nrec = ds.size("Time") if "Time" in ds.dims else 0
    if not 0 <= record < max(nrec, 1):
        raise IndexError(f"{ds.path.name} holds {nrec} boundary record(s); record {record} was requested")
    out = {}
    for name, var in ds.numeric():
        dims = tuple(var.dimensions)
        a = np.asarray(var.data)
        if dims[:1] == ("Time",):
            a, dims = a[record], dims[1:]
        out[name.lower()] = _boundary_order(a, dims)
    return out


def load_wrfbdy(path: str | Path) -> BoundaryForcing:
    """Every numeric wrfbdy variable with all its records, in file order."""
    with _Dataset(path) as ds:
        return _forcing(ds)


def load_wrfbdy_fields(path: str | Path, *, time_index: int = 0) -> dict:
    """One wrfbdy record in the reference's boundary order, by lower-case name."""
    with _Dataset(path) as ds:
        return _boundary_record(ds, int(time_index))


def wrfbdy_record_count(path: str | Path) -> int:
    with _Dataset(path) as ds:
        return ds.size("Time") if "Time" in ds.dims else 0


# ---------------------------------------------------------------------------
# run directory
# ---------------------------------------------------------------------------
def load_case(case_dir: str | Path) -> CaseData:
    """``namelist.input``, ``wrfinput_d01`` and ``wrfbdy_d01`` (record 0 merged
    into the fields) of a WRF run directory."""
    root = Path(case_dir)
    paths = {role: root / fname for role, fname in
             (("namelist", "namelist.input"), ("input", "wrfinput_d01"), ("boundary", "wrfbdy_d01"))}
    config = load_case_config(paths["namelist"])
    with _Dataset(paths["input"]) as ds:
        grid, state, fields, attrs = _grid(ds, config), _state(ds), _input_fields(ds), ds.attrs()
    with _Dataset(paths["boundary"]) as ds:
        forcing = _forcing(ds)
        fields.update(_boundary_record(ds, 0))
    return CaseData(config=config, grid=grid, state=state, forcing=forcing, fields=fields, attrs=attrs,
                    input_path=paths["input"], boundary_path=paths["boundary"], namelist_path=paths["namelist"])


def assert_state_finite(state: WrfState) -> None:
    """FloatingPointError naming each of U, V, T, QVAPOR that holds NaN or
    infinity, with the count of such values."""
    counts = {name: int(np.size(a) - np.count_nonzero(np.isfinite(a)))
              for name, a in zip(STATE_VARIABLES, (state.u, state.v, state.t, state.qvapor))}
    bad = [f"{name}: {n}" for name, n in counts.items() if n]
    if bad:
        raise FloatingPointError("non-finite values in the initial state (" + ", ".join(bad) + ")")


__all__ = [
    "BoundaryForcing", "CaseData", "STATE_VARIABLES", "WrfGrid", "WrfState", "assert_state_finite", "load_case",
    "load_wrfbdy", "load_wrfbdy_fields", "load_wrfinput", "load_wrfinput_attrs", "load_wrfinput_fields",
    "wrfbdy_record_count",
]
