"""Write a dycore case as a WRF run directory (NetCDF classic).

``export_case`` writes ``namelist.input``, ``wrfinput_d01`` and
``wrfbdy_d01`` with WRF's variable names, dimension names and staggering
(the layout src/wrf_jax/data_io.py:63-105 reads), so a seeded synthetic
workload can be run through the same file path as a real case, and files
from this package round-trip through ``data_io.load_case``.  Files are
NetCDF 64-bit-offset (``scipy.io.netcdf_file``, version 2).
"""

from __future__ import annotations

from pathlib import Path
from typing import Iterable, Mapping

import numpy as np

from .config import DycoreConfig
from .fields import Dims
from .real_case import FILE_NAMES, MOIST_SPECIES, SCALAR_SPECIES
from .step_em import from_registry

# file dimensions of each variable (Time first where the variable has it)
_3D = ("Time", "bottom_top", "south_north", "west_east")
_DIMS = {
    "U": ("Time", "bottom_top", "south_north", "west_east_stag"),
    "V": ("Time", "bottom_top", "south_north_stag", "west_east"),
    "W": ("Time", "bottom_top_stag", "south_north", "west_east"),
    "PH": ("Time", "bottom_top_stag", "south_north", "west_east"),
    "PHB": ("Time", "bottom_top_stag", "south_north", "west_east"),
    "T": _3D, "P": _3D, "PB": _3D, "AL": _3D, "ALB": _3D,
    "MU": ("Time", "south_north", "west_east"), "MUB": ("Time", "south_north", "west_east"),
    "HGT": ("Time", "south_north", "west_east"), "F": ("Time", "south_north", "west_east"),
    "E": ("Time", "south_north", "west_east"), "SINALPHA": ("Time", "south_north", "west_east"),
    "COSALPHA": ("Time", "south_north", "west_east"), "CLAT": ("Time", "south_north", "west_east"),
    "MAPFAC_MX": ("Time", "south_north", "west_east"), "MAPFAC_MY": ("Time", "south_north", "west_east"),
    "MAPFAC_UX": ("Time", "south_north", "west_east_stag"), "MAPFAC_UY": ("Time", "south_north", "west_east_stag"),
    "MAPFAC_VX": ("Time", "south_north_stag", "west_east"), "MAPFAC_VY": ("Time", "south_north_stag", "west_east"),
    "MF_VX_INV": ("Time", "south_north_stag", "west_east"),
    "ZNW": ("Time", "bottom_top_stag"), "C1F": ("Time", "bottom_top_stag"), "C2F": ("Time", "bottom_top_stag"),
    "C3F": ("Time", "bottom_top_stag"), "C4F": ("Time", "bottom_top_stag"),
}
for _n in ("ZNU", "DNW", "RDNW", "DN", "RDN", "FNM", "FNP", "C1H", "C2H", "C3H", "C4H", "U_BASE", "V_BASE"):
    _DIMS[_n] = ("Time", "bottom_top")
# registry name -> file variable written for it
_FILE_VAR = {
    "u_2": "U", "v_2": "V", "w_2": "W", "t_2": "T", "ph_2": "PH", "phb": "PHB", "p": "P", "pb": "PB", "al": "AL",
    "alb": "ALB", "mu_2": "MU", "mub": "MUB", "ht": "HGT", "msfux": "MAPFAC_UX", "msfuy": "MAPFAC_UY",
    "msfvx": "MAPFAC_VX", "msfvy": "MAPFAC_VY", "msftx": "MAPFAC_MX", "msfty": "MAPFAC_MY", "msfvx_inv": "MF_VX_INV",
    "f": "F", "e": "E", "sina": "SINALPHA", "cosa": "COSALPHA", "clat": "CLAT", "znw": "ZNW", "znu": "ZNU",
    "dnw": "DNW", "rdnw": "RDNW", "dn": "DN", "rdn": "RDN", "fnm": "FNM", "fnp": "FNP", "c1h": "C1H", "c2h": "C2H",
    "c1f": "C1F", "c2f": "C2F", "c3h": "C3H", "c4h": "C4H", "c3f": "C3F", "c4f": "C4F", "u_base": "U_BASE",
    "v_base": "V_BASE",
}
assert set(_FILE_VAR) == set(FILE_NAMES)
_BDY_VARS = (("u", "U", "bottom_top", "south_north", "west_east_stag"),
             ("v", "V", "bottom_top", "south_north_stag", "west_east"),
             ("t", "T", "bottom_top", "south_north", "west_east"),
             ("ph", "PH", "bottom_top_stag", "south_north", "west_east"),
             ("qv", "QVAPOR", "bottom_top", "south_north", "west_east"),
             ("mu", "MU", None, "south_north", "west_east"))


def _sizes(d: Dims) -> dict:
    return {"west_east": d.nx, "west_east_stag": d.nx + 1, "south_north": d.ny, "south_north_stag": d.ny + 1,
            "bottom_top": d.nz, "bottom_top_stag": d.nz + 1, "bdy_width": d.spec_bdy_width, "DateStrLen": 19}


def _file_order(ref: np.ndarray, dims: tuple, size: dict) -> np.ndarray:
    """Reference memory order (i, k, j) / (i, j) / (k,) cropped to the file
    extents and put in file order (k, j, i) / (j, i) / (k,)."""
    body = dims[1:]
    if len(body) == 3:
        k, j, i = (size[n] for n in body)
        return np.ascontiguousarray(ref[:i, :k, :j].transpose(1, 2, 0))
    if len(body) == 2:
        j, i = (size[n] for n in body)
        return np.ascontiguousarray(ref[:i, :j].transpose(1, 0))
    return np.ascontiguousarray(ref[: size[body[0]]])


def _times(stamps: Iterable[str]) -> np.ndarray:
    rows = []
    for s in stamps:
        b = s.encode("ascii")
        if len(b) != 19:
            raise ValueError(f"WRF timestamp must be 19 ASCII characters, got {s!r}")
        rows.append(list(b))
    return np.array(rows, dtype=np.uint8).view("S1")


def _open(path: Path, size: dict, names: Iterable[str]):
    from scipy.io import netcdf_file

    nc = netcdf_file(str(path), "w", version=2)
    nc.createDimension("Time", None)
    for n in names:
        nc.createDimension(n, size[n])
    return nc


def _stamp(start: str, seconds: float) -> str:
    from datetime import datetime, timedelta

    t = datetime.strptime(start, "%Y-%m-%d_%H:%M:%S") + timedelta(seconds=float(seconds))
    return t.strftime("%Y-%m-%d_%H:%M:%S")


def write_namelist(path: str | Path, cfg: DycoreConfig, *, run_seconds: int, interval_seconds: int,
                   history_interval_seconds: int, start: str) -> None:
    end = _stamp(start, run_seconds)
    y, mo, d = start[:10].split("-")
    h, mi, s = start[11:].split(":")
    ey, emo, ed = end[:10].split("-")
    eh, emi, es = end[11:].split(":")
    text = f"""&time_control
 run_days = 0,
 run_hours = 0,
 run_minutes = 0,
 run_seconds = {int(run_seconds)},
 start_year = {y},
 start_month = {mo},
 start_day = {d},
 start_hour = {h},
 start_minute = {mi},
 start_second = {s},
 end_year = {ey},
 end_month = {emo},
 end_day = {ed},
 end_hour = {eh},
 end_minute = {emi},
 end_second = {es},
 interval_seconds = {int(interval_seconds)},
 history_interval_s = {int(history_interval_seconds)},
/
&domains
 time_step = {int(cfg.dt)},
 max_dom = 1,
 e_we = {cfg.nx + 1},
 e_sn = {cfg.ny + 1},
 e_vert = {cfg.nz + 1},
 dx = {cfg.dx!r},
 dy = {cfg.dy!r},
/
&dynamics
 rk_ord = 3,
 diff_opt = 2,
 km_opt = 4,
 time_step_sound = {cfg.time_step_sound},
 moist_adv_opt = {cfg.pd_moist},
 scalar_adv_opt = {cfg.pd_scalar},
 epssm = {cfg.epssm!r},
 smdiv = {cfg.smdiv!r},
 emdiv = {cfg.emdiv!r},
 c_s = {cfg.c_s!r},
 non_hydrostatic = .true.,
 hypsometric_opt = 2,
/
&bdy_control
 spec_bdy_width = {cfg.spec_bdy_width},
 spec_zone = {cfg.spec_zone},
 relax_zone = {cfg.relax_zone},
 specified = .true.,
/
"""
    Path(path).write_text(text, encoding="ascii")


def write_wrfinput(path: str | Path, cfg: DycoreConfig, arrays: Mapping[str, np.ndarray], *, start: str) -> None:
    """``wrfinput_d01`` of registry arrays (record 0 at ``start``)."""
    d = Dims(nx=cfg.nx, ny=cfg.ny, nz=cfg.nz, nmoist=cfg.nmoist, nscalar=cfg.nscalar, spec_bdy_width=cfg.spec_bdy_width)
    size = _sizes(d)
    nc = _open(Path(path), size, ("DateStrLen", "west_east", "west_east_stag", "south_north", "south_north_stag",
                                  "bottom_top", "bottom_top_stag"))
    try:
        nc.TITLE = "dycore case written by wrfdc.case_files"
        nc.DX, nc.DY = np.float32(cfg.dx), np.float32(cfg.dy)
        nc.MAP_PROJ = np.int32(0)
        v = nc.createVariable("Times", "c", ("Time", "DateStrLen"))
        v[:] = _times([start])
        for reg, var in _FILE_VAR.items():
            if reg not in arrays:
                continue
            dims = _DIMS[var]
            data = _file_order(from_registry(reg, arrays[reg], d), dims, size)
            nv = nc.createVariable(var, "f", dims)
            nv[:] = data[None]
        for stack, species, n in (("moist", MOIST_SPECIES, cfg.nmoist), ("scalar", SCALAR_SPECIES, cfg.nscalar)):
            if stack not in arrays or n == 0:
                continue
            ref = from_registry(stack, arrays[stack], d)
            for s in range(n):
                nv = nc.createVariable(species[s][-1].upper(), "f", _3D)
                nv[:] = _file_order(ref[..., s + 1], _3D, size)[None]
        for name, value in (("P_TOP", cfg.ptop), ("CF1", cfg.cf1), ("CF2", cfg.cf2), ("CF3", cfg.cf3), ("CFN", cfg.cfn),
                            ("CFN1", cfg.cfn1)):
            nv = nc.createVariable(name, "f", ("Time",))
            nv[:] = np.array([value], dtype=np.float32)
    finally:
        nc.close()


def write_wrfbdy(path: str | Path, cfg: DycoreConfig, records: list, *, start: str, interval_seconds: int) -> None:
    """``wrfbdy_d01`` with one record per entry of ``records`` (each a mapping
    of registry boundary arrays ``u_bxs`` ... ``mu_btye``; missing arrays are
    written as zeros), record n at ``start`` + n * ``interval_seconds``."""
    d = Dims(nx=cfg.nx, ny=cfg.ny, nz=cfg.nz, nmoist=cfg.nmoist, nscalar=cfg.nscalar, spec_bdy_width=cfg.spec_bdy_width)
    size = _sizes(d)
    bw = d.spec_bdy_width
    nc = _open(Path(path), size, ("DateStrLen", "west_east", "west_east_stag", "south_north", "south_north_stag",
                                  "bottom_top", "bottom_top_stag", "bdy_width"))
    try:
        nc.TITLE = "dycore boundary records written by wrfdc.case_files"
        v = nc.createVariable("Times", "c", ("Time", "DateStrLen"))
        v[:] = _times([_stamp(start, n * interval_seconds) for n in range(len(records))])
        for var, file_var, kdim, jdim, idim in _BDY_VARS:
            for part, tag in (("b", "B"), ("bt", "BT")):
                for edge in ("xs", "xe", "ys", "ye"):
                    reg = f"{var}_{part}{edge}"
                    edim = jdim if edge[0] == "x" else idim
                    dims = ("Time", "bdy_width") + ((kdim,) if kdim else ()) + (edim,)
                    nv = nc.createVariable(f"{file_var}_{tag}{edge.upper()}", "f", dims)
                    data = np.zeros((len(records),) + tuple(size[n] for n in dims[1:]), dtype=np.float32)
                    for r, rec in enumerate(records):
                        if reg not in rec:
                            continue
                        a = np.asarray(rec[reg], dtype=np.float32)  # registry (bw, kde, edge) or (bw, edge)
                        if kdim:
                            data[r] = a[:bw, : size[kdim], : size[edim]]
                        else:
                            data[r] = a[:bw, : size[edim]]
                    nv[:] = data
    finally:
        nc.close()


def export_case(directory: str | Path, cfg: DycoreConfig, arrays: Mapping[str, np.ndarray], *,
                records: list | None = None, interval_seconds: int | None = None, run_seconds: int | None = None,
                history_interval_seconds: int | None = None, start: str = "2000-01-01_00:00:00") -> Path:
    """Write ``namelist.input``, ``wrfinput_d01`` and ``wrfbdy_d01`` for the
    registry arrays of one domain.  ``records`` defaults to one record taken
    from the arrays' own boundary values and tendencies."""
    root = Path(directory)
    root.mkdir(parents=True, exist_ok=True)
    if records is None:
        records = [{k: v for k, v in arrays.items() if "_b" in k and k.split("_b")[0] in ("u", "v", "t", "ph", "qv", "mu")}]
    dt = int(cfg.dt)
    interval = int(interval_seconds if interval_seconds is not None else 10 * dt)
    run = int(run_seconds if run_seconds is not None else interval * len(records))
    write_namelist(root / "namelist.input", cfg, run_seconds=run, interval_seconds=interval,
                   history_interval_seconds=int(history_interval_seconds or run), start=start)
    write_wrfinput(root / "wrfinput_d01", cfg, arrays, start=start)
    write_wrfbdy(root / "wrfbdy_d01", cfg, records, start=start, interval_seconds=interval)
    return root


__all__ = ["export_case", "write_namelist", "write_wrfbdy", "write_wrfinput"]
