"""A WRF run directory -> the dycore's configuration and device state.

The reference assembles its runtime state from the loaded case in
src/wrf_jax/runtime.py:4312-4431 (``_fields``: map-factor and trigonometry
aliases, ``_1``/``_2`` copies of the prognostic fields, zero work arrays, the
CONUS moist and scalar stacks of runtime.py:4128-4139) and its configuration
in runtime.py:3641-3870 (``_runtime_grid``, ``_config_flags``: global
attributes of wrfinput first, the namelist second).  This module does the
same for the arrays and options the HIP dycore reads:

* ``runtime_fields(case)`` -> reference-layout template (3-D (i, k, j) and
  2-D (i, j) padded to the staggered extents, moisture (i, k, j, n) with slot
  0 unused, boundary arrays (edge, k, bdy_width)), the ``fields_template`` of
  ``build_step_em_core_context``;
* ``dycore_config(case)`` -> ``DycoreConfig`` (options the path does not
  implement are rejected by ``DycoreConfig.validate``);
* ``boundary_record(path, index, dims)`` -> one wrfbdy record as registry
  arrays, for the boundary refresh of a time loop.

Start-up work that WRF's start_em does beyond copying (diagnosing AL or the
surface w from the input) is not repeated: AL, ALB, P and PB are required in
the input file.
"""

from __future__ import annotations

from pathlib import Path
from typing import Any

import numpy as np

from .config import DycoreConfig
from .data_io import CaseData, load_wrfbdy_fields
from .fields import FIELD_KIND, Dims
from .step_em import registry_name, to_registry

# moist / scalar species in stack order (src/wrf_jax/runtime.py:4128-4139);
# each entry lists the accepted file names
MOIST_SPECIES = (("qvapor",), ("qcloud",), ("qrain",), ("qice",), ("qsnow",), ("qgraup",))
SCALAR_SPECIES = (("qni", "qnice"), ("qnr", "qnrain"))

# registry field <- lower-case file variable(s), first present wins
FILE_NAMES = {
    "u_2": ("u",), "v_2": ("v",), "w_2": ("w",), "t_2": ("t",), "ph_2": ("ph",), "phb": ("phb",),
    "p": ("p",), "pb": ("pb",), "al": ("al",), "alb": ("alb",),
    "mu_2": ("mu",), "mub": ("mub",), "ht": ("ht", "hgt"),
    "msfux": ("msfux", "mapfac_ux"), "msfuy": ("msfuy", "mapfac_uy"), "msfvx": ("msfvx", "mapfac_vx"),
    "msfvy": ("msfvy", "mapfac_vy"), "msftx": ("msftx", "mapfac_mx"), "msfty": ("msfty", "mapfac_my"),
    "msfvx_inv": ("msfvx_inv", "mf_vx_inv"), "f": ("f",), "e": ("e",), "sina": ("sina", "sinalpha"),
    "cosa": ("cosa", "cosalpha"), "clat": ("clat",),
    "znw": ("znw",), "znu": ("znu",), "dnw": ("dnw",), "rdnw": ("rdnw",), "dn": ("dn",), "rdn": ("rdn",),
    "fnm": ("fnm",), "fnp": ("fnp",), "c1h": ("c1h",), "c2h": ("c2h",), "c1f": ("c1f",), "c2f": ("c2f",),
    "c3h": ("c3h",), "c4h": ("c4h",), "c3f": ("c3f",), "c4f": ("c4f",), "u_base": ("u_base",), "v_base": ("v_base",),
}
REQUIRED = ("u_2", "v_2", "w_2", "t_2", "ph_2", "phb", "p", "pb", "al", "alb", "mu_2", "mub", "znw", "znu", "dnw", "rdnw",
            "dn", "rdn", "fnm", "fnp", "c1h", "c2h", "c1f", "c2f", "c3h", "c4h", "c3f", "c4f")
# the reference's _1 / _2 copies at start-up (runtime.py:4356-4372)
COPIES = {"u_1": "u_2", "v_1": "v_2", "w_1": "w_2", "t_1": "t_2", "ph_1": "ph_2", "mu_1": "mu_2"}
# boundary arrays a wrfbdy record provides (registry names)
BOUNDARY_VARS = ("u", "v", "t", "ph", "qv", "mu")


def _pick(fields: dict, names: tuple) -> Any:
    for n in names:
        if n in fields:
            return fields[n]
    return None


def _pad(a: np.ndarray, shape: tuple, horizontal_edge: bool = False) -> np.ndarray:
    """Pad ``a`` at the high end of every axis to ``shape``: with zeros, or
    (``horizontal_edge``, for (i, k, j) / (i, j) fields) by repeating the
    last i column and j row, as the reference expands input-shaped arrays to
    its memory extents (runtime.py:3907-3935, 3946-3986; the level axis is
    zero-filled here where the reference keeps the shorter extent)."""
    a = np.asarray(a, dtype=np.float32)
    if a.ndim != len(shape) or any(s > t for s, t in zip(a.shape, shape)):
        raise ValueError(f"array of shape {a.shape} does not fit the staggered extents {shape}")
    if horizontal_edge and a.ndim in (2, 3):
        widths = [(0, t - s) for s, t in zip(a.shape, shape)]
        if a.ndim == 3:
            kpad = widths[1]
            widths[1] = (0, 0)
            a = np.pad(a, widths, mode="edge")
            return np.pad(a, [(0, 0), kpad, (0, 0)])
        return np.pad(a, widths, mode="edge")
    out = np.zeros(shape, dtype=np.float32)
    out[tuple(slice(0, s) for s in a.shape)] = a
    return out


def _ref_shape(kind: str, dims: Dims) -> tuple:
    ide, jde, kde, bw = dims.ide, dims.jde, dims.kde, dims.spec_bdy_width
    return {"3": (ide, kde, jde), "2": (ide, jde), "1": (kde,), "X": (jde, kde, bw), "Y": (ide, kde, bw),
            "x": (jde, 1, bw), "y": (ide, 1, bw)}[kind]


def _scalar(fields: dict, attrs: dict, name: str, default: float | None = None) -> float:
    if name in fields:
        return float(np.asarray(fields[name]).reshape(-1)[0])
    if name.upper() in attrs:
        return float(attrs[name.upper()])
    if default is None:
        raise KeyError(f"required WRF variable {name.upper()!r} is missing from the input")
    return float(default)


def _option(case: CaseData, attr: str, group: str, key: str, default: Any) -> Any:
    """Global attribute of wrfinput, else the namelist (domain 1), else the default."""
    if attr in case.attrs:
        return case.attrs[attr]
    v = case.config.namelist.get(group, {}).get(key, default)
    return v[0] if isinstance(v, list) else v


def species_counts(case: CaseData) -> tuple:
    """(moist species, scalar species) held for this case: the six CONUS moist
    species always, the two number concentrations when the input has them."""
    nscalar = 2 if any(_pick(case.fields, names) is not None for names in SCALAR_SPECIES) else 0
    return len(MOIST_SPECIES), nscalar


def dycore_config(case: CaseData) -> DycoreConfig:
    """Configuration of the case: extents and spacing from the namelist and
    the file, vertical constants from the file, options from the global
    attributes and the namelist (runtime.py:3641-3870)."""
    cfg = case.config
    f, a = case.fields, case.attrs
    nx, ny, nz = int(cfg.e_we) - 1, int(cfg.e_sn) - 1, int(cfg.e_vert) - 1
    nmoist, nscalar = species_counts(case)
    spec_bdy_width = int(_option(case, "SPEC_BDY_WIDTH", "bdy_control", "spec_bdy_width", 5))
    spec_zone = int(_option(case, "SPEC_ZONE", "bdy_control", "spec_zone", 1))
    relax_zone = int(_option(case, "RELAX_ZONE", "bdy_control", "relax_zone", spec_bdy_width - spec_zone))
    moist_adv = int(_option(case, "MOIST_ADV_OPT", "dynamics", "moist_adv_opt", 1))
    scalar_adv = int(_option(case, "SCALAR_ADV_OPT", "dynamics", "scalar_adv_opt", 1))
    for name, v in (("moist_adv_opt", moist_adv), ("scalar_adv_opt", scalar_adv)):
        if v not in (0, 1):
            raise NotImplementedError(f"dycore option {name}={v!r} is not implemented (only 0 or 1)")
    options = {
        "rk_ord": int(_option(case, "RK_ORD", "dynamics", "rk_ord", 3)),
        "diff_opt": int(_option(case, "DIFF_OPT", "dynamics", "diff_opt", 2)),
        "km_opt": int(_option(case, "KM_OPT", "dynamics", "km_opt", 4)),
        "damp_opt": int(_option(case, "DAMP_OPT", "dynamics", "damp_opt", 0)),
        "w_damping": int(_option(case, "W_DAMPING", "dynamics", "w_damping", 0)),
        "hypsometric_opt": int(_option(case, "HYPSOMETRIC_OPT", "dynamics", "hypsometric_opt", 2)),
        "use_theta_m": int(_option(case, "USE_THETA_M", "dynamics", "use_theta_m", 0)),
        "non_hydrostatic": bool(_option(case, "NON_HYDROSTATIC", "dynamics", "non_hydrostatic", True)),
        "specified": bool(_option(case, "SPECIFIED", "bdy_control", "specified", True)),
        "zadvect_implicit": int(_option(case, "ZADVECT_IMPLICIT", "dynamics", "zadvect_implicit", 0)),
        "h_mom_adv_order": int(_option(case, "H_MOM_ADV_ORDER", "dynamics", "h_mom_adv_order", 5)),
        "v_mom_adv_order": int(_option(case, "V_MOM_ADV_ORDER", "dynamics", "v_mom_adv_order", 3)),
        "h_sca_adv_order": int(_option(case, "H_SCA_ADV_ORDER", "dynamics", "h_sca_adv_order", 5)),
        "v_sca_adv_order": int(_option(case, "V_SCA_ADV_ORDER", "dynamics", "v_sca_adv_order", 3)),
        "top_lid": bool(_option(case, "TOP_LID", "dynamics", "top_lid", False)),
        # relaxation-zone sponge exponent (runtime.py:3742; lbc_fcx_gcx scales
        # the weights by exp(-(b - spec_zone - 1) * spec_exp)) and lateral
        # boundary data for species beyond QV (runtime.py:3855-3856): the
        # path implements spec_exp 0 and QV-only boundary data
        "spec_exp": float(_option(case, "SPEC_EXP", "bdy_control", "spec_exp", 0.0)),
        "have_bcs_moist": bool(_option(case, "HAVE_BCS_MOIST", "bdy_control", "have_bcs_moist", False)),
        "have_bcs_scalar": bool(_option(case, "HAVE_BCS_SCALAR", "bdy_control", "have_bcs_scalar", False)),
    }
    out = DycoreConfig(
        nx=nx, ny=ny, nz=nz, ip=Dims(nx=nx, ny=ny, nz=nz, nmoist=nmoist, nscalar=nscalar).ip, nmoist=nmoist, nscalar=nscalar,
        spec_bdy_width=spec_bdy_width, spec_zone=spec_zone, relax_zone=relax_zone,
        time_step_sound=int(_option(case, "TIME_STEP_SOUND", "dynamics", "time_step_sound", 4)),
        pd_moist=moist_adv, pd_scalar=scalar_adv, top_lid=0,
        dx=float(a.get("DX", cfg.dx)), dy=float(a.get("DY", cfg.dy)), dt=float(cfg.time_step),
        ptop=_scalar(f, a, "p_top"),
        epssm=float(_option(case, "EPSSM", "dynamics", "epssm", 0.1)),
        smdiv=float(_option(case, "SMDIV", "dynamics", "smdiv", 0.1)),
        emdiv=float(_option(case, "EMDIV", "dynamics", "emdiv", 0.01)),
        c_s=float(_option(case, "C_S", "dynamics", "c_s", 0.25)),
        cf1=_scalar(f, a, "cf1"), cf2=_scalar(f, a, "cf2"), cf3=_scalar(f, a, "cf3"),
        cfn=_scalar(f, a, "cfn"), cfn1=_scalar(f, a, "cfn1"),
        options=options,
    )
    return out.validate()


def runtime_fields(case: CaseData, dims: Dims) -> dict:
    """Reference-layout template of every registry field the input defines
    (the reference's names: ``qvapor_b*`` for the QV boundary arrays, the
    moisture and scalar stacks with slot 0 unused)."""
    f = case.fields
    out: dict = {}
    for reg, names in FILE_NAMES.items():
        v = _pick(f, names)
        if v is None:
            if reg in REQUIRED:
                raise KeyError(f"required WRF variable {names[-1].upper()!r} is missing from {case.input_path}")
            continue
        out[reg] = _pad(v, _ref_shape(FIELD_KIND[reg], dims), horizontal_edge=True)
    if "msfvx_inv" not in out and "msfvx" in out:
        out["msfvx_inv"] = np.float32(1.0) / out["msfvx"]  # runtime.py:4349-4350
    for dst, src in COPIES.items():
        out[dst] = out[src].copy()
    shape3 = _ref_shape("3", dims)
    moist = np.zeros(shape3 + (dims.nmoist + 1,), dtype=np.float32)
    for s, names in enumerate(MOIST_SPECIES[: dims.nmoist]):
        v = _pick(f, names)
        if v is not None:
            moist[..., s + 1] = _pad(v, shape3, horizontal_edge=True)
    out["moist"] = moist
    if dims.nscalar:
        scalar = np.zeros(shape3 + (dims.nscalar + 1,), dtype=np.float32)
        for s, names in enumerate(SCALAR_SPECIES[: dims.nscalar]):
            v = _pick(f, names)
            if v is not None:
                scalar[..., s + 1] = _pad(v, shape3, horizontal_edge=True)
        out["scalar"] = scalar
    out.update(_boundary_template(f, dims))
    return out


def _boundary_template(record: dict, dims: Dims) -> dict:
    out = {}
    for var in BOUNDARY_VARS:
        file_var = "qvapor" if var == "qv" else var
        for part in ("b", "bt"):
            for edge in ("xs", "xe", "ys", "ye"):
                reg = f"{var}_{part}{edge}"
                v = record.get(f"{file_var}_{part}{edge}")
                if v is None:
                    continue
                out["qvapor" + reg[2:] if var == "qv" else reg] = _pad(v, _ref_shape(FIELD_KIND[reg], dims))
    return out


def boundary_record(path: str | Path, index: int, dims: Dims) -> dict:
    """Registry arrays of wrfbdy record ``index`` (values and tendencies)."""
    ref = _boundary_template(load_wrfbdy_fields(path, time_index=index), dims)
    return {registry_name(k): to_registry(registry_name(k), v, dims) for k, v in ref.items()}


def registry_arrays(case: CaseData, cfg: DycoreConfig) -> dict:
    """The case as registry arrays, ready for ``Domain.upload``."""
    dims = dims_of(cfg)
    return {registry_name(k): to_registry(registry_name(k), v, dims) for k, v in runtime_fields(case, dims).items()}


def dims_of(cfg: DycoreConfig) -> Dims:
    return Dims(nx=cfg.nx, ny=cfg.ny, nz=cfg.nz, nmoist=cfg.nmoist, nscalar=cfg.nscalar, spec_bdy_width=cfg.spec_bdy_width)


__all__ = ["BOUNDARY_VARS", "FILE_NAMES", "MOIST_SPECIES", "SCALAR_SPECIES", "boundary_record", "dims_of", "dycore_config",
           "registry_arrays", "runtime_fields", "species_counts"]
