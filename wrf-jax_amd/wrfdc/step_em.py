"""The reference's typed one-step API over the HIP dycore.

Mirrors src/wrf_jax/core/step_em.py:15-93 (``StepEmCoreContext``,
``StepEmCoreResult``, ``build_step_em_core_context``,
``initial_step_em_state``, ``step_em_core_once``) and the state containers of
src/wrf_jax/core/state.py:12-96 (``TimeState``, ``ModelState``), so an
em_real time loop written against the reference can call this module
instead.  Arrays cross this boundary in the reference's layout: 3-D fields
(i, k, j) at staggered extents (ide, kde, jde), 2-D fields (i, j), 1-D
columns (k,), moisture as (ide, kde, jde, n) with slot 0 unused
(PARAM_FIRST_SCALAR = 2), x-edge boundary arrays (j, k, bdy_width), y-edge
(i, k, bdy_width), 2-D boundary arrays (j|i, 1, bdy_width).  The context
keeps the whole registry resident on the GPU; conversion happens only for
the arrays the state carries.
"""

from __future__ import annotations

from collections.abc import Iterable, Mapping
from dataclasses import dataclass, field, replace
from typing import Any

import numpy as np

from .config import DycoreConfig
from .fields import FIELD_SPECS, FIELD_KIND, Dims

# reference name of a registry field where the two differ
_REF_NAME = {f"qv_{p}{e}": f"qvapor_{p}{e}" for p in ("b", "bt") for e in ("xs", "xe", "ys", "ye")}
_REG_NAME = {v: k for k, v in _REF_NAME.items()}


def reference_name(name: str) -> str:
    return _REF_NAME.get(name, name)


def registry_name(ref: str) -> str:
    return _REG_NAME.get(ref, ref)


# ---------------------------------------------------------------------------
# state containers (src/wrf_jax/core/state.py:12-96)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class TimeState:
    """Model clock: step counter, elapsed seconds and time since the last
    lateral-boundary update (int32 / float32 like the reference)."""

    itimestep: Any
    elapsed_seconds: Any
    dtbc: Any

    @classmethod
    def from_grid(cls, grid: Any) -> "TimeState":
        return cls(
            itimestep=np.int32(int(getattr(grid, "itimestep", 0))),
            elapsed_seconds=np.float32(float(getattr(grid, "current_time", 0.0))),
            dtbc=np.float32(float(getattr(grid, "dtbc", 0.0))),
        )

    def advance(self, dt_seconds: float, boundary_interval_seconds: int) -> "TimeState":
        elapsed = np.float32(self.elapsed_seconds) + np.float32(dt_seconds)
        dtbc = np.float32(self.dtbc) + np.float32(dt_seconds)
        interval = int(boundary_interval_seconds)
        if interval > 0 and elapsed > 0.0 and np.mod(elapsed, np.float32(interval)) == 0.0:
            dtbc = np.float32(0.0)
        return replace(self, itimestep=np.int32(int(self.itimestep) + 1), elapsed_seconds=np.float32(elapsed), dtbc=np.float32(dtbc))


@dataclass(frozen=True)
class ModelState:
    """Fixed-name, sorted set of arrays plus the model time."""

    array_names: tuple
    arrays: tuple
    time: TimeState

    def __post_init__(self) -> None:
        if len(self.array_names) != len(self.arrays):
            raise ValueError("ModelState array_names and arrays must have the same length")
        if tuple(sorted(self.array_names)) != tuple(self.array_names):
            raise ValueError("ModelState array_names must be sorted for stable PyTree structure")
        if len(set(self.array_names)) != len(self.array_names):
            raise ValueError("ModelState array_names must be unique")

    def as_dict(self) -> dict:
        return dict(zip(self.array_names, self.arrays, strict=True))

    def select_arrays(self, names: Iterable[str]) -> "ModelState":
        selected = tuple(sorted(names))
        values = self.as_dict()
        return ModelState(selected, tuple(values[n] for n in selected), self.time)

    def with_arrays(self, updates: dict) -> "ModelState":
        if not updates:
            return self
        values = self.as_dict()
        unknown = set(updates) - set(values)
        if unknown:
            raise KeyError(f"cannot update unknown ModelState arrays: {', '.join(sorted(unknown))}")
        values.update(updates)
        return replace(self, arrays=tuple(values[n] for n in self.array_names))

    def require(self, *names: str) -> tuple:
        values = self.as_dict()
        return tuple(values[n] for n in names)


# ---------------------------------------------------------------------------
# layout conversion: reference arrays <-> registry arrays
# ---------------------------------------------------------------------------
def to_registry(name: str, ref: np.ndarray, dims: Dims) -> np.ndarray:
    """Reference-layout array of registry field ``name`` -> registry layout."""
    kind = FIELD_KIND[name]
    a = np.asarray(ref, dtype=np.float32)
    out = np.zeros(dims.shape(kind), dtype=np.float32)
    if kind == "3":
        out[:, :, : dims.ide] = a.transpose(2, 1, 0)
    elif kind == "2":
        out[:, : dims.ide] = a.transpose(1, 0)
    elif kind == "1":
        out[:] = a.reshape(-1)[: dims.kde]
    elif kind in "MS":
        n = out.shape[0]
        for s in range(n):
            out[s, :, :, : dims.ide] = a[..., s + 1].transpose(2, 1, 0)
    elif kind == "X":
        out[:] = a.transpose(2, 1, 0)
    elif kind == "Y":
        out[:, :, : dims.ide] = a.transpose(2, 1, 0)
    elif kind == "x":
        out[:] = a[:, 0, :].transpose(1, 0)
    elif kind == "y":
        out[:, : dims.ide] = a[:, 0, :].transpose(1, 0)
    return out


def from_registry(name: str, reg: np.ndarray, dims: Dims) -> np.ndarray:
    kind = FIELD_KIND[name]
    a = np.asarray(reg)
    if kind == "3":
        return np.ascontiguousarray(a[:, :, : dims.ide].transpose(2, 1, 0))
    if kind == "2":
        return np.ascontiguousarray(a[:, : dims.ide].transpose(1, 0))
    if kind == "1":
        return a.copy()
    if kind in "MS":
        n = a.shape[0]
        out = np.zeros((dims.ide, dims.kde, dims.jde, n + 1), dtype=np.float32)
        for s in range(n):
            out[..., s + 1] = a[s, :, :, : dims.ide].transpose(2, 1, 0)
        return out
    if kind == "X":
        return np.ascontiguousarray(a.transpose(2, 1, 0))
    if kind == "Y":
        return np.ascontiguousarray(a[:, :, : dims.ide].transpose(2, 1, 0))
    if kind == "x":
        return np.ascontiguousarray(a.transpose(1, 0)[:, None, :])
    if kind == "y":
        return np.ascontiguousarray(a[:, : dims.ide].transpose(1, 0)[:, None, :])
    raise ValueError(kind)


# the prognostic arrays a ModelState carries between steps (reference names)
DYNAMIC_ARRAY_NAMES = tuple(sorted(reference_name(n) for n, _ in FIELD_SPECS))


# ---------------------------------------------------------------------------
# step API (src/wrf_jax/core/step_em.py:15-93)
# ---------------------------------------------------------------------------
@dataclass
class SolveResult:
    """Outcome of one HIP step: FP64 checksums [sum(mu_2+mub), sum(t_2),
    sum(moist), n_nonfinite, dry-air mass, total energy] when requested."""

    checksums: Any = None


@dataclass(frozen=True)
class StepEmCoreContext:
    domain: Any
    grid: DycoreConfig
    config_flags: Mapping
    fields_template: Mapping
    dynamic_only: bool = True
    boundary_interval_seconds: int = 0
    check_finite: bool = False
    dims: Dims = field(default=None)


@dataclass(frozen=True)
class StepEmCoreResult:
    state: ModelState
    fields: Mapping
    solve_result: SolveResult


def build_step_em_core_context(*, grid: DycoreConfig, config_flags: Mapping | None = None, fields_template: Mapping,
                               callbacks: Mapping | None = None, num_moist: int = 2, num_scalar: int = 0,
                               dynamic_only: bool = True, boundary_interval_seconds: int = 0,
                               check_finite: bool = False, device: str = "cuda:0") -> StepEmCoreContext:
    """Validate the configuration, allocate the device-resident domain and
    upload the template (reference layout).  ``callbacks`` is accepted for
    signature compatibility; every stage runs on the GPU.  ``num_moist``
    counts the reference's moist slots including the unused slot 0."""
    from .device import Domain

    opts = dict(config_flags or {})
    cfg = replace(grid, nmoist=int(num_moist) - 1, nscalar=int(num_scalar), options=opts).validate()
    dims = Dims(nx=cfg.nx, ny=cfg.ny, nz=cfg.nz, nmoist=cfg.nmoist, nscalar=cfg.nscalar, spec_bdy_width=cfg.spec_bdy_width)
    dom = Domain(cfg, device)
    template = dict(fields_template)
    upload = {}
    for name, _ in FIELD_SPECS:
        ref = reference_name(name)
        if ref in template:
            upload[name] = to_registry(name, template[ref], dims)
    dom.upload(upload)
    return StepEmCoreContext(domain=dom, grid=cfg, config_flags=opts, fields_template=template, dynamic_only=bool(dynamic_only),
                             boundary_interval_seconds=int(boundary_interval_seconds), check_finite=bool(check_finite), dims=dims)


def initial_step_em_state(context: StepEmCoreContext) -> ModelState:
    t = context.fields_template
    names = tuple(sorted(n for n in DYNAMIC_ARRAY_NAMES if n in t)) if context.dynamic_only else tuple(sorted(t))
    return ModelState(array_names=names, arrays=tuple(t[n] for n in names), time=TimeState.from_grid(context.grid))


def step_em_core_once(state: ModelState, context: StepEmCoreContext) -> StepEmCoreResult:
    """Advance one RK3 step: the state's arrays are uploaded (reference layout
    -> device layout), the step runs on the GPU, every carried array is read
    back and the clock advances by dt."""
    dom, dims = context.domain, context.dims
    upload = {}
    for ref, arr in state.as_dict().items():
        name = registry_name(ref)
        if name in FIELD_KIND:
            upload[name] = to_registry(name, arr, dims)
    dom.upload(upload)
    dom.step(1, float(state.time.dtbc))
    ck = dom.checksums() if context.check_finite else None
    got = dom.download([registry_name(r) for r in state.array_names if registry_name(r) in FIELD_KIND])
    new_arrays = dict(state.as_dict())
    for name, a in got.items():
        new_arrays[reference_name(name)] = from_registry(name, a, dims)
    time = state.time.advance(context.grid.dt, context.boundary_interval_seconds)
    next_state = ModelState(state.array_names, tuple(new_arrays[n] for n in state.array_names), time)
    fields = dict(context.fields_template)
    fields.update(new_arrays)
    return StepEmCoreResult(state=next_state, fields=fields, solve_result=SolveResult(checksums=ck))


__all__ = [
    "DYNAMIC_ARRAY_NAMES", "ModelState", "SolveResult", "StepEmCoreContext", "StepEmCoreResult", "TimeState",
    "build_step_em_core_context", "from_registry", "initial_step_em_state", "reference_name", "registry_name",
    "step_em_core_once", "to_registry",
]
