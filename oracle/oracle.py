"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/_build/liboracle_f32.so`` (or the FP64 build, or the FP32
build with OpenMP, ``f32_omp``, which steps several domains at once) compiled from
``oracle/src`` and runs the loop-form restatement of the reference dycore on
numpy arrays laid out like the product's field registry.  Only tests/,
``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import this
module; the product never does.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(REPO, "wrf-jax_amd"))

from wrfdc.config import DycoreConfig  # noqa: E402
from wrfdc.fields import FIELD_SPECS, Dims  # noqa: E402

_LIBS: dict[str, ctypes.CDLL] = {}


def build() -> None:
    """Compile the oracle libraries (make in oracle/)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


# A rebuilt copy of the restatement (the mutation tests build one per planted
# slip) is loaded instead of oracle/_build when this variable names its folder.
BUILD_ENV = "WRFDC_ORACLE_BUILD"


def _lib(precision: str) -> ctypes.CDLL:
    if precision not in _LIBS:
        build_dir = os.environ.get(BUILD_ENV) or os.path.join(HERE, "_build")
        path = os.path.join(build_dir, f"liboracle_{precision}.so")
        if not os.path.exists(path):
            if os.environ.get(BUILD_ENV):
                raise FileNotFoundError(path)
            build()
        lib = ctypes.CDLL(path)
        lib.o_create.restype = ctypes.c_void_p
        lib.o_create.argtypes = [ctypes.c_void_p]
        lib.o_destroy.argtypes = [ctypes.c_void_p]
        lib.o_bind.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
        lib.o_bind.restype = ctypes.c_int
        lib.o_run_stage.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
        lib.o_run_stage.restype = ctypes.c_int
        lib.o_set_dtbc.argtypes = [ctypes.c_void_p, ctypes.c_double]
        lib.o_step_many.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        lib.o_step_many.restype = ctypes.c_int
        lib.o_field_name.restype = ctypes.c_char_p
        lib.o_field_name.argtypes = [ctypes.c_int]
        lib.o_field_kind.argtypes = [ctypes.c_int]
        # the registry of the C restatement must match the product registry
        names = [(lib.o_field_name(i).decode(), chr(lib.o_field_kind(i))) for i in range(lib.o_field_count())]
        if tuple(names) != FIELD_SPECS:
            raise RuntimeError("oracle field table differs from wrfdc.fields.FIELD_SPECS")
        _LIBS[precision] = lib
    return _LIBS[precision]


class Oracle:
    """One domain held in host numpy arrays, advanced by the C restatement."""

    def __init__(self, config: DycoreConfig, arrays: dict, precision: str = "f32"):
        self.precision = precision
        self.dtype = np.float32 if precision.startswith("f32") else np.float64
        self.lib = _lib(precision)
        self.config = config
        self.dims = Dims(nx=config.nx, ny=config.ny, nz=config.nz, nmoist=config.nmoist, nscalar=config.nscalar,
                         spec_bdy_width=config.spec_bdy_width)
        self.arrays = {}
        for name, kind in FIELD_SPECS:
            shape = self.dims.shape(kind)
            if kind in "32MS":
                shape = shape[:-1] + (config.ip,)
            src = arrays.get(name)
            a = np.zeros(shape, dtype=self.dtype) if src is None else np.array(src, dtype=self.dtype, copy=True)
            if a.shape != shape:
                raise ValueError(f"field {name}: shape {a.shape} != {shape}")
            self.arrays[name] = np.ascontiguousarray(a)
        self._cfg = config.as_struct()
        self.ctx = self.lib.o_create(ctypes.byref(self._cfg))
        for name, a in self.arrays.items():
            if self.lib.o_bind(self.ctx, name.encode(), a.ctypes.data_as(ctypes.c_void_p)) != 0:
                raise KeyError(name)

    def close(self) -> None:
        if getattr(self, "ctx", None):
            self.lib.o_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    def set_dtbc(self, dtbc: float) -> None:
        self.lib.o_set_dtbc(self.ctx, float(dtbc))

    def stage(self, name: str, rk_step: int = 1, iteration: int = 0) -> None:
        if self.lib.o_run_stage(self.ctx, name.encode(), int(rk_step), int(iteration)) != 0:
            raise KeyError(f"unknown oracle stage {name!r}")

    def step(self, nsteps: int = 1, dtbc0: float = 0.0) -> None:
        """Advance nsteps RK3 steps; dtbc = time since the boundary update."""
        for n in range(nsteps):
            self.set_dtbc(dtbc0 + (n + 1) * self.config.dt)
            self.stage("step", 3, 0)

    @staticmethod
    def step_many(oracles, nsteps: int, dtbc0: float = 0.0, threads: int = 1) -> None:
        """Advance independent oracles of one library nsteps steps each,
        ``threads`` at a time (the f32_omp build runs them concurrently)."""
        lib = oracles[0].lib
        if any(o.lib is not lib for o in oracles):
            raise ValueError("step_many needs oracles of one build")
        ctxs = (ctypes.c_void_p * len(oracles))(*[o.ctx for o in oracles])
        if lib.o_step_many(ctxs, len(oracles), int(nsteps), float(dtbc0), int(threads)) != 0:
            raise ValueError("o_step_many: bad arguments")

    def copy(self) -> dict:
        return {k: v.copy() for k, v in self.arrays.items()}


__all__ = ["Oracle", "build"]
