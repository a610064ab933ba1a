"""Device-resident dycore domain driven through the C ABI.

``Domain`` owns one float32 torch buffer per registry field (layout of
wrfdc.fields: [j][k][i] with a padded i pitch) on one GPU, binds them to a
``wrfdc_ctx`` and advances them with ``wrfdc_step`` / ``wrfdc_stage``.
PyTorch is used for device memory and the stream only; every operation on the
state is a HIP kernel of libwrfdc.so.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .config import DycoreConfig
from .fields import FIELD_SPECS, Dims

# values returned by wrfdc_checksums (WRFDC_NCHECKSUMS in include/wrfdc.h)
NCHECKSUMS = 6


def _torch():
    import torch

    return torch


class Domain:
    def __init__(self, config: DycoreConfig, device: str = "cuda:0"):
        torch = _torch()
        config.validate()
        self.lib = _lib.load()
        if not torch.cuda.is_available():
            raise RuntimeError("wrfdc.Domain needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.config = config
        self.device = torch.device(device)
        self.dims = Dims(nx=config.nx, ny=config.ny, nz=config.nz, nmoist=config.nmoist, nscalar=config.nscalar,
                         spec_bdy_width=config.spec_bdy_width)
        if self.dims.ip != config.ip:
            raise ValueError(f"config.ip={config.ip} differs from the registry pitch {self.dims.ip}")
        self._cfg = config.as_struct()
        ctx = ctypes.c_void_p()
        rc = self.lib.wrfdc_create(ctypes.byref(self._cfg), ctypes.byref(ctx))
        _lib.raise_for(rc, self.lib, None, "wrfdc_create")
        self.ctx = ctx
        self.buffers = {}
        with torch.cuda.device(self.device):
            for name, kind in FIELD_SPECS:
                t = torch.zeros(self.dims.shape(kind), dtype=torch.float32, device=self.device)
                self.buffers[name] = t
                ptr = t.data_ptr() if t.numel() else None
                self._check(self.lib.wrfdc_bind(self.ctx, name.encode(), ptr), f"wrfdc_bind({name})")
            self.stream = torch.cuda.current_stream(self.device)
            self._check(self.lib.wrfdc_set_stream(self.ctx, ctypes.c_void_p(self.stream.cuda_stream)), "wrfdc_set_stream")

    # ------------------------------------------------------------------
    def _check(self, rc: int, what: str) -> None:
        _lib.raise_for(rc, self.lib, self.ctx, what)

    def close(self) -> None:
        if getattr(self, "ctx", None):
            self.lib.wrfdc_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------
    def upload(self, arrays: dict) -> None:
        """Copy host arrays (registry layout) into the device buffers."""
        torch = _torch()
        for name, a in arrays.items():
            buf = self.buffers[name]
            src = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
            if tuple(src.shape) != tuple(buf.shape):
                raise ValueError(f"field {name}: shape {src.shape} != {tuple(buf.shape)}")
            if buf.numel():
                buf.copy_(torch.from_numpy(src))
        self.stream.synchronize()

    def download(self, names=None) -> dict:
        self.synchronize()
        names = [n for n, _ in FIELD_SPECS] if names is None else list(names)
        return {n: self.buffers[n].cpu().numpy() for n in names}

    def step(self, nsteps: int = 1, dtbc0: float = 0.0) -> None:
        """Advance nsteps RK3 steps (asynchronous on the domain's stream)."""
        self._check(self.lib.wrfdc_step(self.ctx, int(nsteps), float(dtbc0)), "wrfdc_step")

    def stage(self, name: str, rk_step: int = 1, iteration: int = 0) -> None:
        self._check(self.lib.wrfdc_stage(self.ctx, name.encode(), int(rk_step), int(iteration)), f"wrfdc_stage({name})")

    def set_tile(self, name: str, value: int) -> None:
        """A launch-shape setting (wrfdc_set_tile): "face_levels" = levels per
        level group of the face-block kernels, 0 for the per-kernel defaults.
        Results do not depend on it."""
        self._check(self.lib.wrfdc_set_tile(self.ctx, name.encode(), int(value)), f"wrfdc_set_tile({name})")

    def set_dtbc(self, dtbc: float) -> None:
        """Boundary time (s since the boundary record) that stage calls read."""
        self._check(self.lib.wrfdc_set_dtbc(self.ctx, float(dtbc)), "wrfdc_set_dtbc")

    def checksums(self) -> np.ndarray:
        """FP64 [sum(mu_2+mub), sum(t_2), sum(moist), n_nonfinite, dry-air mass (kg),
        total energy (J)]; raises FloatingPointError on NaN/Inf."""
        out = (ctypes.c_double * NCHECKSUMS)()
        self._check(self.lib.wrfdc_checksums(self.ctx, out), "wrfdc_checksums")
        return np.array(list(out))

    def synchronize(self) -> None:
        self._check(self.lib.wrfdc_sync(self.ctx), "wrfdc_sync")


__all__ = ["Domain"]
