"""Learned eddy viscosity for smag_km: data, training and use in the step.

The network (55 -> 64 -> 64 -> 1, bf16 MFMA hidden layers) and every kernel
that evaluates or trains it live in ``csrc/k_closure.h`` behind the C ABI
(``wrfdc_set_closure``, ``wrfdc_closure_rows``, ``wrfdc_subgrid_stress``,
``wrfdc_coarse_grain``, ``wrfdc_mlp_*``).  This module is the host side:

* ``coarse_grain`` / ``training_rows``: a fine domain (dx) run by the dycore
  is block-averaged onto a domain coarser by ``ratio``; each coarse point gives
  one row of features (the 3x3x3 stencils of log|D| and N^2, and log dx) and
  targets (the fine run's sub-grid stresses t1 = (tau11 - tau22)/2, t2 = tau12,
  the coarse deformation s1 = (D11 - D22)/2, s2 = D12, the slope divisor and
  the Smagorinsky km);
* ``Network``: parameters in device memory, Adam training on minibatches
  (the minibatch gather is torch indexing, the rest HIP kernels);
* ``flux_error``: the score of a diffusivity, sum((t + K s)^2) / sum(t^2) over
  rows, i.e. how much of the fine run's sub-grid stress the eddy-viscosity
  model K*s leaves unexplained (0 perfect, 1 = no model).

The reference computes km with Smagorinsky (calculate_km_kh_vectorized,
src/wrf_jax/dyn_em/module_diffusion_em_fast.py:773-1083); the network
replaces its km = min(c_s^2 l^2 |D|, 10 l), Smagorinsky stays the fallback.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _lib

NPARAMS = 8385
WBF_SIZE = 12288
NFEAT = 64
NTARGET = 6
# input slots: h*32 + s, stencil point s < 27 (di = s%3-1, dk = (s/3)%3-1,
# dj = s/9-1) of log|D| (h = 0) or N^2 (h = 1); slot 27 log(dx / 1 km)
FEAT_CENTRE_D, FEAT_CENTRE_N, FEAT_DX = 13, 32 + 13, 27
OFF = dict(W1=0, b1=4096, W2=4160, b2=8256, w3=8320, b3=8384)


class CNorm(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("d_shift", "d_scale", "n_shift", "n_scale", "dx_feature", "kref")]


@dataclass(frozen=True)
class ClosureNorm:
    """Feature normalisation and output scale (wrfdc_closure_norm)."""

    d_shift: float = 0.0
    d_scale: float = 1.0
    n_shift: float = 0.0
    n_scale: float = 1.0
    dx_feature: float = 0.0
    kref: float = 1.0

    def as_struct(self) -> CNorm:
        return CNorm(self.d_shift, self.d_scale, self.n_shift, self.n_scale, self.dx_feature, self.kref)


def _torch():
    import torch

    return torch


def _stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def init_theta(seed: int = 0) -> np.ndarray:
    """He-style initial parameters, seeded."""
    rng = np.random.default_rng(seed)
    t = np.zeros(NPARAMS)
    t[OFF["W1"]:OFF["b1"]] = rng.standard_normal(4096) * np.sqrt(2.0 / 55.0)
    t[OFF["W2"]:OFF["b2"]] = rng.standard_normal(4096) * np.sqrt(2.0 / 64.0)
    t[OFF["w3"]:OFF["b3"]] = rng.standard_normal(64) * np.sqrt(1.0 / 64.0)
    return t


def row_count(config) -> int:
    return (config.nx - 2) * config.nz * (config.ny - 2)


# ---------------------------------------------------------------------------
# data
# ---------------------------------------------------------------------------
def coarse_grain(fine, coarse, ratio: int) -> None:
    """Block means of the fine domain's dynamic fields into the coarse one."""
    fine._check(fine.lib.wrfdc_coarse_grain(fine.ctx, coarse.ctx, int(ratio)), "wrfdc_coarse_grain")


def training_rows(coarse, norm: ClosureNorm, fine=None, ratio: int = 0):
    """(X bf16 [n, 64], T f32 [n, 6]) of every point of the coarse domain's
    smag2d_km range; with a fine domain, columns 0..1 of T take its sub-grid
    stresses (otherwise they are zero)."""
    torch = _torch()
    n = row_count(coarse.config)
    X = torch.zeros((n, NFEAT), dtype=torch.bfloat16, device=coarse.device)
    T = torch.zeros((n, NTARGET), dtype=torch.float32, device=coarse.device)
    nm = norm.as_struct()
    coarse._check(coarse.lib.wrfdc_closure_rows(coarse.ctx, ctypes.byref(nm), _ptr(X), _ptr(T)), "wrfdc_closure_rows")
    if fine is not None:
        c = coarse.config
        fine._check(fine.lib.wrfdc_subgrid_stress(fine.ctx, c.nx, c.ny, int(ratio), _ptr(T)), "wrfdc_subgrid_stress")
    torch.cuda.current_stream().synchronize()
    return X, T


def fit_norm(coarse, c_s: float = 0.25) -> ClosureNorm:
    """Feature normalisation from the coarse state: mean and spread of log|D|
    and N^2 over its points (read from rows built with the identity
    normalisation), log(dx / 1 km), and kref = c_s^2 dx dy exp(mean log|D|),
    the Smagorinsky km of a typical deformation."""
    cfg = coarse.config
    X, _ = training_rows(coarse, ClosureNorm(dx_feature=0.0, kref=1.0))
    x = X.float().cpu().numpy()
    ld, n2 = x[:, FEAT_CENTRE_D].astype(np.float64), x[:, FEAT_CENTRE_N].astype(np.float64)
    d_shift, d_scale = float(ld.mean()), float(max(ld.std(), 1e-3))
    n_shift, n_scale = float(n2.mean()), float(max(n2.std(), 1e-8))
    kref = c_s * c_s * cfg.dx * cfg.dy * math.exp(d_shift)
    return ClosureNorm(d_shift, d_scale, n_shift, n_scale, math.log(cfg.dx / 1000.0), kref)


def flux_error(K, T) -> float:
    """sum((t1 + K s1)^2 + (t2 + K s2)^2) / sum(t1^2 + t2^2) over the rows (FP64)."""
    K = np.asarray(K, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    r1 = T[:, 0] + K * T[:, 2]
    r2 = T[:, 1] + K * T[:, 3]
    return float(np.sum(r1 * r1 + r2 * r2) / np.sum(T[:, 0] ** 2 + T[:, 1] ** 2))


# ---------------------------------------------------------------------------
# network
# ---------------------------------------------------------------------------
class Network:
    """Parameters (FP32 master + bf16 weight copies) and Adam moments on one GPU."""

    def __init__(self, norm: ClosureNorm, device: str = "cuda:0", seed: int = 0, theta: np.ndarray | None = None):
        torch = _torch()
        self.lib = _lib.load()
        self.norm = norm
        self.device = torch.device(device)
        init = init_theta(seed) if theta is None else np.asarray(theta, dtype=np.float64)
        self.theta = torch.tensor(init.astype(np.float32), device=self.device)
        self.wbf = torch.zeros(WBF_SIZE, dtype=torch.bfloat16, device=self.device)
        self.m = torch.zeros(NPARAMS, dtype=torch.float32, device=self.device)
        self.v = torch.zeros(NPARAMS, dtype=torch.float32, device=self.device)
        self.steps = 0
        self._work = None
        self._loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        self._check(self.lib.wrfdc_mlp_pack(_ptr(self.theta), _ptr(self.wbf), _stream(torch)), "wrfdc_mlp_pack")

    def _check(self, rc, what):
        _lib.raise_for(rc, self.lib, None, what)

    def _work_for(self, n: int):
        torch = _torch()
        need = int(self.lib.wrfdc_mlp_work_floats(int(n)))
        if self._work is None or self._work.numel() < need:
            self._work = torch.empty(need, dtype=torch.float32, device=self.device)
        return self._work

    def gradient(self, X, T, lossnorm: float):
        """(loss, gradient) of the loss sum over the rows, on the GPU."""
        torch = _torch()
        grad = torch.zeros(NPARAMS, dtype=torch.float32, device=self.device)
        n = X.shape[0]
        self._check(self.lib.wrfdc_mlp_grad(_ptr(self.theta), _ptr(self.wbf), _ptr(X), _ptr(T), int(n), float(lossnorm),
                                            float(self.norm.kref), _ptr(self._work_for(n)), _ptr(grad), _ptr(self._loss),
                                            _stream(torch)), "wrfdc_mlp_grad")
        torch.cuda.current_stream().synchronize()
        return float(self._loss.item()), grad.cpu().numpy().astype(np.float64)

    def train_step(self, X, T, lossnorm: float, lr: float) -> None:
        """One Adam step on the rows (stream-ordered, no host wait)."""
        torch = _torch()
        self.steps += 1
        n = X.shape[0]
        self._check(self.lib.wrfdc_mlp_train_step(_ptr(self.theta), _ptr(self.wbf), _ptr(X), _ptr(T), int(n), float(lossnorm),
                                                  float(self.norm.kref), _ptr(self._work_for(n)), _ptr(self.m), _ptr(self.v),
                                                  int(self.steps), float(lr), _ptr(self._loss), _stream(torch)),
                    "wrfdc_mlp_train_step")

    def last_loss(self) -> float:
        return float(self._loss.item())

    def z(self, X):
        """Network output per row (before softplus)."""
        torch = _torch()
        out = torch.zeros(X.shape[0], dtype=torch.float32, device=self.device)
        self._check(self.lib.wrfdc_mlp_forward(_ptr(self.theta), _ptr(self.wbf), _ptr(X), int(X.shape[0]), _ptr(out),
                                               _stream(torch)), "wrfdc_mlp_forward")
        torch.cuda.current_stream().synchronize()
        return out.cpu().numpy().astype(np.float64)

    def km(self, X, T) -> np.ndarray:
        """The diffusivity the step would use at each row: kref*softplus(z) / divisor."""
        z = self.z(X)
        sp = np.where(z > 20.0, z, np.log1p(np.exp(np.minimum(z, 20.0))))
        return self.norm.kref * sp / T[:, 4].astype(np.float64)

    def fit(self, X, T, epochs: int, batch: int = 4096, lr: float = 2e-3, seed: int = 0, log=None) -> list:
        """Adam over shuffled minibatches; returns the loss per epoch."""
        torch = _torch()
        n = X.shape[0]
        t = T.float()
        lossnorm = 1.0 / max(float((t[:, 0] ** 2 + t[:, 1] ** 2).mean().item()) * batch, 1e-30)
        gen = torch.Generator(device="cpu")
        gen.manual_seed(seed)
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen).to(self.device)
            tot = 0.0
            nb = 0
            for b0 in range(0, n - batch + 1, batch):
                idx = perm[b0:b0 + batch]
                self.train_step(X[idx].contiguous(), T[idx].contiguous(), lossnorm, lr)
                tot += self.last_loss()
                nb += 1
            hist.append(tot / max(nb, 1))
            if log:
                log(f"epoch {ep + 1}: loss {hist[-1]:.4f}")
        return hist

    def attach(self, domain) -> None:
        """Use the network for km in the domain's step."""
        nm = self.norm.as_struct()
        self._norm_struct = nm  # the library keeps the pointers of theta and wbf, the norm by value
        domain._check(domain.lib.wrfdc_set_closure(domain.ctx, _ptr(self.theta), _ptr(self.wbf), ctypes.byref(nm)),
                      "wrfdc_set_closure")

    @staticmethod
    def detach(domain) -> None:
        domain._check(domain.lib.wrfdc_set_closure(domain.ctx, None, None, None), "wrfdc_set_closure")


# ---------------------------------------------------------------------------
# a truth run and a trained network
# ---------------------------------------------------------------------------
def truth_pair(device: str, nxc: int = 24, nz: int = 10, ratio: int = 3, fine_steps: int = 2, dx: float = 3000.0, dt: float = 18.0,
               nmoist: int = 2):
    """(fine, coarse) Domains of the seeded hill case: the fine one at
    dx / ratio and dt / ratio (nxc * ratio columns a side), advanced
    fine_steps steps; the coarse one at dx, dt (nxc columns a side)."""
    from .cases import make_case
    from .device import Domain

    fine_case = make_case(nxc * ratio, nxc * ratio, nz, dx=dx / ratio, dt=dt / ratio, nmoist=nmoist, nscalar=0, hill_height=300.0)
    coarse_case = make_case(nxc, nxc, nz, dx=dx, dt=dt, nmoist=nmoist, nscalar=0, hill_height=300.0)
    fine = Domain(fine_case.config, device)
    fine.upload(fine_case.arrays)
    if fine_steps:
        fine.step(fine_steps, 0.0)
    coarse = Domain(coarse_case.config, device)
    coarse.upload(coarse_case.arrays)
    return fine, coarse


def train_on_truth(device: str, nxc: int = 32, nz: int = 16, ratio: int = 3, fine_steps: int = 30, epochs: int = 30,
                   batch: int = 1024, lr: float = 2e-3, dx: float = 3000.0, dt: float = 18.0, log=None):
    """A network trained on the sub-grid stresses of a fine run of this
    dycore (dx / ratio) coarse-grained to dx: returns (network, training
    loss per epoch).  The fine and coarse domains are released."""
    fine, coarse = truth_pair(device, nxc=nxc, nz=nz, ratio=ratio, fine_steps=fine_steps, dx=dx, dt=dt)
    try:
        coarse_grain(fine, coarse, ratio)
        norm = fit_norm(coarse)
        X, T = training_rows(coarse, norm, fine, ratio)
        net = Network(norm, device, seed=0)
        hist = net.fit(X, T, epochs=epochs, batch=batch, lr=lr, log=log)
    finally:
        fine.close()
        coarse.close()
    return net, hist


# 2 * (55 * 64 + 64 * 64 + 64) floating-point operations per point of the
# network's km (both hidden layers counted with their 55 / 64 real inputs)
FLOP_PER_POINT = 2 * (55 * 64 + 64 * 64 + 64)


__all__ = ["FEAT_CENTRE_D", "FEAT_CENTRE_N", "FEAT_DX", "ClosureNorm", "FLOP_PER_POINT", "Network", "coarse_grain", "fit_norm", "flux_error", "init_theta", "row_count", "train_on_truth",
           "training_rows", "truth_pair"]
