"""Shared helpers of the parity tests: field comparison against the oracle."""

from __future__ import annotations

import numpy as np

from wrfdc.fields import FIELD_SPECS

# per-field bar after one RK3 step / one stage (relative L2, FP32 vs FP32 oracle)
REL_L2_STEP = 1e-5
# NRMSE bar after 100 steps
NRMSE_LONG = 1e-3


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    nd = np.linalg.norm(a - b)
    if nb == 0.0:
        return 0.0 if nd == 0.0 else float("inf")
    return float(nd / nb)


def nrmse(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    rng = float(b.max() - b.min())
    rmse = float(np.sqrt(np.mean((a - b) ** 2)))
    if rng == 0.0:
        return 0.0 if rmse == 0.0 else float("inf")
    return rmse / rng


# Diagnostic outputs no later stage of the step reads, which the fused step
# does not write: the last species' advective tendency of the reference's
# species stage (runtime.py:14830), and calculate_km_kh's N^2 with km_opt 4
# (module_diffusion_em_fast.py:933-1005; the step computes it only when the
# learned closure takes it as an input).
STEP_DIAGNOSTICS = ("advect_tend", "bn2")


def compare(gpu: dict, ref: dict, names=None, tol: float = REL_L2_STEP, skip=()) -> dict:
    """Relative L2 of every named field; returns {name: err} of the failures."""
    names = [n for n, _ in FIELD_SPECS if n not in skip] if names is None else names
    bad = {}
    for n in names:
        e = rel_l2(gpu[n], ref[n])
        if not (e <= tol):
            bad[n] = e
    return bad
