"""Writes the committed golden vectors of the dycore step.

The reference ships no golden vectors for this path, so these fixtures are
outputs of the CPU oracle (FP32 and FP64 builds) on small seeded ideal cases,
kept to catch regressions of the oracle, the case generator and the HIP path.
Each fixture stores a digest of the generated inputs, the step count and the
prognostic fields after the steps.

    python tests/golden/make_golden.py
"""

from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "wrf-jax_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))

from oracle import Oracle  # noqa: E402
from wrfdc.cases import make_case  # noqa: E402

OUTPUTS = ("u_2", "v_2", "w_2", "t_2", "ph_2", "mu_2", "p", "al", "moist", "scalar", "ww", "ru_m")

FIXTURES = {
    # name: (make_case args, steps, precision)
    "step1_13x17x7_f32": (dict(nx=13, ny=17, nz=7, nmoist=6, nscalar=2, hydrometeors=True, hill_height=500.0), 1, "f32"),
    "step3_16x14x8_f32": (dict(nx=16, ny=14, nz=8, nmoist=2, nscalar=1, hydrometeors=True, u0=15.0), 3, "f32"),
    "step1_13x17x7_f64": (dict(nx=13, ny=17, nz=7, nmoist=6, nscalar=2, hydrometeors=True, hill_height=500.0), 1, "f64"),
}


def input_digest(arrays: dict) -> str:
    h = hashlib.sha256()
    for name in sorted(arrays):
        a = np.ascontiguousarray(arrays[name])
        h.update(name.encode())
        h.update(a.tobytes())
    return h.hexdigest()


def generate(name: str):
    args, steps, precision = FIXTURES[name]
    dtype = np.float32 if precision == "f32" else np.float64
    case = make_case(**args, dtype=dtype)
    o = Oracle(case.config, case.arrays, precision)
    o.step(steps, 0.0)
    out = {k: o.arrays[k].copy() for k in OUTPUTS}
    o.close()
    return case, input_digest(case.arrays), steps, out


def main():
    for name in FIXTURES:
        case, digest, steps, out = generate(name)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, input_digest=np.array(digest), steps=np.array(steps), **out)
        print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
