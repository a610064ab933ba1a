"""Build libwrfdc.so from csrc/ with hipcc for gfx950 (in-tree output).

FP contraction is off so every kernel rounds exactly the operations it writes,
in the order it writes them, like the FP32 arithmetic of the reference path;
packed FP32 instructions are disabled for all device code.
"""

from __future__ import annotations

import os
import subprocess

from ._lib import LIB_PATH

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(PKG), "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Werror=uninitialized",
]


def build(verbose: bool = False) -> str:
    src = os.path.join(CSRC, "wrfdc.hip")
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, src]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stderr}")
    if verbose and res.stderr:
        print(res.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
