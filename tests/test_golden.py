"""Committed golden vectors (tests/golden/*.npz, written by make_golden.py):
the case generator reproduces the recorded inputs and the oracle the
recorded outputs, bit for bit."""

import os

import numpy as np
import pytest

from golden.make_golden import FIXTURES, OUTPUTS, generate

HERE = os.path.dirname(os.path.abspath(__file__))


def load(name):
    with np.load(os.path.join(HERE, "golden", f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(FIXTURES))
def test_oracle_reproduces_golden(name):
    ref = load(name)
    case, digest, steps, out = generate(name)
    assert str(ref["input_digest"]) == digest, "case generator drifted from the recorded inputs"
    assert int(ref["steps"]) == steps
    for k in OUTPUTS:
        np.testing.assert_array_equal(out[k], ref[k], err_msg=k)


def test_golden_f32_tracks_f64():
    a = load("step1_13x17x7_f32")
    b = load("step1_13x17x7_f64")
    for k in ("u_2", "t_2", "ph_2", "moist"):
        x = a[k].astype(np.float64)
        assert np.linalg.norm(x - b[k]) <= 1e-3 * np.linalg.norm(b[k]), k
