"""The reference's typed step API: state containers, the clock and the
reference-layout <-> device-layout conversions (CPU), and one
step_em_core_once through the C ABI against the oracle (GPU)."""

import numpy as np
import pytest

from wrfdc.cases import make_case
from wrfdc.fields import FIELD_SPECS, Dims
from wrfdc.step_em import (ModelState, TimeState, from_registry, reference_name, registry_name, to_registry)


def _dims(case):
    c = case.config
    return Dims(c.nx, c.ny, c.nz, c.nmoist, c.nscalar, c.spec_bdy_width)


def test_layout_round_trip_every_field():
    case = make_case(13, 17, 7, nmoist=3, nscalar=2, hydrometeors=True)
    d = _dims(case)
    for name, kind in FIELD_SPECS:
        a = case.arrays[name]
        ref = from_registry(name, a, d)
        back = to_registry(name, ref, d)
        if kind in "32MSYy":  # the i padding beyond ide is not part of the state
            np.testing.assert_array_equal(back[..., : d.ide], a[..., : d.ide], err_msg=name)
        else:
            np.testing.assert_array_equal(back, a, err_msg=name)


def test_reference_shapes():
    case = make_case(13, 17, 7, nmoist=3, nscalar=2)
    d = _dims(case)
    A = case.arrays
    assert from_registry("u_2", A["u_2"], d).shape == (14, 8, 18)
    assert from_registry("mu_2", A["mu_2"], d).shape == (14, 18)
    m = from_registry("moist", A["moist"], d)
    assert m.shape == (14, 8, 18, 4) and not m[..., 0].any()  # slot 0 unused
    # the reference's (i, k, j) index of QV at (i=3, k=2, j=5) is registry [0, j-1, k-1, i-1]
    assert m[2, 1, 4, 1] == A["moist"][0, 4, 1, 2]
    assert from_registry("u_bxs", A["u_bxs"], d).shape == (18, 8, 5)
    assert from_registry("u_bys", A["u_bys"], d).shape == (14, 8, 5)
    assert from_registry("mu_bxs", A["mu_bxs"], d).shape == (18, 1, 5)
    assert reference_name("qv_btye") == "qvapor_btye" and registry_name("qvapor_bxs") == "qv_bxs"


def test_time_state_advance_matches_reference_rule():
    t = TimeState(np.int32(0), np.float32(0.0), np.float32(0.0))
    for _ in range(3):
        t = t.advance(18.0, 54)
    assert int(t.itimestep) == 3 and float(t.elapsed_seconds) == 54.0
    assert float(t.dtbc) == 0.0  # reset on the boundary-update time
    t = t.advance(18.0, 54)
    assert float(t.dtbc) == 18.0
    assert TimeState(np.int32(0), np.float32(0), np.float32(5)).advance(18.0, 0).dtbc == np.float32(23.0)


def test_model_state_validation():
    t = TimeState(np.int32(0), np.float32(0), np.float32(0))
    with pytest.raises(ValueError):
        ModelState(("b", "a"), (1, 2), t)
    with pytest.raises(ValueError):
        ModelState(("a", "a"), (1, 2), t)
    s = ModelState(("a", "b"), (1, 2), t)
    assert s.with_arrays({"b": 5}).require("b") == (5,)
    with pytest.raises(KeyError):
        s.with_arrays({"c": 1})
    assert s.select_arrays(["b"]).array_names == ("b",)


@pytest.mark.gpu
def test_step_em_core_once_matches_oracle(gpu):
    from oracle import Oracle
    from parity import REL_L2_STEP, rel_l2
    from wrfdc.step_em import build_step_em_core_context, initial_step_em_state, step_em_core_once

    case = make_case(20, 18, 9, nmoist=2, nscalar=1, hydrometeors=True, hill_height=300.0)
    d = _dims(case)
    template = {reference_name(n): from_registry(n, a, d) for n, a in case.arrays.items()}
    grid = case.config
    ctx = build_step_em_core_context(grid=grid, config_flags={"diff_opt": 2, "km_opt": 4}, fields_template=template,
                                     num_moist=grid.nmoist + 1, num_scalar=grid.nscalar, check_finite=True, device=gpu)
    state = initial_step_em_state(ctx)
    res = step_em_core_once(state, ctx)
    o = Oracle(case.config, case.arrays)
    o.step(1, 0.0)
    for name in ("u_2", "v_2", "w_2", "t_2", "ph_2", "mu_2", "moist", "scalar"):
        got = to_registry(name, res.state.as_dict()[name], d)
        assert rel_l2(got[..., : d.ide], o.arrays[name][..., : d.ide]) <= REL_L2_STEP, name
    assert int(res.state.time.itimestep) == 1 and float(res.state.time.dtbc) == grid.dt
    assert res.solve_result.checksums[3] == 0
    ctx.domain.close()
