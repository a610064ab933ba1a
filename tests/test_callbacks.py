"""The reference-named stage callbacks and the solve_em driver (CPU).

The driver and the callback dict are the library's; the domain here is the
oracle, which has the same ``config`` / ``stage`` / ``set_dtbc`` interface
and runs every stage name through its own C restatement.  Checks:

* one step driven by ``solve_em`` through ``strict_solve_callbacks()`` is
  bit-identical to the oracle's own step (whose order restates
  direct_step.py:3037-3422 and solve_em.py:700-790), with and without
  positive-definite species advection, with scalars, and at a nonzero
  boundary time;
* the event order matches solve_em.py:470-790 in the properties that a
  wrong order would break (acoustic-loop counts, first-stage-only stages,
  the final boundary stages last);
* a caller's replacement of one callback is what runs;
* every numerical stage name is accepted by the oracle's stage dispatch,
  and an unknown name is refused.
"""

import numpy as np
import pytest

from oracle import Oracle

from wrfdc.callbacks import IDENTITY_STAGES, NUMERICAL_STAGES, rk_step_sizes, solve_em, strict_solve_callbacks
from wrfdc.cases import make_case


def _case(**kw):
    args = dict(nmoist=3, nscalar=1, hydrometeors=True, hill_height=400.0)
    args.update(kw)
    return make_case(20, 18, 9, **args)


def _forcing(case, seed=7):
    rng = np.random.default_rng(seed)
    for name in ("u_btxs", "v_btye", "t_btys", "mu_btxe", "qv_btxs", "ph_btys"):
        case.arrays[name][:] = (1e-3 * rng.standard_normal(case.arrays[name].shape)).astype(np.float32)


@pytest.mark.parametrize("pd", [1, 0], ids=["pd", "plain"])
def test_driver_matches_oracle_step(pd):
    from dataclasses import replace

    case = _case()
    _forcing(case)
    cfg = replace(case.config, pd_moist=pd, pd_scalar=pd)
    dtbc0 = 2.5 * cfg.dt
    ref = Oracle(cfg, case.arrays)
    ref.step(1, dtbc0)
    via = Oracle(cfg, case.arrays)
    res = solve_em(via, strict_solve_callbacks(), dtbc0=dtbc0)
    assert res.dtbc == dtbc0 + cfg.dt
    for name, a in ref.arrays.items():
        assert np.array_equal(a.view(np.uint32), via.arrays[name].view(np.uint32)), name
    ref.close()
    via.close()


def test_two_steps_and_no_scalars():
    case = _case(nscalar=0, nmoist=1, hydrometeors=False)
    cfg = case.config
    ref = Oracle(cfg, case.arrays)
    ref.step(2, 0.0)
    via = Oracle(cfg, case.arrays)
    solve_em(via, dtbc0=0.0)
    solve_em(via, dtbc0=cfg.dt)
    for name, a in ref.arrays.items():
        assert np.array_equal(a.view(np.uint32), via.arrays[name].view(np.uint32)), name
    ref.close()
    via.close()


def test_event_order():
    case = _case()
    cfg = case.config
    calls = []

    class Recorder:
        config = cfg

        def set_dtbc(self, dtbc):
            calls.append(("dtbc", dtbc))

        def stage(self, name, rk, it):
            calls.append((name, rk, it))

    res = solve_em(Recorder(), strict_solve_callbacks(), dtbc0=0.0)
    ev = res.events
    ns = cfg.num_sound_steps()
    assert ev[0] == "initialize_moist_old" and ev[-1] == "after_all_rk_steps"
    assert ev.count("rk_step_prep") == 3
    assert ev.count("first_rk_step_part1") == 1 and ev.count("relax_bdy_dry") == 1
    assert ev.count("advance_uv") == sum(rk_step_sizes(rk, cfg.dt, ns)[2] for rk in (1, 2, 3))
    assert ev.count("advance_uv") == 1 + ns // 2 + ns
    assert ev.count("calc_p_rho") == 3 + ev.count("advance_uv")
    assert ev.index("relax_bdy_dry") < ev.index("rk_addtend_dry") < ev.index("spec_bdy_dry") < ev.index("small_step_prep")
    assert ev.count("rk_phys_bc_dry_2") == 2 and ev.count("rk_update_scalar_pd:moist_old") == 1
    last_rk = max(i for i, e in enumerate(ev) if e == "calc_p_rho_phi")
    tail = ev[last_rk:]
    assert tail.index("set_phys_bc_dry_2") < tail.index("spec_bdy_final:u") < tail.index("set_w_surface")
    # iteration numbers seen by the stages: 0 for calc_p_rho before each
    # acoustic loop, 1..n inside it
    it_seen = [it for name, rk, it in calls[1:] if name == "calc_p_rho"]
    want = []
    for rk in (1, 2, 3):
        want += [0] + list(range(1, rk_step_sizes(rk, cfg.dt, ns)[2] + 1))
    assert it_seen == want
    # halo names are called and skipped (no entry)
    assert "halo_em_a_sub" in ev and "halo_em_a_sub" not in strict_solve_callbacks()
    assert not any(name in IDENTITY_STAGES for name, *_ in calls[1:])


def test_replaced_callback_runs():
    case = _case()
    cfg = case.config
    cbs = strict_solve_callbacks()
    seen = []
    inner = cbs["calc_p_rho_phi"]

    def mine(state):
        seen.append(state["rk_step"])
        return inner(state)

    cbs["calc_p_rho_phi"] = mine
    o = Oracle(cfg, case.arrays)
    solve_em(o, cbs)
    assert seen == [1, 2, 3]
    o.close()


def test_every_stage_name_runs_on_the_oracle():
    case = _case()
    o = Oracle(case.config, case.arrays)
    o.step(1, 0.0)
    for name in NUMERICAL_STAGES:
        if name == "spec_bdyupdate:w":
            continue
        o.stage(name, 2, 1)
    with pytest.raises(KeyError):
        o.stage("set_physical_bc3d:no_such_field", 2, 0)
    with pytest.raises(KeyError):
        o.stage("set_physical_bc3d:al:large", 2, 0)
    o.close()
