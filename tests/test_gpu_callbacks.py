"""GPU: one RK3 step driven through the reference-named stage callbacks.

``solve_em`` calls ``strict_solve_callbacks()`` in the order of
src/wrf_jax/dyn_em/solve_em.py:470-790; every numerical stage runs the HIP
kernels through wrfdc_stage.  After one and two steps, with time-varying
lateral forcing, every registry field matches the FP32 oracle's step
(relative L2 1e-5) and the library's fused step (twice that: both
sit within 1e-5 of the oracle).
"""

import numpy as np
import pytest

from parity import REL_L2_STEP, STEP_DIAGNOSTICS, compare

pytestmark = pytest.mark.gpu


def _case():
    from wrfdc.cases import make_case

    case = make_case(28, 24, 12, nmoist=6, nscalar=2, hydrometeors=True, hill_height=500.0)
    rng = np.random.default_rng(31)
    for name in ("u_btxs", "u_btye", "v_btxe", "v_btys", "t_btxs", "t_btye", "mu_btxe", "ph_btys", "qv_btxs"):
        case.arrays[name][:] = (1e-3 * rng.standard_normal(case.arrays[name].shape)).astype(np.float32)
    return case


def test_solve_em_through_callbacks(gpu):
    from oracle import Oracle
    from wrfdc.callbacks import solve_em, strict_solve_callbacks
    from wrfdc.device import Domain

    case = _case()
    cfg = case.config
    print(f"[callbacks] solve_em via strict_solve_callbacks, {cfg.nx}x{cfg.ny}x{cfg.nz}, 2 steps", flush=True)
    dom = Domain(cfg, gpu)
    dom.upload(case.arrays)
    fused = Domain(cfg, gpu)
    fused.upload(case.arrays)
    o = Oracle(cfg, case.arrays)
    cbs = strict_solve_callbacks()
    dtbc0 = 1.5 * cfg.dt
    for n in range(2):
        res = solve_em(dom, cbs, dtbc0=dtbc0 + n * cfg.dt)
        assert res.events.count("advance_uv") == 1 + res.num_sound_steps // 2 + res.num_sound_steps
        o.step(1, dtbc0 + n * cfg.dt)
        fused.step(1, dtbc0 + n * cfg.dt)
        got = dom.download()
        bad = compare(got, o.arrays, tol=REL_L2_STEP, skip=STEP_DIAGNOSTICS)
        assert not bad, (n, "vs oracle", bad)
        bad = compare(got, fused.download(), tol=2 * REL_L2_STEP, skip=STEP_DIAGNOSTICS)
        assert not bad, (n, "vs fused step", bad)
    o.close()
    dom.close()
    fused.close()


def test_unsupported_stage_is_refused(gpu):
    from wrfdc.device import Domain

    case = _case()
    dom = Domain(case.config, gpu)
    dom.upload(case.arrays)
    for name in ("spec_bdyupdate:w", "set_physical_bc3d:no_such_field", "no_such_stage"):
        with pytest.raises(Exception) as e:
            dom.stage(name, 2, 1)
        print("[callbacks] refused:", name, "->", e.value, flush=True)
    dom.close()
