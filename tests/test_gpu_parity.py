"""GPU parity: HIP kernels (through the C ABI) vs the CPU oracle.

Every test uploads the same seeded state to the GPU and to the oracle, runs
the same stage or step on both and compares every registry field.  Bars
(FP32 vs FP32 oracle): relative L2 <= 1e-5 per field after one stage or one
RK3 step, NRMSE <= 1e-3 per prognostic field after 100 steps.
"""

import numpy as np
import pytest

from parity import NRMSE_LONG, REL_L2_STEP, STEP_DIAGNOSTICS, compare, nrmse, rel_l2

pytestmark = pytest.mark.gpu


def _case(nx, ny, nz, **kw):
    from wrfdc.cases import make_case

    return make_case(nx, ny, nz, **kw)


@pytest.fixture(scope="module")
def stepped_state():
    """A state after one oracle step: every work array holds realistic values."""
    from oracle import Oracle

    case = _case(28, 24, 12, nmoist=6, nscalar=2, hydrometeors=True, hill_height=500.0)
    o = Oracle(case.config, case.arrays)
    o.step(1, 0.0)
    o.set_dtbc(2 * case.config.dt)
    state = o.copy()
    o.close()
    return case, state


@pytest.fixture(scope="module")
def domain_for(gpu):
    from wrfdc.device import Domain

    cache = {}

    def get(config):
        key = config
        if key not in cache:
            cache[key] = Domain(config, gpu)
        return cache[key]

    yield get
    for d in cache.values():
        d.close()


STAGES = [
    ("rk_step_prep", 1, 0), ("first_rk_step_part1", 1, 0), ("compute_diff_metrics", 1, 0),
    ("cal_deform_and_div", 1, 0), ("smag2d_km", 1, 0), ("calculate_n2", 1, 0), ("calculate_km_kh", 1, 0),
    ("horizontal_diffusion_2", 1, 0),
    ("first_rk_step_part2", 1, 0), ("advect_u", 2, 0), ("advect_v", 2, 0), ("advect_w", 2, 0),
    ("advect_scalar", 2, 0), ("advect_scalar_pd", 3, 0), ("rhs_ph", 2, 0),
    ("horizontal_pressure_gradient", 2, 0), ("pg_buoy_w", 2, 0), ("rk_tendency", 2, 0),
    ("dry_boundary_tendency", 1, 0), ("dry_boundary_tendency", 2, 0), ("small_step_prep", 1, 0),
    ("small_step_prep", 2, 0), ("calc_p_rho", 2, 0), ("calc_p_rho", 2, 1), ("calc_coef_w", 2, 0),
    ("advance_uv", 2, 1), ("advance_mu_t", 2, 1), ("advance_w", 2, 1), ("sumflux", 2, 1),
    ("sumflux", 2, 2), ("sumflux", 1, 1), ("small_step_finish", 2, 0), ("small_step_finish", 3, 0),
    ("calc_p_rho_phi", 2, 0), ("acoustic_substep", 1, 1), ("acoustic_substep", 2, 1), ("acoustic_substep", 2, 2),
    ("acoustic_substep", 3, 1), ("acoustic_substep", 3, 3), ("rk_stage", 1, 0), ("rk_stage", 2, 0), ("rk_stage", 3, 0),
    # the remaining stages solve_em calls by name (src/wrf_jax/dyn_em/solve_em.py:470-790)
    ("initialize_moist_old", 1, 0), ("rk_phys_bc_dry_1", 2, 0), ("relax_bdy_dry", 1, 0), ("rk_addtend_dry", 1, 0),
    ("rk_addtend_dry", 3, 0), ("spec_bdy_dry", 2, 0), ("spec_bdyupdate:u", 2, 1), ("spec_bdyupdate:v", 2, 1),
    ("spec_bdyupdate:t", 3, 2), ("spec_bdyupdate:mu_2", 2, 1), ("spec_bdyupdate:muts", 2, 1), ("spec_bdyupdate_ph", 2, 1),
    ("zero_grad_bdy", 2, 1), ("calc_mu_uv_1", 2, 0), ("rk_update_scalar_pd:moist_old", 3, 0),
    ("rk_update_scalar_pd:scalar_old", 3, 0), ("moist_scalar_advance", 1, 0), ("moist_scalar_advance", 2, 0),
    ("moist_scalar_advance", 3, 0), ("scalar_advance", 1, 0), ("scalar_advance", 3, 0), ("calc_p_rho_phi:final", 3, 0),
    ("rk_phys_bc_dry_2", 2, 0), ("set_phys_bc_dry_2", 3, 0), ("spec_bdy_final:u", 3, 0), ("spec_bdy_final:v", 3, 0),
    ("spec_bdy_final:t", 3, 0), ("spec_bdy_final:ph", 3, 0), ("spec_bdy_final:mu", 3, 0), ("spec_bdy_final:moist", 3, 0),
    ("set_w_surface", 3, 0), ("set_physical_bc3d:ru_tend", 2, 0), ("set_physical_bc3d:t_save", 2, 0),
    ("set_physical_bc3d:moist", 2, 0), ("set_physical_bc3d:scalar_old", 3, 0), ("set_physical_bc3d:p:small", 2, 1),
    ("set_physical_bc3d:ww_m", 3, 0), ("set_physical_bc2d:mudf:small", 2, 1), ("set_physical_bc2d:muts", 3, 0),
]


@pytest.mark.parametrize("name,rk,it", STAGES, ids=[f"{n}-rk{r}-it{i}" for n, r, i in STAGES])
def test_stage_parity(stepped_state, domain_for, name, rk, it):
    from oracle import Oracle

    case, state = stepped_state
    print(f"[stage parity] {name} rk={rk} it={it} shape={case.config.nx}x{case.config.ny}x{case.config.nz}", flush=True)
    dom = domain_for(case.config.with_dtbc(2 * case.config.dt))
    o = Oracle(case.config.with_dtbc(2 * case.config.dt), state)
    o.stage(name, rk, it)
    dom.upload(state)
    dom.stage(name, rk, it)
    got = dom.download()
    skip = STEP_DIAGNOSTICS if name in ("rk_stage", "step") else ()
    bad = compare(got, o.arrays, tol=REL_L2_STEP, skip=skip)
    o.close()
    assert not bad, f"{name}: fields above rel-L2 {REL_L2_STEP}: {bad}"


# 65 and 129 columns: the i ranges end on, or one past, a 64-point tile edge
SHAPES = [(28, 24, 12), (13, 17, 7), (40, 33, 20), (65, 21, 8), (129, 13, 7)]


SPECIES_MIXES = [(1, 0, 1), (1, 0, 0), (3, 1, 0), (2, 3, 1)]


@pytest.mark.parametrize("nmoist,nscalar,pd", SPECIES_MIXES, ids=[f"m{a}s{b}pd{c}" for a, b, c in SPECIES_MIXES])
def test_one_step_parity_species_mixes(domain_for, nmoist, nscalar, pd):
    """One RK3 step at 28x24x12 with other species counts and the
    positive-definite advection on or off: the species kernels' lists, the
    second stage's fold of the last stage's q_old update (PD only) and the
    diffusion of every scalar follow the configuration."""
    import dataclasses

    from oracle import Oracle

    print(f"[step parity] 28x24x12, {nmoist} moist + {nscalar} scalars, PD {pd}", flush=True)
    case = _case(28, 24, 12, nmoist=nmoist, nscalar=nscalar, hydrometeors=nmoist > 1, hill_height=500.0)
    cfg = dataclasses.replace(case.config, pd_moist=pd, pd_scalar=pd)
    dom = domain_for(cfg)
    o = Oracle(cfg, case.arrays)
    o.step(1, 0.0)
    dom.upload(case.arrays)
    dom.step(1, 0.0)
    got = dom.download()
    bad = compare(got, o.arrays, tol=REL_L2_STEP, skip=STEP_DIAGNOSTICS)
    o.close()
    assert not bad, f"fields above rel-L2 {REL_L2_STEP} after one RK3 step: {bad}"


@pytest.mark.parametrize("levels", [1, 3, 16, 32])
def test_face_levels_do_not_change_results(gpu, levels):
    """The face-block kernels' level-group size (wrfdc_set_tile) is a launch
    shape only: two RK3 steps at 30x26x13 (13 levels: ragged for every
    group size above 1) give bit-identical fields to the default shapes, and
    out-of-range settings are refused."""
    from wrfdc.device import Domain

    print(f"[tile] face_levels={levels} vs default, 30x26x13, 2 steps", flush=True)
    case = _case(30, 26, 13, nmoist=3, nscalar=1, hydrometeors=True, hill_height=500.0)
    ref, dom = Domain(case.config, gpu), Domain(case.config, gpu)
    try:
        for d in (ref, dom):
            d.upload(case.arrays)
        dom.set_tile("face_levels", levels)
        ref.step(2, 0.0)
        dom.step(2, 0.0)
        a, b = ref.download(), dom.download()
        diff = [n for n in a if not np.array_equal(a[n].view(np.uint32), b[n].view(np.uint32))]
        assert not diff, f"fields that differ with face_levels={levels}: {diff}"
        for bad in (-1, 33):
            with pytest.raises(ValueError, match="face_levels"):
                dom.set_tile("face_levels", bad)
        with pytest.raises(ValueError, match="unknown tile setting"):
            dom.set_tile("no_such_setting", 4)
    finally:
        ref.close()
        dom.close()


def _with_map_factors(case):
    """The case with smooth non-unit map factors (a conformal-projection-like
    field, 0.95..1.05), each staggering sampled at its own points and
    msfvx_inv = 1 / msfvx in FP32: every division by a map factor then has a
    divisor other than 1."""
    A = case.arrays
    jde, ip = A["msftx"].shape
    cfg = case.config
    nx, ny = cfg.nx, cfg.ny

    def field(di, dj):
        ii = np.arange(ip, dtype=np.float64)[None, :] + di
        jj = np.arange(jde, dtype=np.float64)[:, None] + dj
        return (1.0 + 0.05 * np.sin(2.0 * np.pi * ii / nx) * np.cos(np.pi * jj / ny) + 0.01 * jj / ny).astype(np.float32)

    for key, (di, dj) in {"msftx": (0.0, 0.0), "msfty": (0.0, 0.0), "msfux": (-0.5, 0.0), "msfuy": (-0.5, 0.0),
                          "msfvx": (0.0, -0.5), "msfvy": (0.0, -0.5)}.items():
        A[key][...] = field(di, dj) * (1.0 if key.endswith("x") else np.float32(1.002))
    A["msfvx_inv"][...] = np.float32(1.0) / A["msfvx"]
    return case


def test_one_step_parity_map_factors(domain_for):
    """One RK3 step with non-unit map factors against the oracle: the kernels
    divide by map factors through per-column reciprocals (div_rn), the
    oracle with IEEE divisions; every field must agree."""
    from oracle import Oracle

    print("[step parity] 40x33x20 with map factors 0.95..1.05", flush=True)
    case = _with_map_factors(_case(40, 33, 20, nmoist=6, nscalar=2, hydrometeors=True, hill_height=500.0))
    dom = domain_for(case.config)
    o = Oracle(case.config, case.arrays)
    o.step(1, 0.0)
    dom.upload(case.arrays)
    dom.step(1, 0.0)
    got = dom.download()
    diff = {n: int(np.count_nonzero(np.asarray(got[n]) != np.asarray(o.arrays[n]))) for n in ("u_2", "v_2", "w_2", "t_2", "ph_2", "p", "al")}
    print(f"[step parity] differing elements: {diff}", flush=True)
    bad = compare(got, o.arrays, tol=REL_L2_STEP, skip=STEP_DIAGNOSTICS)
    o.close()
    assert not bad, f"fields above rel-L2 {REL_L2_STEP} after one RK3 step: {bad}"


@pytest.mark.parametrize("shape", SHAPES, ids=[f"{a}x{b}x{c}" for a, b, c in SHAPES])
def test_one_step_parity(domain_for, shape):
    from oracle import Oracle

    nx, ny, nz = shape
    print(f"[step parity] shape={shape}", flush=True)
    case = _case(nx, ny, nz, nmoist=6, nscalar=2, hydrometeors=True, hill_height=500.0)
    dom = domain_for(case.config)
    o = Oracle(case.config, case.arrays)
    o.step(1, 0.0)
    dom.upload(case.arrays)
    dom.step(1, 0.0)
    got = dom.download()
    bad = compare(got, o.arrays, tol=REL_L2_STEP, skip=STEP_DIAGNOSTICS)
    o.close()
    assert not bad, f"fields above rel-L2 {REL_L2_STEP} after one RK3 step: {bad}"


def test_boundary_refresh(domain_for):
    """Time-varying lateral forcing across a boundary-record change, as the
    em_real loop drives it (src/wrf_jax/trajectory.py:1504-1526 loads the next
    wrfbdy record and resets dtbc to 0): three steps relaxing towards edge
    values that drift with nonzero tendencies, then the next record (edge
    values advanced over the interval, new tendencies) with dtbc restarting
    at 0, three more steps.  Every registry field against the oracle."""
    from oracle import Oracle
    from wrfdc.fields import FIELD_SPECS

    print("[boundary refresh] shape=(28, 24, 12), 3 + 3 steps", flush=True)
    case = _case(28, 24, 12, nmoist=6, nscalar=2, hydrometeors=True, hill_height=500.0)
    A = {k: v.copy() for k, v in case.arrays.items()}
    dt = case.config.dt
    rng = np.random.default_rng(20260)
    bnames = [n for n, kind in FIELD_SPECS if kind in "XYxy" and "_b" in n and "_bt" not in n]
    assert bnames
    for n in bnames:  # record 1: the edge state drifts by up to +-5 % over 600 s
        t = n.replace("_b", "_bt", 1)
        A[t] = (A[n] * rng.uniform(-0.05, 0.05, A[n].shape) / 600.0).astype(np.float32)
    assert any(np.abs(A[n.replace("_b", "_bt", 1)]).max() > 0 for n in bnames)
    dom = domain_for(case.config)
    o = Oracle(case.config, A)
    dom.upload(A)
    o.step(3, 0.0)
    dom.step(3, 0.0)
    # record 2: edge values advanced over the three steps, halved tendencies
    nxt = {}
    for n in bnames:
        t = n.replace("_b", "_bt", 1)
        nxt[n] = (A[n] + A[t] * np.float32(3 * dt)).astype(np.float32)
        nxt[t] = (A[t] * np.float32(0.5)).astype(np.float32)
    dom.upload(nxt)
    for k, v in nxt.items():
        o.arrays[k][...] = v
    o.step(3, 0.0)
    dom.step(3, 0.0)
    got = dom.download()
    bad = compare(got, o.arrays, tol=REL_L2_STEP, skip=STEP_DIAGNOSTICS)
    o.close()
    assert not bad, f"fields above rel-L2 {REL_L2_STEP} after a boundary refresh: {bad}"


def test_hundred_steps_nrmse(domain_for):
    from oracle import Oracle

    print("[100-step NRMSE] shape=(32, 28, 12)", flush=True)
    case = _case(32, 28, 12, nmoist=6, nscalar=2, hydrometeors=True, hill_height=500.0)
    dom = domain_for(case.config)
    o = Oracle(case.config, case.arrays)
    o.step(100, 0.0)
    dom.upload(case.arrays)
    dom.step(100, 0.0)
    got = dom.download()
    errs = {}
    for n in ("u_2", "v_2", "w_2", "t_2", "ph_2", "mu_2", "p", "al", "moist", "scalar"):
        errs[n] = nrmse(got[n], o.arrays[n])
    o.close()
    print("NRMSE after 100 steps:", {k: f"{v:.2e}" for k, v in errs.items()})
    assert all(np.isfinite(v) and v <= NRMSE_LONG for v in errs.values()), errs


def test_checksums_and_nonfinite(domain_for):
    case = _case(28, 24, 12, nmoist=6, nscalar=2, hydrometeors=True)
    dom = domain_for(case.config)
    dom.upload(case.arrays)
    ck = dom.checksums()
    A = case.arrays
    mass = (A["mu_2"][:-1, :case.config.nx].astype(np.float64) + A["mub"][:-1, :case.config.nx]).sum()
    theta = A["t_2"][:-1, :-1, :case.config.nx].astype(np.float64).sum()
    assert rel_l2(ck[0], mass) < 1e-12 and rel_l2(ck[1], theta) < 1e-12 and ck[3] == 0
    bad = dict(A)
    bad["t_2"] = A["t_2"].copy()
    bad["t_2"][3, 2, 5] = np.nan
    dom.upload({"t_2": bad["t_2"]})
    with pytest.raises(FloatingPointError):
        dom.checksums()
