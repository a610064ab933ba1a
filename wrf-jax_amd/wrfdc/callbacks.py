"""Stage callbacks by the reference's names and the solve_em driver that
calls them.

The reference advances a domain through ``solve_em``
(src/wrf_jax/dyn_em/solve_em.py:260-800), which calls named stages from a
dict of callbacks, ``StageFn = Callable[[dict], Mapping | None]``: each
takes the driver's state dict and returns updates to merge into it, and a
name missing from the dict is skipped (solve_em.py:171-188).
``strict_solve_callbacks()`` builds the dict the reference's runtime passes
(src/wrf_jax/runtime.py:16235-16301).

Here the state stays in HBM: every numerical stage of the dict runs the
library's HIP kernels on ``state["domain"]`` through ``wrfdc_stage`` and
returns None.  A caller may replace any entry by its own function (for
instance one that works on the torch views ``state["domain"].buffers``)
and keep the rest.  The physics stages are identities: the library
supports the configurations with physics off, where they write nothing the
dycore reads (advance_ppt with cu_physics 0, phy_prep_part2 whose outputs
feed microphysics only, after_all_rk_steps, the physics drivers of
first_rk_step_part1).

``solve_em(domain, callbacks, dtbc0)`` calls the dict in the order of
solve_em.py:470-790 for what the library supports: rk_ord 3, specified
lateral boundaries, non-hydrostatic, one tile, no physics, no adaptive time
step.  Halo and periodic-boundary exchanges are called by name as there and
find no entry (one tile on one GPU has no halos).  The domain is anything
with ``config``, ``stage(name, rk_step, iteration)`` and ``set_dtbc(dtbc)``:
a ``wrfdc.device.Domain``.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Callable, Mapping

StageFn = Callable[[dict], "Mapping[str, Any] | None"]

# stages that run the library's kernels, under the names solve_em calls
NUMERICAL_STAGES = (
    "initialize_moist_old",
    "set_physical_bc3d:moist_old",
    "rk_step_prep",
    "rk_phys_bc_dry_1",
    "set_physical_bc3d:rho",
    "set_physical_bc3d:al",
    "set_physical_bc3d:ph_2",
    "first_rk_step_part1",
    "first_rk_step_part2",
    "rk_tendency",
    "relax_bdy_dry",
    "rk_addtend_dry",
    "spec_bdy_dry",
    "dry_boundary_tendency",
    "small_step_prep",
    "calc_p_rho",
    "calc_coef_w",
    "set_physical_bc3d:ru_tend",
    "set_physical_bc3d:rv_tend",
    "set_physical_bc3d:p",
    "set_physical_bc3d:t_1",
    "set_physical_bc3d:t_save",
    "set_physical_bc2d:mu_1",
    "set_physical_bc2d:mu_2",
    "set_physical_bc2d:mudf",
    "advance_uv",
    "spec_bdyupdate:u",
    "spec_bdyupdate:v",
    "advance_mu_t",
    "spec_bdyupdate:t",
    "spec_bdyupdate:mu_2",
    "spec_bdyupdate:muts",
    "advance_w",
    "sumflux",
    "spec_bdyupdate_ph",
    "zero_grad_bdy",
    "spec_bdyupdate:w",
    "set_physical_bc3d:ph_2:small",
    "set_physical_bc3d:al:small",
    "set_physical_bc3d:p:small",
    "set_physical_bc2d:muts:small",
    "set_physical_bc2d:mu_2:small",
    "set_physical_bc2d:mudf:small",
    "calc_mu_uv_1",
    "small_step_finish",
    "set_physical_bc3d:ru_m",
    "set_physical_bc3d:rv_m",
    "set_physical_bc3d:ww_m",
    "set_physical_bc2d:mut",
    "set_physical_bc2d:muts",
    "rk_update_scalar_pd:moist_old",
    "rk_update_scalar_pd:scalar_old",
    "set_physical_bc3d:scalar_old",
    "moist_scalar_advance",
    "scalar_advance",
    "calc_p_rho_phi",
    "calc_p_rho_phi:final",
    "rk_phys_bc_dry_2",
    "set_physical_bc3d:moist",
    "set_physical_bc3d:scalar",
    "set_physical_bc3d:h_diabatic",
    "set_phys_bc_dry_2",
    "set_physical_bc3d:tke_1",
    "set_physical_bc3d:tke_2",
    "set_physical_bc3d:moist_qv",
    "spec_bdy_final:u",
    "spec_bdy_final:v",
    "spec_bdy_final:t",
    "spec_bdy_final:ph",
    "spec_bdy_final:mu",
    "spec_bdy_final:moist",
    "spec_bdy_final:moist_qv",
    "set_w_surface",
)

# identities with physics off (see the module docstring)
IDENTITY_STAGES = (
    "advance_ppt",
    "phy_prep_part2",
    "after_all_rk_steps",
)


def _library_stage(name: str) -> StageFn:
    def run(state: dict) -> None:
        # the stages before the RK loop (initialize_moist_old and its
        # boundary update) read no stage sizes
        state["domain"].stage(name, int(state.get("rk_step", 1)), int(state.get("iteration", 0) or 0))
        return None

    run.__name__ = "stage_" + name.replace(":", "_")
    run.__qualname__ = run.__name__
    return run


def _identity(_state: dict) -> None:
    return None


def strict_solve_callbacks() -> dict[str, StageFn]:
    """Stage name -> StageFn for every stage solve_em calls that the library
    implements (runtime.py:16235-16301)."""
    callbacks: dict[str, StageFn] = {name: _library_stage(name) for name in NUMERICAL_STAGES}
    callbacks.update({name: _identity for name in IDENTITY_STAGES})
    return callbacks


def rk_step_sizes(rk_step: int, dt: float, num_sound_steps: int) -> tuple[float, float, int]:
    """(dt_rk, dts_rk, number_of_small_timesteps) of an RK3 stage
    (solve_em.py:243-257)."""
    dts = dt / float(num_sound_steps)
    if rk_step == 1:
        return dt / 3.0, dt / 3.0, 1
    if rk_step == 2:
        return 0.5 * dt, dts, max(num_sound_steps // 2, 1)
    if rk_step == 3:
        return dt, dts, num_sound_steps
    raise ValueError(f"rk_step {rk_step} outside 1..3")


@dataclass
class SolveEmResult:
    rk_order: int
    num_sound_steps: int
    dts: float
    dtbc: float
    events: list = field(default_factory=list)


def solve_em(domain, callbacks: Mapping[str, StageFn] | None = None, dtbc0: float = 0.0,
             record_events: bool = True, state: dict | None = None) -> SolveEmResult:
    """One time step of ``domain`` by the named stages of ``callbacks``
    (default: strict_solve_callbacks()), in the order of
    solve_em.py:470-790.  dtbc0 is the boundary time before the step; the
    step runs at dtbc0 + dt (solve_em.py:455)."""
    cfg = domain.config
    cfg.validate()
    cbs = strict_solve_callbacks() if callbacks is None else callbacks
    events: list[str] = []
    st = {} if state is None else state
    dt = float(cfg.dt)
    dtbc = float(dtbc0) + dt
    domain.set_dtbc(dtbc)
    ns = cfg.num_sound_steps()
    st.update({"domain": domain, "config": cfg, "dtbc": dtbc, "iteration": 0})
    have_moist = cfg.nmoist >= 1
    have_scalar = cfg.nscalar >= 1

    def stage(name: str, **ctx) -> None:
        if record_events:
            events.append(name)
        if ctx:
            st.update(ctx)
        fn = cbs.get(name)
        if fn is None:
            return
        updates = fn(st)
        if updates:
            st.update(updates)

    stage("initialize_moist_old")
    if have_moist:
        stage("halo_em_moist_old_e_7_sub")
        stage("period_bdy_em_moist_old_sub")
        stage("set_physical_bc3d:moist_old")

    for rk in (1, 2, 3):
        dt_rk, dts_rk, nsmall = rk_step_sizes(rk, dt, ns)
        st.update({"rk_step": rk, "dt_rk": dt_rk, "dts_rk": dts_rk, "number_of_small_timesteps": nsmall, "iteration": 0})
        stage("rk_step_prep")
        stage("halo_em_a_sub")
        stage("period_bdy_em_a_sub")
        stage("rk_phys_bc_dry_1")
        stage("set_physical_bc3d:rho")
        stage("set_physical_bc3d:al")
        stage("set_physical_bc3d:ph_2")
        if rk == 1:
            stage("first_rk_step_part1")
            stage("first_rk_step_part2")
        stage("rk_tendency")
        if rk == 1:
            stage("relax_bdy_dry")
        stage("rk_addtend_dry")
        stage("spec_bdy_dry")
        stage("small_step_prep")
        stage("calc_p_rho", iteration=0)
        stage("calc_coef_w")
        stage("halo_em_b_sub")
        stage("period_bdy_em_b_sub")
        for name in ("set_physical_bc3d:ru_tend", "set_physical_bc3d:rv_tend", "set_physical_bc3d:ph_2",
                     "set_physical_bc3d:al", "set_physical_bc3d:p", "set_physical_bc3d:t_1", "set_physical_bc3d:t_save",
                     "set_physical_bc2d:mu_1", "set_physical_bc2d:mu_2", "set_physical_bc2d:mudf"):
            stage(name)
        for it in range(1, nsmall + 1):
            st["iteration"] = it
            stage("period_bdy_em_b_sub")
            stage("advance_uv")
            stage("spec_bdyupdate:u")
            stage("spec_bdyupdate:v")
            stage("halo_em_c_sub")
            stage("advance_mu_t")
            stage("spec_bdyupdate:t")
            stage("spec_bdyupdate:mu_2")
            stage("spec_bdyupdate:muts")
            stage("advance_w")
            stage("sumflux")
            stage("spec_bdyupdate_ph")
            stage("zero_grad_bdy")
            stage("calc_p_rho", iteration=it)
            stage("halo_em_c2_sub")
            stage("period_bdy_em_b3_sub")
            for name in ("set_physical_bc3d:ph_2:small", "set_physical_bc3d:al:small", "set_physical_bc3d:p:small",
                         "set_physical_bc2d:muts:small", "set_physical_bc2d:mu_2:small", "set_physical_bc2d:mudf:small"):
                stage(name)
        stage("calc_mu_uv_1")
        stage("small_step_finish")
        if rk == 3:
            for name in ("set_physical_bc3d:ru_m", "set_physical_bc3d:rv_m", "set_physical_bc3d:ww_m",
                         "set_physical_bc2d:mut", "set_physical_bc2d:muts"):
                stage(name)
        if have_moist and rk == 3 and cfg.pd_moist:
            stage("rk_update_scalar_pd:moist_old")
            stage("halo_em_moist_old_e_7_sub")
            stage("period_bdy_em_moist_old_sub")
            stage("set_physical_bc3d:moist_old")
        if have_scalar and rk == 3 and cfg.pd_scalar:
            stage("rk_update_scalar_pd:scalar_old")
            stage("halo_em_scalar_old_e_7_sub")
            stage("period_bdy_em_scalar_old_sub")
            stage("set_physical_bc3d:scalar_old")
        if have_moist:
            stage("moist_scalar_advance")
        if have_scalar:
            stage("scalar_advance")
        stage("calc_p_rho_phi")
        if rk < 3:
            stage("halo_em_d2_5_sub")
            stage("period_bdy_em_d_sub")
            stage("period_bdy_em_moist2_sub")
            stage("rk_phys_bc_dry_2")
            if have_moist:
                stage("set_physical_bc3d:moist")
            if have_scalar:
                stage("set_physical_bc3d:scalar")
            stage("halo_em_moist_e_5_sub")
            if have_scalar:
                stage("halo_em_scalar_e_5_sub")

    stage("advance_ppt")
    stage("phy_prep_part2")
    stage("set_phys_bc_dry_2")
    stage("set_physical_bc3d:tke_1")
    stage("set_physical_bc3d:tke_2")
    if have_moist:
        stage("set_physical_bc3d:moist_qv")
    for name in ("spec_bdy_final:u", "spec_bdy_final:v", "spec_bdy_final:t", "spec_bdy_final:ph", "spec_bdy_final:mu"):
        stage(name)
    if have_moist:
        stage("spec_bdy_final:moist")
    stage("halo_em_c_sub")
    stage("period_bdy_em_e_sub")
    stage("set_w_surface")
    stage("after_all_rk_steps")
    return SolveEmResult(rk_order=3, num_sound_steps=ns, dts=dt / ns, dtbc=dtbc, events=events)


__all__ = ["IDENTITY_STAGES", "NUMERICAL_STAGES", "SolveEmResult", "StageFn", "rk_step_sizes", "solve_em",
           "strict_solve_callbacks"]
