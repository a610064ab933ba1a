"""Algorithmic HBM byte models of the dycore (version 1).

Step model: compulsory traffic of the minimal fused schedule, 4 B per mass
point per 3-D field pass (2-D/1-D arrays and the +1 staggering ignored):

    B_alg(step) = 4 * N_mass * (274 + 52 * (1 + 1.5 * ns) + 12 * N_s)

with ns acoustic substeps per dt and N_s advected species; at ns = 4 this
is 4 * N_mass * (638 + 12 * N_s).  Per RK stage: the tendency pass
(rk_step_prep .. rk_tendency, diffusion at rk1) 29/27/27 field passes, the
acoustic prologue (small_step_prep + calc_p_rho + calc_coef_w) 38/33/33, per
acoustic substep 14 (advance_uv) + 38 (advance_mu_t + advance_w + sumflux +
calc_p_rho), the epilogue (small_step_finish) 18, species 3 shared + 4 per
species, the equation of state 8.

Kernel model: the distinct 3-D fields a kernel must read and write once
each (a read-modify-write field counts twice), 4 B * N_mass per field.
"""

from __future__ import annotations

import os
import re
from dataclasses import dataclass

MODEL_VERSION = 1


def step_passes(n_species: int, ns: int = 4) -> float:
    return 274 + 52 * (1 + 1.5 * ns) + 12 * n_species


def step_bytes(n_mass: int, n_species: int, ns: int = 4) -> float:
    return 4.0 * n_mass * step_passes(n_species, ns)


@dataclass(frozen=True)
class KernelModel:
    name: str          # HIP kernel symbol
    stage: str         # wrfdc_stage entry that launches exactly this kernel ("" if none)
    rk_step: int
    iteration: int
    reads: tuple
    writes: tuple
    per_species: int = 0  # further passes per advected species (moist + scalar)
    per_moist: int = 0    # further passes per moist species
    per_scalar_diffused: int = 0  # further passes per diffused scalar (theta + every species)

    def passes(self, nmoist: int = 0, nscalar: int = 0) -> int:
        return (len(self.reads) + len(self.writes) + self.per_species * (nmoist + nscalar) + self.per_moist * nmoist +
                self.per_scalar_diffused * (1 + nmoist + nscalar))

    def bytes(self, cfg) -> float:
        return 4.0 * cfg.nx * cfg.ny * cfg.nz * self.passes(cfg.nmoist, cfg.nscalar)


KERNELS = {
    "k_acoustic_col": KernelModel(
        "k_acoustic_col", "acoustic_column", 3, 2,
        reads=("u_2", "u_save", "v_2", "v_save", "ww", "ww_1", "t_2", "t_save", "t_tend", "ph_save", "phb", "ph_tend",
               "w_2", "ph_2", "c2a", "cqw", "alt", "rw_tend", "a", "alpha", "gamma", "ru_m", "rv_m", "ww_m", "pm1"),
        writes=("ww", "t_2", "t_2ave", "w_2", "ph_2", "ru_m", "rv_m", "ww_m", "al", "p", "pm1")),
    "k_uv_substep": KernelModel(
        "k_uv_substep", "acoustic_uv", 3, 2,
        reads=("u_2", "v_2", "ru_tend", "rv_tend", "cqu", "cqv", "p", "pb", "ph_2", "alt", "al", "php"),
        writes=("u_2", "v_2")),
    "k_advance_w": KernelModel(
        "k_advance_w", "advance_w", 3, 1,
        reads=("w_2", "rw_tend", "ww", "t_2", "t_2ave", "t_save", "ph_save", "phb", "ph_tend", "ph_2", "c2a", "cqw",
               "alt", "a", "alpha", "gamma"),
        writes=("t_2ave", "w_2", "ph_2")),
    "k_advance_mu_t": KernelModel(
        "k_advance_mu_t", "advance_mu_t", 3, 1,
        reads=("u_2", "u_save", "v_2", "v_save", "ww", "ww_1", "t_2", "t_save", "t_tend"),
        writes=("ww", "t_2", "t_2ave")),
    "k_advect_u": KernelModel("k_advect_u", "advect_u", 2, 0, reads=("u_2", "ru", "rv", "ww", "ru_tend"), writes=("ru_tend",)),
    "k_advect_scalar": KernelModel("k_advect_scalar", "advect_scalar", 2, 0, reads=("t_2", "ru", "rv", "ww", "t_tend"),
                                   writes=("t_tend",)),
    "k_calc_p_rho": KernelModel("k_calc_p_rho", "calc_p_rho", 3, 1, reads=("alt", "t_save", "ph_2", "t_2", "c2a", "pm1"),
                                writes=("al", "p", "pm1")),
    "k_rhs_ph": KernelModel("k_rhs_ph", "rhs_ph", 2, 0, reads=("u_2", "v_2", "ww", "ph_2", "phb", "w_2", "ph_tend"),
                            writes=("ph_tend",)),
    # kernels of the step's fused path (a stage name only where one launches
    # the kernel alone); the species kernels carry every species of a stage
    # q in, q_old, tendency, q out; at rk 2 also q_old' and the cleared
    # tendency of every PD species (the last stage's rk_update_scalar_pd):
    # 5 per species is the mean of the step's two launches with all species PD
    "k_species_update": KernelModel("k_species_update", "", 2, 0, reads=("ru_m", "rv_m", "ww_m"), writes=(),
                                    per_species=5),
    "k_species_pd_prep": KernelModel("k_species_pd_prep", "", 3, 0, reads=(), writes=(), per_species=4),  # q_old, tend r/w
    "k_species_pd_scale": KernelModel("k_species_pd_scale", "", 3, 0, reads=("ru_m", "rv_m", "ww_m"), writes=(),
                                      per_species=3),  # q in, q_old, factor out
    "k_species_pd_update": KernelModel("k_species_pd_update", "", 3, 0, reads=("ru_m", "rv_m", "ww_m"), writes=(),
                                       per_species=4),  # q in, q_old, factor, q out
    "k_advect_u": KernelModel("k_advect_u", "advect_u", 2, 0, reads=("u_2", "ru", "rv", "ww", "ru_tend"), writes=("ru_tend",)),
    "k_advect_u<true>": KernelModel("k_advect_u<true>", "", 2, 0,
                                    reads=("u_2", "ru", "rv", "ww", "ph_2", "php", "alt", "al", "p", "pb", "v_2", "w_2", "ru_tendf",
                                           "u_save"), writes=("ru_tend",)),
    "k_advect_v<true>": KernelModel("k_advect_v<true>", "", 2, 0,
                                    reads=("v_2", "ru", "rv", "ww", "ph_2", "php", "alt", "al", "p", "pb", "u_2", "w_2", "rv_tendf",
                                           "v_save"), writes=("rv_tend",)),
    "k_advect_w<true>": KernelModel("k_advect_w<true>", "", 2, 0,
                                    reads=("w_2", "ru", "rv", "ww", "t_2", "t_save", "alt", "al", "p", "pb", "cqw", "u_2", "v_2",
                                           "rw_tendf", "w_save"), writes=("rw_tend", "cqw")),
    "k_advect_scalar<true>": KernelModel("k_advect_scalar<true>", "", 2, 0, reads=("t_2", "ru", "rv", "ww", "t_tendf", "t_save"),
                                         writes=("t_tend",)),
    "k_tend_ph": KernelModel("k_tend_ph", "", 2, 0, reads=("u_2", "v_2", "ww", "ph_2", "phb", "w_2", "ph_tendf", "ph_save"),
                             writes=("ph_tend",)),
    "k_rk_prep3": KernelModel("k_rk_prep3", "", 2, 0, reads=("u_2", "v_2", "w_2", "al", "alb", "phb", "ph_2"),
                              writes=("ru", "rv", "rw", "alt", "php", "cqu", "cqv", "cqw"), per_moist=1),
    "k_ssp_3d<false>": KernelModel("k_ssp_3d<false>", "", 2, 0,
                                   reads=("pb", "p", "alt", "u_2", "v_2", "t_2", "w_2", "ph_2", "ww", "u_1", "v_1", "t_1", "w_1",
                                          "ph_1"),
                                   writes=("c2a", "u_save", "u_2", "v_save", "v_2", "t_save", "t_2", "w_save", "w_2", "ph_save",
                                           "ph_2", "ww_1")),
    "k_ssf<false>": KernelModel("k_ssf<false>", "", 2, 0,
                                reads=("u_2", "u_save", "v_2", "v_save", "w_2", "w_save", "ph_2", "ph_save", "ww", "ww_1", "t_2",
                                       "t_save"), writes=("u_2", "v_2", "w_2", "ph_2", "ww", "t_2")),
    "k_calc_p_rho_phi": KernelModel("k_calc_p_rho_phi", "calc_p_rho_phi", 2, 0, reads=("ph_2", "phb", "alb", "pb", "t_2"),
                                    writes=("al", "p"), per_moist=0),
    "k_hdiff_s_faces": KernelModel("k_hdiff_s_faces", "", 1, 0, reads=("xkhh", "rho", "zx", "zy", "rdzw"), writes=(),
                                   per_scalar_diffused=3),  # q in, H1 and H2 out
    "k_hdiff_s_tend": KernelModel("k_hdiff_s_tend", "", 1, 0, reads=("zx", "zy", "rdzw"), writes=(),
                                  per_scalar_diffused=4),  # H1, H2 in, tendency r/w
}

DOMINANT = "k_acoustic_col"


def dominant_kernel(cfg) -> KernelModel:
    return KERNELS[DOMINANT]


def measured_traffic(kernel: str, workload: str, path: str | None = None):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (already scaled by the copy-kernel calibration), or None."""
    if path is None:
        repo = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
        path = os.path.join(repo, "profiles", f"traffic_{workload}.md")
    if not os.path.exists(path):
        return None
    # a row of the raw counter table: kernel | fetch MiB | write MiB |
    # traffic bytes | dispatches (the symbol may carry template arguments,
    # k_acoustic_col<8>); rows of the other tables have more columns
    num = r"\s*([0-9.eE+]+)\s*\|"
    pat = re.compile(rf"^\|\s*{re.escape(kernel)}(<[^>|]*>)?\s*\|{num}{num}{num}{num}$")
    with open(path) as fh:
        for line in fh:
            m = pat.match(line.strip())
            if m:
                return float(m.group(4))
    return None


__all__ = ["KERNELS", "KernelModel", "MODEL_VERSION", "dominant_kernel", "measured_traffic", "step_bytes", "step_passes"]
