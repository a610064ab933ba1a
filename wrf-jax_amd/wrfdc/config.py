"""Scalar configuration of one dycore domain.

Mirrors the subset of the reference's ``config_flags`` / grid scalars that the
RK3 dycore reads (src/wrf_jax/runtime.py:3712-3870, grid dx/dy/dt/p_top and
the vertical-interpolation constants cf1..cf3, cfn, cfn1).  The options the
path supports are fixed: rk_ord 3, 5th/3rd-order advection, non-hydrostatic,
hypsometric_opt 2, diff_opt 2 with km_opt 4, specified lateral boundaries,
use_theta_m 0, damp_opt 0, no lid.  Unsupported options are rejected by
``DycoreConfig.validate`` with the reference's exception types.
"""

from __future__ import annotations

import ctypes
from dataclasses import asdict, dataclass, field, replace


@dataclass(frozen=True)
class DycoreConfig:
    nx: int
    ny: int
    nz: int
    ip: int
    nmoist: int
    nscalar: int
    spec_bdy_width: int = 5
    spec_zone: int = 1
    relax_zone: int = 4
    time_step_sound: int = 4
    pd_moist: int = 1
    pd_scalar: int = 1
    top_lid: int = 0
    pad_: int = 0
    dx: float = 3000.0
    dy: float = 3000.0
    dt: float = 18.0
    ptop: float = 5000.0
    p0: float = 100000.0
    t0: float = 300.0
    g: float = 9.81
    r_d: float = 287.0
    r_v: float = 461.6
    cp: float = 7.0 * 287.0 / 2.0
    cpovcv: float = 1.4
    cvpm: float = -1.0 / 1.4
    rvovrd: float = 461.6 / 287.0
    epssm: float = 0.1
    smdiv: float = 0.1
    emdiv: float = 0.01
    c_s: float = 0.25
    prandtl: float = 1.0 / 3.0
    cf1: float = 0.0
    cf2: float = 0.0
    cf3: float = 0.0
    cfn: float = 0.0
    cfn1: float = 0.0
    dtbc: float = 0.0
    reradius: float = 1.0 / 6370.0e03
    # options of the reference that this path implements one value of
    options: dict = field(default_factory=dict, compare=False, hash=False)

    def num_sound_steps(self) -> int:
        """WRF acoustic steps per dt (src/wrf_jax/dyn_em/solve_em.py:220-240)."""
        if self.time_step_sound:
            return int(self.time_step_sound)
        spacing = min(self.dx, self.dy)
        if spacing == 0.0:
            raise ValueError("solve_em requires non-zero grid spacing")
        return max(2 * (int(300.0 * self.dt / spacing - 0.01) + 1), 4)

    def validate(self) -> "DycoreConfig":
        supported = {
            "rk_ord": 3, "h_mom_adv_order": 5, "v_mom_adv_order": 3, "h_sca_adv_order": 5,
            "v_sca_adv_order": 3, "diff_opt": 2, "km_opt": 4, "damp_opt": 0, "w_damping": 0,
            "hypsometric_opt": 2, "use_theta_m": 0, "non_hydrostatic": True, "specified": True,
            "zadvect_implicit": 0, "periodic_x": False, "top_lid": False,
            "spec_exp": 0.0, "have_bcs_moist": False, "have_bcs_scalar": False, "constant_bc": False, "nested": False,
            "mp_physics": 0, "cu_physics": 0, "use_adaptive_time_step": False,
        }
        for key, value in self.options.items():
            if key in supported and value != supported[key]:
                if key == "rk_ord":
                    raise ValueError(f"solve_em unknown rk_order={value}")
                raise NotImplementedError(f"dycore option {key}={value!r} is not implemented (only {supported[key]!r})")
        if self.nx < 12 or self.ny < 12 or self.nz < 6:
            raise ValueError("domain too small for the 5th-order stencils and the 5-point boundary zone")
        if self.spec_bdy_width != self.spec_zone + self.relax_zone or self.relax_zone <= 1:
            raise ValueError("spec_bdy_width must equal spec_zone + relax_zone with relax_zone > 1")
        if self.ip < self.nx + 1:
            raise ValueError("i pitch smaller than the staggered extent")
        if self.nz > 150:
            raise ValueError("at most 150 vertical levels are supported (LDS-resident columns)")
        if (self.ny + 1) * (self.nz + 1) * self.ip >= 1 << 30:
            raise ValueError("a 3-D field must hold fewer than 2^30 elements (32-bit byte offsets)")
        return self

    def with_dtbc(self, dtbc: float) -> "DycoreConfig":
        return replace(self, dtbc=float(dtbc))

    def as_struct(self) -> "CConfig":
        d = asdict(self)
        d.pop("options")
        return CConfig(**d)


class CConfig(ctypes.Structure):
    """C layout shared by the oracle (OCfg) and the HIP library (wrfdc_config)."""

    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("ip", ctypes.c_int),
        ("nmoist", ctypes.c_int), ("nscalar", ctypes.c_int), ("spec_bdy_width", ctypes.c_int),
        ("spec_zone", ctypes.c_int), ("relax_zone", ctypes.c_int), ("time_step_sound", ctypes.c_int),
        ("pd_moist", ctypes.c_int), ("pd_scalar", ctypes.c_int), ("top_lid", ctypes.c_int),
        ("pad_", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dt", ctypes.c_double),
        ("ptop", ctypes.c_double), ("p0", ctypes.c_double), ("t0", ctypes.c_double), ("g", ctypes.c_double),
        ("r_d", ctypes.c_double), ("r_v", ctypes.c_double), ("cp", ctypes.c_double), ("cpovcv", ctypes.c_double),
        ("cvpm", ctypes.c_double), ("rvovrd", ctypes.c_double), ("epssm", ctypes.c_double),
        ("smdiv", ctypes.c_double), ("emdiv", ctypes.c_double), ("c_s", ctypes.c_double),
        ("prandtl", ctypes.c_double), ("cf1", ctypes.c_double), ("cf2", ctypes.c_double),
        ("cf3", ctypes.c_double), ("cfn", ctypes.c_double), ("cfn1", ctypes.c_double),
        ("dtbc", ctypes.c_double), ("reradius", ctypes.c_double),
    ]


__all__ = ["CConfig", "DycoreConfig"]
