"""Field registry of the dycore state held on the device.

Every array the RK3 step touches has a name and a kind.  The names follow the
reference's state dictionary (``u_2``, ``mu_tend``, ``ph_btxs``, ...; see
src/wrf_jax/core/field_sets.py:10-205 and src/wrf_jax/runtime.py:4312-4431),
with two departures that are layout, not semantics:

* species are held as species-major stacks (``moist[s]``, ``scalar[s]``)
  without the reference's unused slot 0 (PARAM_FIRST_SCALAR = 2), so moist
  species 0 is QV;
* QVAPOR lateral-boundary arrays are named ``qv_b*`` instead of ``qvapor_b*``.

Storage layout of one domain (all float32, C order):

====  =====================================  ==============================
kind  shape                                  index
====  =====================================  ==============================
'3'   (jde, kde, ip)                         [j-1, k-1, i-1]
'2'   (jde, ip)                              [j-1, i-1]
'1'   (kde,)                                 [k-1]
'M'   (nmoist, jde, kde, ip)                 moist species stack
'S'   (nscalar, jde, kde, ip)                scalar species stack
'X'   (bw, kde, jde)   x-edge boundary       [b-1, k-1, j-1]
'Y'   (bw, kde, ip)    y-edge boundary       [b-1, k-1, i-1]
'x'   (bw, jde)        x-edge boundary (2-D) [b-1, j-1]
'y'   (bw, ip)         y-edge boundary (2-D) [b-1, i-1]
====  =====================================  ==============================

with ide = nx+1, jde = ny+1, kde = nz+1 (staggered extents, as in the
reference) and an i pitch ``ip >= ide`` padded for aligned rows.
"""

from __future__ import annotations

from dataclasses import dataclass

FIELD_SPECS: tuple[tuple[str, str], ...] = (
    ("u_1", "3"), ("u_2", "3"), ("v_1", "3"), ("v_2", "3"), ("w_1", "3"), ("w_2", "3"),
    ("t_1", "3"), ("t_2", "3"), ("ph_1", "3"), ("ph_2", "3"), ("p", "3"), ("al", "3"),
    ("pb", "3"), ("phb", "3"), ("alb", "3"), ("ru", "3"), ("rv", "3"), ("rw", "3"),
    ("ww", "3"), ("php", "3"), ("alt", "3"), ("cqu", "3"), ("cqv", "3"), ("cqw", "3"),
    ("c2a", "3"), ("pm1", "3"), ("a", "3"), ("alpha", "3"), ("gamma", "3"),
    ("ru_tend", "3"), ("rv_tend", "3"), ("rw_tend", "3"), ("t_tend", "3"),
    ("ph_tend", "3"), ("ru_tendf", "3"), ("rv_tendf", "3"), ("rw_tendf", "3"),
    ("t_tendf", "3"), ("ph_tendf", "3"), ("u_save", "3"), ("v_save", "3"),
    ("w_save", "3"), ("t_save", "3"), ("ph_save", "3"), ("ww_1", "3"), ("t_2ave", "3"),
    ("ru_m", "3"), ("rv_m", "3"), ("ww_m", "3"), ("h_diabatic", "3"),
    ("advect_tend", "3"), ("rho", "3"), ("z", "3"), ("rdz", "3"), ("rdzw", "3"),
    ("zx", "3"), ("zy", "3"), ("div", "3"), ("defor11", "3"), ("defor22", "3"),
    ("defor33", "3"), ("defor12", "3"), ("defor13", "3"), ("defor23", "3"),
    ("xkmh", "3"), ("xkmv", "3"), ("xkhh", "3"), ("xkhv", "3"), ("bn2", "3"),
    ("mu_1", "2"), ("mu_2", "2"), ("mub", "2"), ("mut", "2"), ("muu", "2"), ("muv", "2"),
    ("muts", "2"), ("muus", "2"), ("muvs", "2"), ("muave", "2"), ("mudf", "2"),
    ("mu_tend", "2"), ("mu_tendf", "2"), ("mu_save", "2"), ("ht", "2"), ("msfux", "2"),
    ("msfuy", "2"), ("msfvx", "2"), ("msfvx_inv", "2"), ("msfvy", "2"), ("msftx", "2"),
    ("msfty", "2"), ("f", "2"), ("e", "2"), ("sina", "2"), ("cosa", "2"), ("clat", "2"),
    ("znw", "1"), ("znu", "1"), ("dnw", "1"), ("rdnw", "1"), ("dn", "1"), ("rdn", "1"),
    ("fnm", "1"), ("fnp", "1"), ("c1h", "1"), ("c2h", "1"), ("c1f", "1"), ("c2f", "1"),
    ("c3h", "1"), ("c4h", "1"), ("c3f", "1"), ("c4f", "1"), ("u_base", "1"),
    ("v_base", "1"), ("moist", "M"), ("moist_old", "M"), ("moist_tend", "M"),
    ("scalar", "S"), ("scalar_old", "S"), ("scalar_tend", "S"),
) + tuple(
    (f"{var}_{part}{edge}", kind)
    for var in ("u", "v", "t", "ph", "qv")
    for part in ("b", "bt")
    for edge, kind in (("xs", "X"), ("xe", "X"), ("ys", "Y"), ("ye", "Y"))
) + tuple(
    (f"mu_{part}{edge}", kind)
    for part in ("b", "bt")
    for edge, kind in (("xs", "x"), ("xe", "x"), ("ys", "y"), ("ye", "y"))
)

FIELD_KIND = dict(FIELD_SPECS)

# Prognostic and static arrays that define a model state between steps.  The
# rest are per-step work arrays (tendencies, coefficients, diagnostics) that
# every step recomputes before reading, except the few that the reference
# carries across steps (xkmh ring values, moist_tend relaxation, pm1 ...);
# those are kept resident too, so the whole table is the state.
PROGNOSTIC = (
    "u_2", "v_2", "w_2", "t_2", "ph_2", "mu_2", "p", "al", "moist", "scalar",
)


@dataclass(frozen=True)
class Dims:
    """Extents of one domain (mass-point counts plus derived staggered sizes)."""

    nx: int
    ny: int
    nz: int
    nmoist: int
    nscalar: int
    spec_bdy_width: int = 5
    pitch_align: int = 16

    @property
    def ide(self) -> int:
        return self.nx + 1

    @property
    def jde(self) -> int:
        return self.ny + 1

    @property
    def kde(self) -> int:
        return self.nz + 1

    @property
    def ip(self) -> int:
        a = self.pitch_align
        return ((self.ide + a - 1) // a) * a

    @property
    def n_mass(self) -> int:
        return self.nx * self.ny * self.nz

    def shape(self, kind: str) -> tuple[int, ...]:
        jde, kde, ip, bw = self.jde, self.kde, self.ip, self.spec_bdy_width
        if kind == "3":
            return (jde, kde, ip)
        if kind == "2":
            return (jde, ip)
        if kind == "1":
            return (kde,)
        if kind == "M":
            return (self.nmoist, jde, kde, ip)
        if kind == "S":
            return (max(self.nscalar, 0), jde, kde, ip)
        if kind == "X":
            return (bw, kde, jde)
        if kind == "Y":
            return (bw, kde, ip)
        if kind == "x":
            return (bw, jde)
        if kind == "y":
            return (bw, ip)
        raise ValueError(f"unknown field kind {kind!r}")

    def field_shape(self, name: str) -> tuple[int, ...]:
        return self.shape(FIELD_KIND[name])


__all__ = ["Dims", "FIELD_KIND", "FIELD_SPECS", "PROGNOSTIC"]
