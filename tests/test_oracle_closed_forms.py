"""Closed-form pins of the oracle's momentum, diffusion, acoustic and
boundary stages.

The HIP kernels are compared with the CPU oracle to the last bit, so every
misreading of the reference that the oracle carries would pass unnoticed.
Each test below sets up a field for which the mathematics fixes the answer
without reference to how the stage is coded: polynomial exactness of the
flux operators, dissipation of a spike by the upwind term, the f-plane and
earth-curvature terms of the momentum equations, diffusion reducing to
mu*K times a second derivative under hydrostatic metrics, the implicit w
solve against a banded LU from scipy, the acoustic velocity update with no
gradients / with one pressure gradient / with a step in the mass tendency,
the substep average of constant mass fluxes, and the Davies relaxation
weights.  A slip of sign, index or coefficient in the restatement moves at
least one of these answers (tests/test_oracle_mutations.py plants such
slips and checks that they are caught).

Everything runs on the FP64 build of the oracle on small grids.
"""

import numpy as np
import pytest
from scipy.linalg import solve_banded

from oracle import Oracle
from wrfdc.cases import vertical_coordinate
from wrfdc.config import DycoreConfig
from wrfdc.fields import FIELD_SPECS, Dims

NX, NY, NZ = 24, 20, 10
F64 = "f64"


def blank(nx=NX, ny=NY, nz=NZ, uniform_eta=False, **cfg):
    """A zero state with unit map factors, terrain-following mass levels
    (c1 = 1, c2 = 0) and either WRF's stretched eta levels or a uniform
    eta spacing (dnw = -1/nz)."""
    vc = vertical_coordinate(nz)
    if uniform_eta:
        kde = nz + 1
        znw = 1.0 - np.arange(kde) / nz
        dnw = np.zeros(kde)
        dnw[:-1] = np.diff(znw)
        vc = dict(vc)
        vc["znw"] = znw
        vc["dnw"] = dnw
        vc["rdnw"] = np.where(dnw != 0.0, 1.0 / np.where(dnw != 0.0, dnw, 1.0), 0.0)
        dn = np.zeros(kde)
        dn[1:-1] = 0.5 * (dnw[1:-1] + dnw[:-2])
        vc["dn"] = dn
        vc["rdn"] = np.where(dn != 0.0, 1.0 / np.where(dn != 0.0, dn, 1.0), 0.0)
        fz = np.zeros(kde)
        fz[1:-1] = 0.5
        vc["fnm"], vc["fnp"] = fz.copy(), fz.copy()
    dims = Dims(nx, ny, nz, nmoist=1, nscalar=0)
    config = DycoreConfig(nx=nx, ny=ny, nz=nz, ip=dims.ip, nmoist=1, nscalar=0, cf1=vc["cf1"], cf2=vc["cf2"],
                          cf3=vc["cf3"], cfn=vc["cfn"], cfn1=vc["cfn1"], **cfg)
    A = {n: np.zeros(dims.shape(k), dtype=np.float64) for n, k in FIELD_SPECS}
    for n in ("msfux", "msfuy", "msfvx", "msfvx_inv", "msfvy", "msftx", "msfty", "cosa"):
        A[n][:] = 1.0
    for n in ("znw", "znu", "dnw", "rdnw", "dn", "rdn", "fnm", "fnp"):
        A[n][:] = vc[n][: A[n].shape[0]]
    A["c1h"][:] = 1.0
    A["c1f"][:] = 1.0
    return config, dims, A


def run(config, A, *stages):
    o = Oracle(config, A, F64)
    for name, rk, it in stages:
        o.stage(name, rk, it)
    out = o.copy()
    o.close()
    return out


def cell_avg(poly, s):
    """Average of the polynomial over [s-1/2, s+1/2]."""
    P = poly.integ()
    return P(s + 0.5) - P(s - 0.5)


def poly(degree, root, scale):
    if degree == 0:
        return np.polynomial.Polynomial([scale])
    return np.polynomial.Polynomial.fromroots([root] * degree) * scale


# ---------------------------------------------------------------------------
# momentum advection: advect_u / advect_v / advect_w
# ---------------------------------------------------------------------------
# (stage, advected field, tendency, axis) -> the index range whose two faces
# both take the high order (5th horizontal, 3rd vertical), as (lo, hi) offsets
# from the staggered extent (ide/jde/kde) of that axis.
MOMENTUM = {
    ("advect_u", "x"): (4, -3), ("advect_u", "y"): (4, -4), ("advect_u", "z"): (3, -3),
    ("advect_v", "x"): (4, -4), ("advect_v", "y"): (4, -3), ("advect_v", "z"): (3, -3),
    ("advect_w", "x"): (4, -4), ("advect_w", "y"): (4, -4), ("advect_w", "z"): (3, -2),
}
FIELD_OF = {"advect_u": ("u_2", "ru_tend"), "advect_v": ("v_2", "rv_tend"), "advect_w": ("w_2", "rw_tend")}
VEL_OF = {"x": "ru", "y": "rv", "z": "ww"}


def _axis_len(d, axis):
    return {"x": d.ide, "y": d.jde, "z": d.kde}[axis]


def _line(a, axis, idx, fixed):
    """View of a 3-D registry array along one axis at 1-based indices."""
    i0, k0, j0 = fixed
    if axis == "x":
        return a[j0 - 1, k0 - 1, idx - 1]
    if axis == "y":
        return a[idx - 1, k0 - 1, i0 - 1]
    return a[j0 - 1, idx - 1, i0 - 1]


def _fill_along(a, axis, values):
    n = len(values)
    if axis == "x":
        a[:, :, :n] = values[None, None, :]
    elif axis == "y":
        a[:n] = values[:, None, None]
    else:
        a[:, :n, :] = values[None, :, None]


@pytest.mark.parametrize("degree", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("stage,axis", sorted(MOMENTUM), ids=lambda v: str(v))
def test_momentum_advection_polynomial_exact(stage, axis, degree):
    """Uniform mass flux V along one axis and a field whose values are cell
    averages of a polynomial g: where both faces of a point take the high
    order, the flux divergence is exactly -rd*V*(g(s+1/2) - g(s-1/2)) (5th
    order horizontally: exact to degree 4; 3rd order vertically on a uniform
    eta grid: exact to degree 2)."""
    if axis == "z" and degree > 2:
        pytest.skip("3rd-order vertical flux is exact to degree 2")
    config, d, A = blank(uniform_eta=(axis == "z"))
    fld, tend = FIELD_OF[stage]
    vel = 3.0 if axis != "y" else -4.0
    g = poly(degree, 7.3, 0.01)
    n = _axis_len(d, axis)
    s = np.arange(1, n + 1, dtype=np.float64)
    _fill_along(A[fld], axis, cell_avg(g, s))
    A[VEL_OF[axis]][:] = vel
    out = run(config, A, (stage, 2, 0))
    lo, hi = MOSTLY = MOMENTUM[(stage, axis)]
    idx = np.arange(lo, n + hi + 1)
    if axis == "x":
        rd = 1.0 / config.dx
    elif axis == "y":
        rd = 1.0 / config.dy
    else:  # rdnw (mass levels) or rdn (w levels), uniform here
        rd = A["rdn"][3] if stage == "advect_w" else A["rdnw"][3]
    expect = -rd * vel * (g(idx + 0.5) - g(idx - 0.5))
    fixed = (9, 4, 8)
    got = _line(out[tend], axis, idx, fixed)
    np.testing.assert_allclose(got, expect, rtol=1e-10, atol=1e-14 * max(1.0, abs(rd * vel)))


@pytest.mark.parametrize("stage,axis", sorted(MOMENTUM), ids=lambda v: str(v))
def test_momentum_advection_upwind_dissipates_spike(stage, axis):
    """A unit spike advected by a uniform mass flux of either sign: the part
    of the tendency odd in the flux vanishes at the spike (centred part) and
    the even part removes mass from it: -rd*|V|*20/60 for the 5th-order
    horizontal flux (5th differences -10 and +10 at its faces) and
    -|rd|*|V|/2 for the 3rd-order vertical flux, with WRF's stretched eta
    levels (rdnw, rdn < 0).  An upwind bias taken in the wrong direction
    makes the even part grow the spike instead."""
    config, d, A = blank()
    fld, tend = FIELD_OF[stage]
    i0, k0, j0 = 11, 5, 9
    A[fld][j0 - 1, k0 - 1, i0 - 1] = 1.0
    res = {}
    for sgn in (1.0, -1.0):
        B = {k: v.copy() for k, v in A.items()}
        B[VEL_OF[axis]][:] = 2.0 * sgn
        res[sgn] = run(config, B, (stage, 2, 0))[tend][j0 - 1, k0 - 1, i0 - 1]
    odd = 0.5 * (res[1.0] - res[-1.0])
    even = 0.5 * (res[1.0] + res[-1.0])
    if axis == "z":
        rd = A["rdn"][k0 - 1] if stage == "advect_w" else A["rdnw"][k0 - 1]
        assert rd < 0.0
        expect = -abs(rd) * 2.0 / 2.0
    else:
        rd = 1.0 / (config.dx if axis == "x" else config.dy)
        expect = -rd * 2.0 * 20.0 / 60.0
    assert abs(odd) <= 1e-15 * abs(expect)
    assert even == pytest.approx(expect, rel=1e-12)


@pytest.mark.parametrize("stage,axis,vel", [("advect_u", "x", "ru"), ("advect_v", "y", "rv")])
def test_uniform_momentum_own_flux_divergence_is_centred(stage, axis, vel):
    """A uniform component C carried along its own axis by a mass flux s*n^2:
    every flux order reproduces C at a face, whose mass flux is the mean of
    its two neighbouring points, so the tendency is the centred derivative
    -rd*C*s*2n at every point of the row, boundary faces included."""
    config, d, A = blank()
    fld, tend = FIELD_OF[stage]
    A[fld][:] = 1.7
    n = _axis_len(d, axis)
    s = np.arange(1, n + 1, dtype=np.float64)
    _fill_along(A[vel], axis, 0.05 * s ** 2)
    out = run(config, A, (stage, 2, 0))
    rd = 1.0 / (config.dx if axis == "x" else config.dy)
    idx = np.arange(2, n)  # u points [2, ide-1] / v rows [2, jde-1]
    got = _line(out[tend], axis, idx, (9, 4, 8))
    np.testing.assert_allclose(got, -rd * 1.7 * 0.05 * 2.0 * idx, rtol=1e-12)


@pytest.mark.parametrize("stage", ["advect_u", "advect_v"])
def test_uniform_momentum_cross_flux_averages_neighbours(stage):
    """u carried across y faces (v across x faces) by the cross mass flux
    s*n^2*m (n along the component's own axis, m across it): the face flux
    of a u point i is the mean of the v-point fluxes at i-1 and i (of a v
    point j, of the u-point fluxes at j-1 and j), so the divergence across
    the row is -rd*C*s*(n^2 + (n-1)^2)/2 per unit of m."""
    config, d, A = blank()
    fld, tend = FIELD_OF[stage]
    A[fld][:] = 1.7
    ii = np.arange(1, d.ip + 1, dtype=np.float64)
    jj = np.arange(1, d.jde + 1, dtype=np.float64)
    s = 0.05
    if stage == "advect_u":  # rv(i, j) = s*i^2*j, divergence along y
        A["rv"][:] = (s * ii[None, :] ** 2 * jj[:, None])[:, None, :]
        rd = 1.0 / config.dy
        i = np.arange(2, d.ide)
        expect = -rd * 1.7 * 0.5 * s * (i ** 2 + (i - 1) ** 2)
        got = run(config, A, (stage, 2, 0))[tend][7, 3, i - 1]
    else:  # ru(i, j) = s*j^2*i, divergence along x
        A["ru"][:] = (s * jj[:, None] ** 2 * ii[None, :])[:, None, :]
        rd = 1.0 / config.dx
        j = np.arange(2, d.jde)
        expect = -rd * 1.7 * 0.5 * s * (j ** 2 + (j - 1) ** 2)
        got = run(config, A, (stage, 2, 0))[tend][j - 1, 3, 8]
    np.testing.assert_allclose(got, expect, rtol=1e-12)


# ---------------------------------------------------------------------------
# Coriolis and curvature (rk_tendency)
# ---------------------------------------------------------------------------
def _interior(a, d, k_range, lo=3, trim=3):
    """Points i in [lo, ide-1-trim], j in [lo, jde-1-trim], k in k_range."""
    return a[lo - 1:d.jde - 1 - trim, k_range[0] - 1:k_range[1], lo - 1:d.ide - 1 - trim]


def test_coriolis_f_plane():
    """Resting fields carrying uniform mass fluxes (RU, RV, RW) on an f-plane
    with a horizontal Coriolis parameter e (grid aligned with the meridian,
    sin(alpha) = 0): the momentum tendencies are those of the rotating
    frame, RU_t = f*RV - e*RW, RV_t = -f*RU, RW_t = e*RU (northward flow is
    turned to the east in the northern hemisphere)."""
    config, d, A = blank()
    f0, e0, R1, R2, R3 = 1.0e-4, 3.0e-5, 7.0, -5.0, 2.0
    A["f"][:] = f0
    A["e"][:] = e0
    A["ru"][:] = R1
    A["rv"][:] = R2
    A["rw"][:] = R3
    out = run(config, A, ("rk_tendency", 2, 0))
    kt = d.kde - 1
    np.testing.assert_allclose(_interior(out["ru_tend"], d, (1, kt)), f0 * R2 - e0 * R3, rtol=1e-12)
    np.testing.assert_allclose(_interior(out["rv_tend"], d, (1, kt)), -f0 * R1, rtol=1e-12)
    np.testing.assert_allclose(_interior(out["rw_tend"], d, (2, kt)), e0 * R1, rtol=1e-12)


def test_curvature_terms():
    """Uniform winds (U, V) carried by uniform mass fluxes (RU, RV, RW) with
    no rotation: only the earth-curvature terms act, RU_t = -U*RW/a,
    RV_t = -V*RW/a, RW_t = (RU*U + RV*V)/a."""
    config, d, A = blank()
    U, V, R1, R2, R3 = 12.0, -6.0, 9.0e4 * 12.0, -9.0e4 * 6.0, 1.0e4
    A["u_2"][:] = U
    A["v_2"][:] = V
    A["ru"][:] = R1
    A["rv"][:] = R2
    A["rw"][:] = R3
    out = run(config, A, ("rk_tendency", 2, 0))
    rer = config.reradius
    kt = d.kde - 1
    np.testing.assert_allclose(_interior(out["ru_tend"], d, (1, kt)), -U * R3 * rer, rtol=1e-10)
    np.testing.assert_allclose(_interior(out["rv_tend"], d, (1, kt)), -V * R3 * rer, rtol=1e-10)
    np.testing.assert_allclose(_interior(out["rw_tend"], d, (2, kt)), rer * (R1 * U + R2 * V), rtol=1e-10)


# ---------------------------------------------------------------------------
# Smagorinsky eddy viscosity (smag2d_km, diff_opt 2 / km_opt 4)
# ---------------------------------------------------------------------------
SMAG_CASES = [
    # D11, D22, D12, |zx|, rdzw
    (1.0e-3, -2.0e-3, 5.0e-4, 0.0, 0.01),     # plain Smagorinsky, flat
    (1.0, 0.0, 0.0, 0.0, 0.01),               # capped at 10*l
    (1.0e-3, -2.0e-3, 5.0e-4, 0.1, 0.01),     # steep slope, weak deformation: km / alpha
    (4.0e-3, -6.0e-3, 1.0e-3, 0.1, 0.01),     # steep slope, strong deformation: km / alpha^2
]


@pytest.mark.parametrize("d11,d22,d12,zx,rdzw", SMAG_CASES)
def test_smag2d_km_closed_form(d11, d22, d12, zx, rdzw):
    """Uniform deformation: km = min(c_s^2 l^2 |D|, 10 l), l = sqrt(dx dy),
    |D|^2 = (D11-D22)^2/4 + D12^2; over a slope s = |zx|*rdzw*dx the
    diffusivity is divided by alpha = max(s, 1), or by alpha^2 when |D|
    exceeds max(10/l, 1e-3); kh = km / Pr with Pr = 1/3, khv = 0."""
    config, d, A = blank()
    A["defor11"][:] = d11
    A["defor22"][:] = d22
    A["defor12"][:] = d12
    A["zx"][:] = zx
    A["rdzw"][:] = rdzw
    out = run(config, A, ("smag2d_km", 1, 0))
    l = np.sqrt(config.dx * config.dy)
    D = np.sqrt(0.25 * (d11 - d22) ** 2 + d12 ** 2)
    km = min(config.c_s ** 2 * l * l * D, 10.0 * l)
    alpha = max(zx * rdzw * config.dx, 1.0)
    km /= alpha * alpha if D > max(10.0 / l, 1.0e-3) else alpha
    sl = (slice(1, d.jde - 2), slice(0, d.kde - 1), slice(1, d.ide - 2))  # i, j in [2, e-2]
    np.testing.assert_allclose(out["xkmh"][sl], km, rtol=1e-12)
    np.testing.assert_allclose(out["xkmv"][sl], km, rtol=1e-12)
    np.testing.assert_allclose(out["xkhh"][sl], 3.0 * km, rtol=1e-12)
    assert np.all(out["xkhv"][sl] == 0.0)


# ---------------------------------------------------------------------------
# horizontal diffusion: mu * K * second derivative under hydrostatic metrics
# ---------------------------------------------------------------------------
MU0, RHO0, KM = 8.0e4, 1.1, 250.0


def _hydrostatic_flat(config, A):
    """Geopotential of a flat column with constant density: g*dz = -mu*deta/rho."""
    dnw = A["dnw"]
    phi = np.concatenate([[0.0], np.cumsum(-MU0 * dnw[:-1] / RHO0)])
    A["phb"][:] = phi[None, :, None]


DIFF_CASES = [
    # field, tendency, coordinate, factor: D11 = 2 u_x (u), D22 = 2 v_y (v), D13 = w_x (w), theta
    ("u_2", "ru_tendf", "x", 4.0),
    ("v_2", "rv_tendf", "y", 4.0),
    ("w_2", "rw_tendf", "x", 2.0),
    ("t_2", "t_tendf", "x", 2.0),
]


@pytest.mark.parametrize("fld,tend,axis,factor", DIFF_CASES)
def test_horizontal_diffusion_is_mu_k_second_derivative(fld, tend, axis, factor):
    """Flat terrain, constant density and diffusivity, a field a*x^2 + b*x + c
    along one axis: the coupled diffusion tendency is mu*K*d/dx(D) with the
    deformation D = 2 u_x for the normal stresses (4 mu K a) and D = w_x for
    w and grad(theta) for theta (2 mu K a); the linear part diffuses nothing.
    mu enters only through the hydrostatic metrics g*dz/deta = -mu/rho that
    compute_diff_metrics derives from the geopotential."""
    config, d, A = blank()
    _hydrostatic_flat(config, A)
    a, b, c = 2.0e-7, 3.0e-4, 5.0
    n = _axis_len(d, axis)
    h = config.dx if axis == "x" else config.dy
    s = (np.arange(1, n + 1, dtype=np.float64) - 1.0) * h
    _fill_along(A[fld], axis, a * s * s + b * s + c)
    o = Oracle(config, A, F64)
    o.stage("compute_diff_metrics", 1, 0)
    o.stage("cal_deform_and_div", 1, 0)
    for n_ in ("xkmh", "xkmv", "xkhh"):
        o.arrays[n_][:] = KM
    o.arrays["rho"][:] = RHO0
    o.stage("horizontal_diffusion_2", 1, 0)
    out = o.copy()
    o.close()
    ks = (2, d.kde - 1) if fld == "w_2" else (1, d.kde - 1)
    got = _interior(out[tend], d, ks, lo=4, trim=4)
    np.testing.assert_allclose(got, factor * MU0 * KM * a, rtol=1e-8)


# ---------------------------------------------------------------------------
# calc_coef_w + advance_w: the implicit w-phi system
# ---------------------------------------------------------------------------
def _coef_w_matrix(A, config, dts, i, j):
    """Tridiagonal of the implicit w-phi solve for unknowns w(2..kde), written
    from calc_coef_w (module_small_step_em.py:1359-1586): sub-diagonal a,
    diagonal b, super-diagonal c; cof = (dts*g*(1+epssm)/2)^2."""
    kde = A["rdn"].shape[0]
    g, eps = config.g, config.epssm
    cof = (0.5 * dts * g * (1.0 + eps)) ** 2
    m = A["mut"][j - 1, i - 1]
    c1h, c2h, c1f, c2f = A["c1h"], A["c2h"], A["c1f"], A["c2f"]
    rdn, rdnw = A["rdn"], A["rdnw"]
    c2a = A["c2a"][j - 1, :, i - 1]
    cqw = A["cqw"][j - 1, :, i - 1]

    def mh(k):
        return c1h[k - 1] * m + c2h[k - 1]

    def mf(k):
        return c1f[k - 1] * m + c2f[k - 1]

    n = kde - 1  # unknowns k = 2..kde
    ab = np.zeros((3, n))
    for k in range(2, kde):
        r = k - 2
        if k >= 3:
            ab[2, r - 1] = -cqw[k - 1] * cof * rdn[k - 1] * rdnw[k - 2] * c2a[k - 2] / (mh(k - 1) * mf(k - 1))
        bu = rdnw[k - 1] * c2a[k - 1] / (mh(k) * mf(k))
        bl = rdnw[k - 2] * c2a[k - 2] / (mh(k - 1) * mf(k))
        ab[1, r] = 1.0 + cqw[k - 1] * cof * rdn[k - 1] * (bu + bl)
        ab[0, r + 1] = -cqw[k - 1] * cof * rdn[k - 1] * rdnw[k - 1] * c2a[k - 1] / (mh(k) * mf(k + 1))
    top = kde - 2  # row of the unknown w(kde)
    ab[2, top - 1] = -2.0 * cof * rdnw[kde - 2] ** 2 * c2a[kde - 2] / (mh(kde - 1) * mf(kde - 1))
    ab[1, top] = 1.0 + 2.0 * cof * rdnw[kde - 2] ** 2 * c2a[kde - 2] / (mh(kde - 1) * mf(kde))
    return ab


def test_advance_w_solves_the_implicit_system():
    """With no explicit forcing except rw_tend (resting w and phi, no
    theta perturbation, no slope), advance_w's new w is the solution of the
    calc_coef_w tridiagonal with right-hand side dts*rw_tend, and phi takes
    the implicit part dts*g*(1+epssm)/2 * w / (c1f*muts + c2f).  Checked
    against scipy's banded LU at every interior column, with seeded column
    masses, level-dependent hybrid mass factors (c1, c2), sound-speed factors
    c2a and moisture factors cqw."""
    config, d, A = blank(nx=16, ny=14, nz=12)
    rng = np.random.default_rng(7)
    A["mut"][:] = 9.0e4 + 5.0e3 * rng.random(A["mut"].shape)
    A["muts"][:] = A["mut"]
    # hybrid mass coordinate: every level has its own (c1, c2)
    kde = d.kde
    A["c1f"][:] = np.linspace(1.0, 0.55, kde)
    A["c2f"][:] = np.linspace(0.0, 2.0e4, kde)
    A["c1h"][:] = np.linspace(0.98, 0.57, kde)
    A["c2h"][:] = np.linspace(1.0e3, 1.9e4, kde)
    A["c2a"][:] = 1.0e5 * (1.0 + 0.3 * rng.random(A["c2a"].shape))
    A["cqw"][:] = 0.9 + 0.1 * rng.random(A["cqw"].shape)
    A["rw_tend"][:] = 50.0 * rng.standard_normal(A["rw_tend"].shape)
    out = run(config, A, ("calc_coef_w", 2, 0), ("advance_w", 2, 1))
    dts = config.dt / config.num_sound_steps()
    kde = d.kde
    for j in range(2, d.jde - 1):
        for i in range(2, d.ide - 1):
            ab = _coef_w_matrix(A, config, dts, i, j)
            rhs = dts * A["rw_tend"][j - 1, 1:kde, i - 1]
            w = solve_banded((1, 1), ab, rhs)
            np.testing.assert_allclose(out["w_2"][j - 1, 1:kde, i - 1], w, rtol=1e-9, atol=1e-12 * np.abs(w).max())
            ph = 0.5 * dts * config.g * (1.0 + config.epssm) * w / (A["c1f"][1:kde] * A["muts"][j - 1, i - 1] + A["c2f"][1:kde])
            np.testing.assert_allclose(out["ph_2"][j - 1, 1:kde, i - 1], ph, rtol=1e-9, atol=1e-12 * np.abs(ph).max())
            assert out["w_2"][j - 1, 0, i - 1] == 0.0


# ---------------------------------------------------------------------------
# advance_uv
# ---------------------------------------------------------------------------
def _uv_state(seed=3):
    config, d, A = blank()
    rng = np.random.default_rng(seed)
    A["u_2"][:] = rng.standard_normal(A["u_2"].shape)
    A["v_2"][:] = rng.standard_normal(A["v_2"].shape)
    A["ru_tend"][:] = rng.standard_normal(A["ru_tend"].shape)
    A["rv_tend"][:] = rng.standard_normal(A["rv_tend"].shape)
    A["muu"][:] = 9.0e4
    A["muv"][:] = 9.0e4
    A["cqu"][:] = 1.0
    A["cqv"][:] = 1.0
    return config, d, A


def _u_pts(a, d):  # u points of advance_uv: i in [2, ide-1], j in [2, jde-2]
    return a[1:d.jde - 2, :d.kde - 1, 1:d.ide - 1]


def _v_pts(a, d):  # v points: i in [2, ide-2], j in [2, jde-1]
    return a[1:d.jde - 1, :d.kde - 1, 1:d.ide - 2]


def test_advance_uv_without_gradients_adds_the_tendency():
    """No pressure, geopotential or mass-tendency gradients: u += dts*ru_tend,
    v += dts*rv_tend."""
    config, d, A = _uv_state()
    out = run(config, A, ("advance_uv", 2, 1))
    dts = config.dt / config.num_sound_steps()
    np.testing.assert_allclose(_u_pts(out["u_2"], d), _u_pts(A["u_2"] + dts * A["ru_tend"], d), rtol=1e-14)
    np.testing.assert_allclose(_v_pts(out["v_2"], d), _v_pts(A["v_2"] + dts * A["rv_tend"], d), rtol=1e-14)


def test_advance_uv_pressure_gradient():
    """Perturbation pressure linear in x (slope P per grid length) with
    constant inverse density alpha and column mass mu: the acoustic momentum
    update is dU = -dts * mu * alpha * dp/dx, and v feels nothing."""
    config, d, A = _uv_state()
    P, alpha = 40.0, 0.9
    A["p"][:] = P * np.arange(1, d.ip + 1, dtype=np.float64)[None, None, :]
    A["alt"][:] = alpha
    out = run(config, A, ("advance_uv", 2, 1))
    dts = config.dt / config.num_sound_steps()
    expect = A["u_2"] + dts * A["ru_tend"] - dts * 9.0e4 * alpha * P / config.dx
    np.testing.assert_allclose(_u_pts(out["u_2"], d), _u_pts(expect, d), rtol=1e-12)
    np.testing.assert_allclose(_v_pts(out["v_2"], d), _v_pts(A["v_2"] + dts * A["rv_tend"], d), rtol=1e-14)


@pytest.mark.parametrize("axis", ["x", "y"])
def test_advance_uv_external_mode_damping(axis):
    """A step M in the column-mass tendency mudf across one face: the external
    mode filter decelerates the flow into the column whose mass grows,
    du = -emdiv * dx * c1 * M at that face (dv = -emdiv * dy * c1 * M for a
    step along y), and nowhere else."""
    config, d, A = _uv_state()
    M, s0 = 0.05, 9
    if axis == "x":
        A["mudf"][:, s0 - 1:] = M
    else:
        A["mudf"][s0 - 1:, :] = M
    out = run(config, A, ("advance_uv", 2, 1))
    dts = config.dt / config.num_sound_steps()
    du = out["u_2"] - (A["u_2"] + dts * A["ru_tend"])
    dv = out["v_2"] - (A["v_2"] + dts * A["rv_tend"])
    c1 = A["c1h"][: d.kde - 1]
    if axis == "x":
        band = du[1:d.jde - 2, :d.kde - 1, s0 - 1]
        np.testing.assert_allclose(band, np.broadcast_to(-config.emdiv * config.dx * c1 * M, band.shape), rtol=1e-10)
        du[:, :, s0 - 1] = 0.0
        assert np.abs(_u_pts(du, d)).max() < 1e-12
        assert np.abs(_v_pts(dv, d)).max() < 1e-12
    else:
        band = dv[s0 - 1, :d.kde - 1, 1:d.ide - 2]
        np.testing.assert_allclose(band, np.broadcast_to((-config.emdiv * config.dy * c1 * M)[:, None], band.shape), rtol=1e-10)
        dv[s0 - 1] = 0.0
        assert np.abs(_v_pts(dv, d)).max() < 1e-12
        assert np.abs(_u_pts(du, d)).max() < 1e-12


# ---------------------------------------------------------------------------
# sumflux
# ---------------------------------------------------------------------------
def test_sumflux_time_mean_of_substep_fluxes():
    """Perturbation mass fluxes growing linearly over the ns substeps of the
    last RK stage (n*U, n*V, n*W at substep n): the time-averaged fluxes
    for scalar transport are their mean (ns+1)/2*(U, V, W) plus the
    stage-start fluxes (mu_u*u_lin, mu_v*v_lin, ww_lin); whatever the
    accumulators held before the first substep is discarded."""
    config, d, A = blank()
    U, V, W, ul, vl, wl, mu = 0.7, -0.4, 0.02, 3.0, -2.0, 0.005, 9.0e4
    A["muu"][:] = mu
    A["muv"][:] = mu
    A["u_save"][:] = ul
    A["v_save"][:] = vl
    A["ww_1"][:] = wl
    for n_ in ("ru_m", "rv_m", "ww_m"):
        A[n_][:] = 123.0
    ns = config.num_sound_steps()
    o = Oracle(config, A, F64)
    for it in range(1, ns + 1):
        o.arrays["u_2"][:] = U * it
        o.arrays["v_2"][:] = V * it
        o.arrays["ww"][:] = W * it
        o.stage("sumflux", 3, it)
    out = o.copy()
    o.close()
    mean = 0.5 * (ns + 1)
    np.testing.assert_allclose(out["ru_m"][:d.jde - 1, :d.kde - 1, :d.ide], U * mean + mu * ul, rtol=1e-13)
    np.testing.assert_allclose(out["rv_m"][:d.jde, :d.kde - 1, :d.ide - 1], V * mean + mu * vl, rtol=1e-13)
    np.testing.assert_allclose(out["ww_m"][:d.jde - 1, :d.kde, :d.ide - 1], W * mean + wl, rtol=1e-13)


# ---------------------------------------------------------------------------
# lateral relaxation zone (relax_bdy_dry)
# ---------------------------------------------------------------------------
def test_relax_zone_davies_weights():
    """Boundary value 0 and a field q = a*(i^2 + j^2) + c, so that
    bdy - q has the 5-point Laplacian -4a everywhere: on relaxation row or
    column b (b = spec_zone+1 .. spec_zone+relax_zone-1 from the edge) the
    tendency is F1(b)*(bdy-q) - F2(b)*Lap(bdy-q) with the weights of
    lbc_fcx_gcx (boundary/lbc_fcx_gcx.py:10-56, spec_exp 0):
    F1 = 0.1/dt * (sz+rz-b)/(rz-1), F2 = 1/(50 dt) * (sz+rz-b)/(rz-1).
    The F2 term diffuses the departure from the boundary value."""
    config, d, A = blank()
    a, c = 0.3, 11.0
    ii = np.arange(1, d.ip + 1, dtype=np.float64)
    jj = np.arange(1, d.jde + 1, dtype=np.float64)
    q = a * (ii[None, :] ** 2 + jj[:, None] ** 2) + c
    A["ru"][:] = q[:, None, :]
    out = run(config, A, ("dry_boundary_tendency", 1, 0))
    sz, rz, dt = config.spec_zone, config.relax_zone, config.dt

    def weights(b):
        return 0.1 / dt * (sz + rz - b) / (rz - 1), 1.0 / dt / 50.0 * (sz + rz - b) / (rz - 1)

    kt = d.kde - 1
    mid_i = np.arange(8, d.ide - 7)
    mid_j = np.arange(8, d.jde - 7)
    for b in range(sz + 1, sz + rz):
        f1, f2 = weights(b)
        # south row j = b, north row j = jde-1-(b-1); west column i = b, east i = ide-(b-1)
        for j in (b, d.jde - 1 - (b - 1)):
            got = out["ru_tendf"][j - 1][:kt][:, mid_i - 1]
            expect = f1 * (-q[j - 1, mid_i - 1]) - f2 * (-4.0 * a)
            np.testing.assert_allclose(got, np.broadcast_to(expect, got.shape), rtol=1e-12)
        for i in (b, d.ide - (b - 1)):
            got = out["ru_tendf"][mid_j - 1][:, :kt, i - 1]  # (j, k)
            expect = f1 * (-q[mid_j - 1, i - 1]) - f2 * (-4.0 * a)
            np.testing.assert_allclose(got, np.broadcast_to(expect[:, None], got.shape), rtol=1e-12)


# ---------------------------------------------------------------------------
# Brunt-Vaisala frequency (calculate_km_kh's N^2)
# ---------------------------------------------------------------------------
def _column_state(theta_of_z, qv_of_z, qc=0.0, nz=12):
    """Flat columns with geopotential levels 0, 500, 1000, ... m (mass levels
    midway), theta(z), qv(z), a constant cloud water qc, and a pressure that
    falls with a 8 km scale height; metrics from compute_diff_metrics."""
    config, d, A = blank(nz=nz)
    nmoist = 2 if qc else 1
    dims = Dims(NX, NY, nz, nmoist=nmoist, nscalar=0)
    vc = vertical_coordinate(nz)
    config = DycoreConfig(nx=NX, ny=NY, nz=nz, ip=dims.ip, nmoist=nmoist, nscalar=0, cf1=vc["cf1"], cf2=vc["cf2"],
                          cf3=vc["cf3"], cfn=vc["cfn"], cfn1=vc["cfn1"])
    B = {n: np.zeros(dims.shape(k), dtype=np.float64) for n, k in FIELD_SPECS}
    for n in A:
        if A[n].shape == B[n].shape:
            B[n][:] = A[n]
    zw = 500.0 * np.arange(nz + 1)
    B["phb"][:] = (config.g * zw)[None, :, None]
    zm = np.zeros(nz + 1)
    zm[:nz] = 0.5 * (zw[:-1] + zw[1:])
    B["t_2"][:] = (theta_of_z(zm) - config.t0)[None, :, None]
    B["pb"][:] = (1.0e5 * np.exp(-zm / 8000.0))[None, :, None]
    B["moist"][0] = qv_of_z(zm)[None, :, None]
    if qc:
        B["moist"][1][:] = qc
    return config, dims, B, zm


def _n2(config, B):
    out = run(config, B, ("compute_diff_metrics", 1, 0), ("calculate_n2", 1, 0))
    return out["bn2"]


def test_n2_dry_linear_theta():
    """Dry air with theta = theta0 + Gamma*z: the centred N^2 of every interior
    level is exactly g*Gamma/theta(z); the top level repeats the one below."""
    G0, th0 = 0.004, 290.0
    config, d, B, zm = _column_state(lambda z: th0 + G0 * z, lambda z: 0.0 * z)
    bn2 = _n2(config, B)
    kt = d.kde - 1
    ks = np.arange(2, kt)  # Fortran levels 2..ktf-1
    expect = config.g * G0 / (th0 + G0 * zm[ks - 1])
    got = bn2[8, ks - 1, 9]
    np.testing.assert_allclose(got, expect, rtol=1e-12)
    assert bn2[8, kt - 1, 9] == bn2[8, kt - 2, 9]


def test_n2_unsaturated_moisture_gradient():
    """Unsaturated air with qv = q0 + beta*z under isentropic theta: the
    virtual-temperature term 1.61*dqv/dz less the total-water term dqv/dz
    leaves N^2 = 0.61*g*beta."""
    q0, beta = 5.0e-4, -8.0e-8  # below saturation at every level
    config, d, B, zm = _column_state(lambda z: 300.0 + 0.0 * z, lambda z: q0 + beta * z)
    bn2 = _n2(config, B)
    ks = np.arange(2, d.kde - 1)
    np.testing.assert_allclose(bn2[8, ks - 1, 9], 0.61 * config.g * beta, rtol=1e-10)


def test_n2_saturated_uses_equivalent_potential_temperature():
    """Cloudy air (qc >= 1e-5 marks it saturated) with no vapour and constant
    total water: N^2 = g/theta * d(theta_e)/dz with theta_e = theta*(1 +
    L*qs/(cp*T)), T = theta*(p/1e5)^(R/cp) and the saturation mixing ratio
    qs = 0.622*es/(p - es) of es = 611.2*exp(17.67*(T-273.15)/(T-29.65)) Pa
    (Bolton 1980), centred over the two neighbouring levels."""
    config, d, B, zm = _column_state(lambda z: 290.0 + 0.003 * z, lambda z: 0.0 * z, qc=2.0e-5)
    bn2 = _n2(config, B)
    L, cp, R = 2.5e6, 3.5 * 287.0, 287.0
    th = 290.0 + 0.003 * zm
    p = 1.0e5 * np.exp(-zm / 8000.0)
    T = th * (p / 1.0e5) ** (R / cp)
    es = 611.2 * np.exp(17.67 * (T - 273.15) / (T - 29.65))
    qs = (287.0 / 461.6) * es / (p - es)
    the = th * (1.0 + L * qs / (cp * T))
    ks = np.arange(2, d.kde - 1)
    expect = config.g / th[ks - 1] * (the[ks] - the[ks - 2]) / (zm[ks] - zm[ks - 2])
    np.testing.assert_allclose(bn2[8, ks - 1, 9], expect, rtol=1e-9)
