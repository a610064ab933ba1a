"""Pins of the CPU oracle that do not depend on the reference's own tests
(the reference ships no golden vectors for the dycore): closed forms of the
flux operators, invariants of the full step, and the WRF step-count formula.
A plausible slip in the restatement (a dropped term, a wrong sign, an index
off by one, a transposed operand) fails at least one of these.
"""

import numpy as np
import pytest

from oracle import Oracle
from wrfdc.cases import make_case, vertical_coordinate
from wrfdc.config import DycoreConfig
from wrfdc.fields import FIELD_SPECS, Dims


def blank(nx=24, ny=20, nz=10, precision="f64", **cfg):
    vc = vertical_coordinate(nz)
    dims = Dims(nx, ny, nz, nmoist=1, nscalar=0)
    config = DycoreConfig(nx=nx, ny=ny, nz=nz, ip=dims.ip, nmoist=1, nscalar=0, cf1=vc["cf1"], cf2=vc["cf2"],
                          cf3=vc["cf3"], cfn=vc["cfn"], cfn1=vc["cfn1"], **cfg)
    dt = np.float64 if precision == "f64" else np.float32
    arrays = {n: np.zeros(dims.shape(k), dtype=dt) for n, k in FIELD_SPECS}
    for n in ("msfux", "msfuy", "msfvx", "msfvx_inv", "msfvy", "msftx", "msfty"):
        arrays[n][:] = 1.0
    for n in ("dnw", "rdnw", "dn", "rdn", "fnm", "fnp", "c1h", "c2h", "c1f", "c2f", "znw", "znu"):
        arrays[n][:] = vc[n][: arrays[n].shape[0]] if n in vc else 0.0
    return config, dims, arrays


# ---------------------------------------------------------------------------
# closed forms: flux operators of advect_scalar
# ---------------------------------------------------------------------------
def cell_avg(poly, s):
    """Exact average of the polynomial over [s-1/2, s+1/2]."""
    P = poly.integ()
    return P(s + 0.5) - P(s - 0.5)


@pytest.mark.parametrize("degree", [0, 1, 2, 3, 4])
def test_advect_scalar_x_polynomial_exact(degree):
    """Constant ru, q(i) the cell averages of a polynomial g of degree <= 4:
    the 5th-order flux reconstructs g exactly at the faces (its centred part
    is exact to degree 5, its upwind part a 5th difference), so
    tend = -rdx*ru*(g(i+1/2) - g(i-1/2)) wherever both faces are 5th order."""
    config, d, A = blank()
    ru0 = 7.0
    g = np.polynomial.Polynomial.fromroots([9.3] * degree) * 0.01 if degree else np.polynomial.Polynomial([0.01])
    i = np.arange(1, d.ip + 1, dtype=np.float64)
    A["t_2"][:] = cell_avg(g, i)[None, None, :]
    A["ru"][:] = ru0
    o = Oracle(config, A, "f64")
    o.stage("advect_scalar", 2, 0)
    ii = np.arange(4, d.ide - 3)  # faces i and i+1 both in [4, ide-3]
    expect = -(1.0 / config.dx) * ru0 * (g(ii + 0.5) - g(ii - 0.5))
    np.testing.assert_allclose(o.arrays["t_tend"][2, 3, ii - 1], expect, rtol=1e-11, atol=1e-15)


@pytest.mark.parametrize("degree", [0, 1, 2, 3, 4])
def test_advect_scalar_y_polynomial_exact(degree):
    config, d, A = blank()
    rv0 = -4.0
    g = np.polynomial.Polynomial.fromroots([7.7] * degree) * 0.02 if degree else np.polynomial.Polynomial([0.02])
    j = np.arange(1, d.jde + 1, dtype=np.float64)
    A["t_2"][:] = cell_avg(g, j)[:, None, None]
    A["rv"][:] = rv0
    o = Oracle(config, A, "f64")
    o.stage("advect_scalar", 2, 0)
    jj = np.arange(4, d.jde - 3)
    expect = -(1.0 / config.dy) * rv0 * (g(jj + 0.5) - g(jj - 0.5))
    np.testing.assert_allclose(o.arrays["t_tend"][jj - 1, 4, 6], expect, rtol=1e-11, atol=1e-15)


@pytest.mark.parametrize("degree", [0, 1, 2])
def test_advect_scalar_z_polynomial_exact(degree):
    """Vertical 3rd-order flux on a uniform grid: exact for cell averages of
    degree <= 2 (its upwind part is a 3rd difference)."""
    config, d, A = blank(nz=12)
    A["fnm"][:] = 0.5
    A["fnp"][:] = 0.5
    A["rdnw"][:] = 1.0
    w0 = 3.0
    g = np.polynomial.Polynomial.fromroots([2.2] * degree) * 0.5 if degree else np.polynomial.Polynomial([0.5])
    k = np.arange(1, d.kde + 1, dtype=np.float64)
    A["t_2"][:] = cell_avg(g, k)[None, :, None]
    A["ww"][:] = w0
    o = Oracle(config, A, "f64")
    o.stage("advect_scalar", 2, 0)
    kk = np.arange(3, d.kde - 2)  # faces k and k+1 both in [3, ktf-1]
    expect = -w0 * (g(kk + 0.5) - g(kk - 0.5))
    np.testing.assert_allclose(o.arrays["t_tend"][5, kk - 1, 5], expect, rtol=1e-11, atol=1e-13)


def test_advect_scalar_upwind_dissipates_a_spike():
    """Unit spike at i0: the centred (velocity-odd) part of the tendency at
    i0 is 0 (the two faces carry 37/60 each) and the upwind (velocity-even)
    part is -rdx*|u|*20/60 (5th differences -10 and +10 at the two faces):
    the upwind bias removes mass from the peak whatever the flow direction."""
    config, d, A = blank()
    i0 = 11
    A["t_2"][:, :, i0 - 1] = 1.0
    res = {}
    for sgn in (1.0, -1.0):
        A["ru"][:] = 2.0 * sgn
        o = Oracle(config, A, "f64")
        o.stage("advect_scalar", 2, 0)
        res[sgn] = o.arrays["t_tend"][3, 3, i0 - 1]
    odd = 0.5 * (res[1.0] - res[-1.0])
    even = 0.5 * (res[1.0] + res[-1.0])
    rdx = 1.0 / config.dx
    assert abs(odd) < 1e-18
    assert even == pytest.approx(-rdx * 2.0 * 20.0 / 60.0, rel=1e-12)


def test_num_sound_steps_formula():
    """WRF's acoustic step count: max(2*(int(300*dt/dx - 0.01)+1), 4)."""
    cases = [(18.0, 3000.0, 4), (6.0, 1000.0, 4), (30.0, 3000.0, 6), (60.0, 3000.0, 12), (54.0, 9000.0, 4)]
    for dt, dx, ns in cases:
        c = DycoreConfig(nx=16, ny=16, nz=8, ip=32, nmoist=1, nscalar=0, dt=dt, dx=dx, dy=dx, time_step_sound=0)
        assert c.num_sound_steps() == ns, (dt, dx)


# ---------------------------------------------------------------------------
# invariants of the full step (FP64 oracle)
# ---------------------------------------------------------------------------
def _speed(A, nx, ny):
    return max(np.abs(A["u_2"][:-1, :-1, :nx + 1]).max(), np.abs(A["v_2"][:, :-1, :nx]).max(), np.abs(A["w_2"][:-1, :, :nx]).max())


@pytest.mark.parametrize("hill", [0.0, 500.0])
def test_dry_rest_state_stays_at_rest(hill):
    """Hydrostatic dry atmosphere at rest (with or without terrain) stays at
    rest: pins the pressure gradient, buoyancy, EOS and the implicit solver."""
    case = make_case(20, 18, 12, u0=0.0, noise=0.0, rh=0.0, hill_height=hill, nmoist=1, nscalar=0, dtype=np.float64)
    o = Oracle(case.config, case.arrays, "f64")
    o.step(3, 0.0)
    assert _speed(o.arrays, 20, 18) < 1e-9
    o.close()


def test_uniform_flow_flat_stays_uniform():
    """Uniform u over flat ground: only the earth-curvature term u^2/R acts."""
    case = make_case(20, 18, 12, u0=10.0, noise=0.0, rh=0.0, hill_height=0.0, nmoist=1, nscalar=0, dtype=np.float64)
    o = Oracle(case.config, case.arrays, "f64")
    o.step(2, 0.0)
    u = o.arrays["u_2"][:-1, :-1, :21]
    assert np.abs(u - 10.0).max() < 5e-3
    assert np.abs(o.arrays["v_2"][:, :-1, :20]).max() < 5e-3
    o.close()


def _column_tracer_mass(A, c, s):
    nx, ny, nz = c.nx, c.ny, c.nz
    mu = A["mu_2"][:ny, :nx].astype(np.float64) + A["mub"][:ny, :nx]
    w = A["c1h"][:nz, None, None] * mu[:, None, :][None, :, 0, :].transpose(1, 0, 2) if False else None
    m = (A["c1h"][None, :nz, None] * mu[:, None, :] + A["c2h"][None, :nz, None]) * A["dnw"][None, :nz, None]
    return float((m * A["scalar"][s][:ny, :nz, :nx]).sum())


def test_scalar_mass_conserved_away_from_boundaries():
    """Flux-form transport with the mass-weighted RK update conserves the
    tracer mass sum((c1*mu+c2)*dnw*q) while the tracer stays off the lateral
    boundary zone (FP64, PD on the last stage)."""
    case = make_case(28, 24, 10, nmoist=1, nscalar=1, dtype=np.float64)
    A = case.arrays
    nx, ny, nz = 28, 24, 10
    blob = np.zeros_like(A["scalar"][0])
    jj, kk, ii = np.meshgrid(np.arange(ny + 1), np.arange(nz + 1), np.arange(blob.shape[2]), indexing="ij")
    blob[:] = np.exp(-(((ii - 14) / 3.0) ** 2 + ((jj - 12) / 3.0) ** 2 + ((kk - 4) / 2.0) ** 2))
    blob[:, :, nx:] = 0.0
    A["scalar"][0] = blob
    A["scalar_old"][0] = blob
    before = _column_tracer_mass(A, case.config, 0)
    o = Oracle(case.config, A, "f64")
    o.step(1, 0.0)
    after = _column_tracer_mass(o.arrays, case.config, 0)
    assert abs(after - before) / before < 1e-10
    o.close()


def test_pd_limiter_guarantees_nonnegative_update():
    """Positive-definite advection (last RK stage): with a sharp box of the
    species, the mass-weighted update (c1*(mu_1+mub)+c2)*q_old + dt*tend stays
    >= 0 on the limited cells, while the unlimited 5th-order tendency drives
    it negative at the box edges (so the check has teeth)."""
    case = make_case(24, 20, 10, nmoist=1, nscalar=0, u0=20.0, dtype=np.float64)
    o = Oracle(case.config, case.arrays, "f64")
    o.step(1, 0.0)
    A = o.copy()
    o.close()
    box = np.zeros_like(A["moist"][0])
    box[6:10, 2:6, 8:11] = 1e-3
    A["moist"][0] = box
    A["moist_old"][0] = box
    nx, ny, nz, dt = 24, 20, 10, case.config.dt
    mass = A["c1h"][None, :nz, None] * (A["mu_1"][:ny, None, :nx] + A["mub"][:ny, None, :nx]) + A["c2h"][None, :nz, None]

    def updated(stage):
        B = {k: v.copy() for k, v in A.items()}
        B["advect_tend"][:] = 0.0
        if stage == "plain":
            B["t_2"] = B["moist"][0].copy()
            B["t_tend"][:] = 0.0
            B["ru"], B["rv"], B["ww"] = B["ru_m"], B["rv_m"], B["ww_m"]
            oo = Oracle(case.config, B, "f64")
            oo.stage("advect_scalar", 3, 0)
            tend = oo.arrays["t_tend"]
        else:
            oo = Oracle(case.config, B, "f64")
            oo.stage("advect_scalar_pd", 3, 0)
            tend = oo.arrays["advect_tend"]
        out = mass * box[:ny, :nz, :nx] + dt * tend[:ny, :nz, :nx] * A["msfty"][:ny, None, :nx]
        oo.close()
        return out[1:ny - 1, :, 1:nx - 1]

    assert updated("plain").min() < -1e-9
    assert updated("pd").min() >= -1e-15


def test_f32_build_tracks_f64_build():
    """The FP32 and FP64 builds of the restatement agree to FP32 noise."""
    c32 = make_case(24, 20, 10, nmoist=6, nscalar=2, hydrometeors=True)
    c64 = make_case(24, 20, 10, nmoist=6, nscalar=2, hydrometeors=True, dtype=np.float64)
    a = Oracle(c32.config, c32.arrays, "f32")
    b = Oracle(c64.config, c64.arrays, "f64")
    a.step(1, 0.0)
    b.step(1, 0.0)
    for n in ("u_2", "v_2", "w_2", "t_2", "ph_2", "mu_2", "moist", "scalar"):
        x = a.arrays[n].astype(np.float64)
        y = b.arrays[n]
        assert np.linalg.norm(x - y) <= 2e-3 * max(np.linalg.norm(y), 1e-30), n
