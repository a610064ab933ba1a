"""Seeded synthetic ideal cases for the dycore (host-side input generation).

The cases follow the workload definitions of the benchmark configurations:
terrain-following eta levels stretched to a 50 hPa top, a dry base state with
constant Brunt-Vaisala frequency over a Gaussian hill, u = 10 m/s plus N(0,1)
noise, v = N(0,1), qv at 70 % relative humidity, hydrometeors either zero
or lognormal blobs, and a 5-point specified/relaxation lateral zone whose
boundary arrays hold the initial edge state with zero tendency.

The base state is integrated hydrostatically with the hypsometric form used by
the dycore (hypsometric_opt 2, src/wrf_jax/runtime.py:12369-12537) so that the
initial dry perturbation pressure is zero; moisture loading is balanced by the
WRF ideal-case top-down integration of the perturbation pressure.  Every random
draw comes from ``numpy.random.Generator(PCG64(20261202 + 1000*config +
member))`` in float64 ``[j][k][i]`` order and is cast to float32 at the end.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .config import DycoreConfig
from .fields import FIELD_SPECS, Dims

SVP1, SVP2, SVP3, SVPT0 = 0.6112, 17.67, 29.65, 273.15
EP_2 = 287.0 / 461.6


@dataclass
class Case:
    name: str
    dims: Dims
    config: DycoreConfig
    arrays: dict  # field name -> float32 ndarray in the registry layout


def vertical_coordinate(nz: int) -> dict:
    """Stretched eta levels and the interpolation constants of WRF's start_em."""
    kde = nz + 1
    k = np.arange(kde, dtype=np.float64)
    znw = (np.exp(-k / nz / 0.4) - np.exp(-1.0 / 0.4)) / (1.0 - np.exp(-1.0 / 0.4))
    znw[0], znw[-1] = 1.0, 0.0
    znu = np.zeros(kde)
    znu[:-1] = 0.5 * (znw[:-1] + znw[1:])
    dnw = np.zeros(kde)
    dnw[:-1] = znw[1:] - znw[:-1]
    rdnw = np.zeros(kde)
    rdnw[:-1] = 1.0 / dnw[:-1]
    dn = np.zeros(kde)
    rdn = np.zeros(kde)
    fnm = np.zeros(kde)
    fnp = np.zeros(kde)
    for kk in range(1, kde - 1):  # Fortran k = kk+1 in [2, kde-1]
        dn[kk] = 0.5 * (dnw[kk] + dnw[kk - 1])
        rdn[kk] = 1.0 / dn[kk]
        fnp[kk] = 0.5 * dnw[kk] / dn[kk]
        fnm[kk] = 0.5 * dnw[kk - 1] / dn[kk]
    # Fortran indices: dn(2) -> dn[1], dnw(1) -> dnw[0]
    cof1 = (2.0 * dn[1] + dn[2]) / (dn[1] + dn[2]) * dnw[0] / dn[1]
    cof2 = dn[1] / (dn[1] + dn[2]) * dnw[0] / dn[2]
    cf1 = fnp[1] + cof1
    cf2 = fnm[1] - cof1 - cof2
    cf3 = cof2
    cfn = (0.5 * dnw[kde - 2] + dn[kde - 2]) / dn[kde - 2]
    cfn1 = -0.5 * dnw[kde - 2] / dn[kde - 2]
    return dict(znw=znw, znu=znu, dnw=dnw, rdnw=rdnw, dn=dn, rdn=rdn, fnm=fnm, fnp=fnp,
                cf1=cf1, cf2=cf2, cf3=cf3, cfn=cfn, cfn1=cfn1)


def _theta_profile(theta_s: float, bvf: float, g: float, cp: float):
    """theta(z) = theta_s exp(N^2 z / g) and its hydrostatic Exner function."""
    a = g * g / (cp * theta_s * bvf * bvf)

    def exner(z):
        return 1.0 + a * (np.exp(-bvf * bvf * z / g) - 1.0)

    def height(pi):
        return -g / (bvf * bvf) * np.log(1.0 + (pi - 1.0) / a)

    def theta(z):
        return theta_s * np.exp(bvf * bvf * z / g)

    return exner, height, theta


def _copy_slots(a: np.ndarray, ide: int, jde: int, xstag: bool, ystag: bool) -> None:
    """Zero-gradient fill of the unused staggered slots (i = ide, j = jde)."""
    if a.ndim == 2:
        if not xstag:
            a[:, ide - 1] = a[:, ide - 2]
        if not ystag:
            a[jde - 1, :] = a[jde - 2, :]
    else:
        if not xstag:
            a[:, :, ide - 1] = a[:, :, ide - 2]
        if not ystag:
            a[jde - 1, :, :] = a[jde - 2, :, :]


def make_case(nx: int, ny: int, nz: int, *, name: str = "custom", config_index: int = 0, member: int = 0,
              dx: float = 3000.0, dt: float = 18.0, nmoist: int = 6, nscalar: int = 0,
              hydrometeors: bool = False, u0: float = 10.0, bvf: float = 0.01, theta_s: float = 290.0,
              hill_height: float = 500.0, hill_halfwidth: float = 10000.0, noise: float = 1.0, rh: float = 0.7,
              time_step_sound: int = 4, pitch_align: int = 16, dtype=np.float32) -> Case:
    dims = Dims(nx=nx, ny=ny, nz=nz, nmoist=nmoist, nscalar=nscalar, pitch_align=pitch_align)
    ide, jde, kde, ip = dims.ide, dims.jde, dims.kde, dims.ip
    vc = vertical_coordinate(nz)
    cfg = DycoreConfig(nx=nx, ny=ny, nz=nz, ip=ip, nmoist=nmoist, nscalar=nscalar, dx=dx, dy=dx, dt=dt,
                       time_step_sound=time_step_sound, cf1=vc["cf1"], cf2=vc["cf2"], cf3=vc["cf3"],
                       cfn=vc["cfn"], cfn1=vc["cfn1"]).validate()
    rng = np.random.Generator(np.random.PCG64(20261202 + 1000 * config_index + member))
    g, r_d, cp, p0, ptop, t0 = cfg.g, cfg.r_d, cfg.cp, cfg.p0, cfg.ptop, cfg.t0
    cvpm, rvovrd = cfg.cvpm, cfg.rvovrd
    A = {n: np.zeros(dims.shape(k), dtype=np.float64) for n, k in FIELD_SPECS}

    # 1-D coordinate arrays
    for key in ("znw", "znu", "dnw", "rdnw", "dn", "rdn", "fnm", "fnp"):
        A[key][:] = vc[key]
    A["c1h"][:] = 1.0
    A["c1f"][:] = 1.0
    A["c3h"][:] = vc["znu"]
    A["c3f"][:] = vc["znw"]

    # map factors, rotation, Coriolis (f-plane off)
    for key in ("msfux", "msfuy", "msfvx", "msfvx_inv", "msfvy", "msftx", "msfty", "cosa"):
        A[key][:] = 1.0

    # terrain: Gaussian hill at the domain centre (mass points)
    ii = np.arange(ide) + 1.0
    jj = np.arange(jde) + 1.0
    xc, yc = 0.5 * (nx + 1), 0.5 * (ny + 1)
    rr2 = ((ii[None, :] - xc) * dx) ** 2 + ((jj[:, None] - yc) * dx) ** 2
    ht = np.zeros((jde, ip))
    ht[:, :ide] = hill_height * np.exp(-rr2 / hill_halfwidth ** 2)
    _copy_slots(ht, ide, jde, False, False)
    A["ht"][:] = ht

    # dry base state on the eta levels
    exner, height, theta = _theta_profile(theta_s, bvf, g, cp)
    psfc = p0 * exner(ht) ** (cp / r_d)
    mub = psfc - ptop
    A["mub"][:] = mub
    pb = vc["znu"][None, :nz, None] * mub[:, None, :] + ptop          # (jde, nz, ip)
    zb = height((pb / p0) ** (r_d / cp))
    thb = theta(zb)
    alb = (r_d / p0) * thb * (pb / p0) ** cvpm
    A["pb"][:, :nz, :] = pb
    A["alb"][:, :nz, :] = alb
    phb = np.zeros((jde, kde, ip))
    phb[:, 0, :] = g * ht
    for k in range(nz):
        pfd = vc["znw"][k] * mub + ptop
        pfu = vc["znw"][k + 1] * mub + ptop
        phb[:, k + 1, :] = phb[:, k, :] + alb[:, k, :] * pb[:, k, :] * np.log(pfd / pfu)
    A["phb"][:] = phb
    t2 = np.zeros((jde, kde, ip))
    t2[:, :nz, :] = thb - t0
    A["t_2"][:] = t2

    # moisture: qv at rh (default 70 %) RH of the base-state temperature and pressure
    temp = thb * (pb / p0) ** (r_d / cp)
    es = SVP1 * np.exp(SVP2 * (temp - SVPT0) / (temp - SVP3))
    qsat = EP_2 * es / (pb / 1000.0 - es)
    qv = np.zeros((jde, kde, ip))
    qv[:, :nz, :] = np.clip(rh * qsat, 0.0, 0.03)
    A["moist"][0] = qv

    zagl = 0.5 * (phb[:, :-1, :] + phb[:, 1:, :]) / g - ht[:, None, :]  # (jde, nz, ip)
    if hydrometeors and nmoist > 1:
        ncell = max(1, int(round(nx * dx * ny * dx / (60.0e3 * 60.0e3))))
        cx = rng.uniform(1.0, nx, size=ncell)
        cy = rng.uniform(1.0, ny, size=ncell)
        radius = 10.0e3
        mask = np.zeros((jde, nz, ip), dtype=bool)
        for c in range(ncell):
            d2 = ((ii[None, :] - cx[c]) * dx) ** 2 + ((jj[:, None] - cy[c]) * dx) ** 2
            inside = np.zeros((jde, ip), dtype=bool)
            inside[:, :ide] = d2 <= radius * radius
            mask |= inside[:, None, :] & (zagl >= 1000.0) & (zagl <= 8000.0)
        for s in range(1, nmoist):
            zdraw = rng.standard_normal(size=(jde, nz, ip))
            q = np.where(mask, 1.0e-4 * np.exp(1.0 * zdraw), 0.0)
            A["moist"][s][:, :nz, :] = q
    if nscalar > 0:
        for s in range(nscalar):
            A["scalar"][s][:, :nz, :] = 10.0 ** rng.uniform(3.0, 6.0, size=(jde, nz, ip))

    # winds: u = u0 + noise on u points, v = noise on v points
    u = np.zeros((jde, kde, ip))
    v = np.zeros((jde, kde, ip))
    u[:ny, :nz, :ide] = u0 + noise * rng.standard_normal(size=(ny, nz, ide))
    v[:jde, :nz, :nx] = noise * rng.standard_normal(size=(jde, nz, nx))
    A["u_2"][:] = u
    A["v_2"][:] = v

    # moist hydrostatic perturbation pressure (WRF ideal-case top-down
    # integration with mu = 0) and the matching specific volume
    mu = np.zeros((jde, ip))
    p = np.zeros((jde, kde, ip))
    al = np.zeros((jde, kde, ip))
    rdnw, rdn = vc["rdnw"], vc["rdn"]
    for k in range(nz - 1, -1, -1):
        if k == nz - 1:
            qvf1 = qv[:, k, :]
            qvf2 = 1.0 / (1.0 + qvf1)
            qvf1 = qvf1 * qvf2
            p[:, k, :] = -0.5 * (mu + qvf1 * mub) / rdnw[k] / qvf2
        else:
            qvf1 = 0.5 * (qv[:, k, :] + qv[:, k + 1, :])
            qvf2 = 1.0 / (1.0 + qvf1)
            qvf1 = qvf1 * qvf2
            p[:, k, :] = p[:, k + 1, :] - (mu + qvf1 * mub) / qvf2 / rdn[k + 1]
        qvf = 1.0 + rvovrd * qv[:, k, :]
        alt = (r_d / p0) * (t2[:, k, :] + t0) * qvf * ((p[:, k, :] + pb[:, k, :]) / p0) ** cvpm
        al[:, k, :] = alt - alb[:, k, :]
    ph = np.zeros((jde, kde, ip))
    for k in range(nz):
        pfd = vc["znw"][k] * mub + ptop
        pfu = vc["znw"][k + 1] * mub + ptop
        ph[:, k + 1, :] = ph[:, k, :] + al[:, k, :] * pb[:, k, :] * np.log(pfd / pfu)
    A["p"][:] = p
    A["al"][:] = al
    A["ph_2"][:] = ph

    # lower-boundary w from the terrain slope (set_w_surface), interior zero
    w = np.zeros((jde, kde, ip))
    cf1, cf2, cf3 = vc["cf1"], vc["cf2"], vc["cf3"]
    for j in range(jde - 1):
        jm, jp = max(j - 1, 0), min(j + 1, jde - 2)
        for_i = np.arange(ide - 1)
        im = np.maximum(for_i - 1, 0)
        ipp = np.minimum(for_i + 1, ide - 2)
        vjp = cf1 * v[jp, 0, for_i] + cf2 * v[jp, 1, for_i] + cf3 * v[jp, 2, for_i]
        vj = cf1 * v[j, 0, for_i] + cf2 * v[j, 1, for_i] + cf3 * v[j, 2, for_i]
        uip = cf1 * u[j, 0, ipp] + cf2 * u[j, 1, ipp] + cf3 * u[j, 2, ipp]
        ui = cf1 * u[j, 0, for_i] + cf2 * u[j, 1, for_i] + cf3 * u[j, 2, for_i]
        yp = 0.5 / dx * ((ht[jp, for_i] - ht[j, for_i]) * vjp + (ht[j, for_i] - ht[jm, for_i]) * vj)
        xp = 0.5 / dx * ((ht[j, ipp] - ht[j, for_i]) * uip + (ht[j, for_i] - ht[j, im]) * ui)
        w[j, 0, :ide - 1] = yp + xp
    A["w_2"][:] = w

    # unused staggered slots: zero-gradient, as the physical BC leaves them
    for key, xs, ys in (("u_2", True, False), ("v_2", False, True), ("w_2", False, False), ("t_2", False, False),
                        ("ph_2", False, False), ("p", False, False), ("al", False, False), ("pb", False, False),
                        ("alb", False, False), ("phb", False, False)):
        _copy_slots(A[key], ide, jde, xs, ys)
    for s in range(nmoist):
        _copy_slots(A["moist"][s], ide, jde, False, False)
    for s in range(nscalar):
        _copy_slots(A["scalar"][s], ide, jde, False, False)
    for key in ("mub",):
        _copy_slots(A[key], ide, jde, False, False)

    # previous-time-level copies
    for a, b in (("u_1", "u_2"), ("v_1", "v_2"), ("w_1", "w_2"), ("t_1", "t_2"), ("ph_1", "ph_2"), ("mu_1", "mu_2")):
        A[a][:] = A[b]

    arrays = {k: np.ascontiguousarray(v.astype(dtype)) for k, v in A.items()}
    # boundary arrays are computed in the target precision from the rounded
    # fields, with the dycore's own operation order, so that the relaxation
    # of the initial state toward them is exactly zero
    _fill_boundary_arrays(arrays, dims)
    return Case(name=name, dims=dims, config=cfg, arrays=arrays)


def _fill_boundary_arrays(A: dict, d: Dims) -> None:
    """Lateral-boundary arrays = coupled initial edge state, zero tendency."""
    ide, jde, kde, bw = d.ide, d.jde, d.kde, d.spec_bdy_width
    mub, mu = A["mub"], A["mu_2"]
    half = mub.dtype.type(0.5)
    mut = mub + mu
    # coupled fields as the dycore's relaxation sees them (calc_mu_uv order)
    ia = np.arange(ide)
    ib = np.clip(ia - 1, 0, None)
    ia_ = np.where(ia == ide - 1, ide - 2, ia)
    ib_ = np.where(ia == ide - 1, ide - 2, ib)
    muu = np.zeros_like(mut)
    muu[:, :ide] = half * (((mu[:, ia_] + mu[:, ib_]) + mub[:, ia_]) + mub[:, ib_])
    ja = np.arange(jde)
    jb = np.clip(ja - 1, 0, None)
    ja_ = np.where(ja == jde - 1, jde - 2, ja)
    jb_ = np.where(ja == jde - 1, jde - 2, jb)
    muv = np.zeros_like(mut)
    muv[:jde, :] = half * (((mu[ja_, :] + mu[jb_, :]) + mub[ja_, :]) + mub[jb_, :])
    ru = A["u_2"] * muu[:, None, :]
    rv = A["v_2"] * muv[:, None, :]
    rt = A["t_2"] * mut[:, None, :]
    rph = A["ph_2"] * mut[:, None, :]
    rqv = A["moist"][0] * mut[:, None, :]
    for name, fld, ibe, jbe in (("u", ru, ide, jde - 1), ("v", rv, ide - 1, jde), ("t", rt, ide - 1, jde - 1),
                                ("ph", rph, ide - 1, jde - 1), ("qv", rqv, ide - 1, jde - 1)):
        for b in range(1, bw + 1):
            A[f"{name}_bxs"][b - 1, :, :] = fld[:, :, b - 1].T
            A[f"{name}_bxe"][b - 1, :, :] = fld[:, :, ibe - b].T
            A[f"{name}_bys"][b - 1, :, :] = fld[b - 1, :, :]
            A[f"{name}_bye"][b - 1, :, :] = fld[jbe - b, :, :]
    for b in range(1, bw + 1):
        A["mu_bxs"][b - 1, :] = mu[:, b - 1]
        A["mu_bxe"][b - 1, :] = mu[:, ide - 1 - b]
        A["mu_bys"][b - 1, :] = mu[b - 1, :]
        A["mu_bye"][b - 1, :] = mu[jde - 1 - b, :]


# The workloads of the benchmark configurations (sizes and species).
CONFIGS = {
    "c0": dict(nx=100, ny=100, nz=45, config_index=0, nmoist=6, nscalar=0, hydrometeors=False),
    "c1": dict(nx=400, ny=400, nz=60, config_index=1, nmoist=6, nscalar=2, hydrometeors=True),
    "c3": dict(nx=200, ny=200, nz=45, config_index=3, nmoist=1, nscalar=0, hydrometeors=False),
}


def named_case(name: str, **overrides) -> Case:
    params = dict(CONFIGS[name])
    params.update(overrides)
    return make_case(name=name, **params)


__all__ = ["CONFIGS", "Case", "make_case", "named_case", "vertical_coordinate"]
