"""The seeded ideal case of ``wrfdc.cases.make_case``, built on the GPU.

``device_case`` creates a ``Domain`` and fills its buffers in place with the
case ``make_case`` defines (the same formulas: stretched eta levels, dry
base state of constant N over a Gaussian hill integrated hydrostatically
with the dycore's hypsometric form, qv at a fixed relative humidity, the
moist hydrostatic perturbation pressure, u = u0 + noise, v = noise, the
terrain-following surface w, boundary arrays holding the coupled initial
edge state).  The arithmetic runs in FP64 on the device, a chunk of rows at
a time, and each field is rounded once into its FP32 buffer, so nothing of
the size of the domain passes through host memory: the 3000x3000x60 truth
run of the learned closure (540 M points, about 196 GB of registry with one
species) is built in seconds.

Differences from ``make_case``: the random draws come from torch's
generator seeded with 20261202 + 1000*config + member (a chunk of rows c
uses that seed + c), not NumPy's PCG64, so the noise fields differ;
hydrometeor blobs and extra scalars are not generated (species beyond qv
start at zero); the surface w is formed from the FP32 winds.
"""

from __future__ import annotations

import math

import numpy as np

from .cases import EP_2, SVP1, SVP2, SVP3, SVPT0, vertical_coordinate
from .config import DycoreConfig
from .fields import Dims


def _torch():
    import torch

    return torch


def _copy_slots(t, ide: int, jde: int, xstag: bool, ystag: bool) -> None:
    """Zero-gradient fill of the unused staggered slots (i = ide, j = jde) of
    a device tensor [j][k][i] or [j][i]."""
    if not xstag:
        t[..., ide - 1] = t[..., ide - 2]
    if not ystag:
        t[jde - 1] = t[jde - 2]


def device_case(nx: int, ny: int, nz: int, *, device: str = "cuda:0", config_index: int = 0, member: int = 0,
                dx: float = 3000.0, dt: float = 18.0, nmoist: int = 1, nscalar: int = 0, u0: float = 10.0,
                bvf: float = 0.01, theta_s: float = 290.0, hill_height: float = 500.0, hill_halfwidth: float = 10000.0,
                noise: float = 1.0, rh: float = 0.7, time_step_sound: int = 4, rows_per_chunk: int = 96):
    """A Domain holding the seeded ideal case (see the module docstring)."""
    torch = _torch()
    from .device import Domain

    dims = Dims(nx=nx, ny=ny, nz=nz, nmoist=nmoist, nscalar=nscalar)
    ide, jde, kde, ip = dims.ide, dims.jde, dims.kde, dims.ip
    vc = vertical_coordinate(nz)
    cfg = DycoreConfig(nx=nx, ny=ny, nz=nz, ip=ip, nmoist=nmoist, nscalar=nscalar, dx=dx, dy=dx, dt=dt,
                       time_step_sound=time_step_sound, cf1=vc["cf1"], cf2=vc["cf2"], cf3=vc["cf3"],
                       cfn=vc["cfn"], cfn1=vc["cfn1"]).validate()
    dom = Domain(cfg, device)
    B = dom.buffers
    dev = dom.device
    f64 = torch.float64
    g, r_d, cp, p0, ptop, t0 = cfg.g, cfg.r_d, cfg.cp, cfg.p0, cfg.ptop, cfg.t0
    cvpm, rvovrd = cfg.cvpm, cfg.rvovrd
    seed = 20261202 + 1000 * config_index + member

    def d1(a):
        return torch.tensor(np.asarray(a, dtype=np.float64), dtype=f64, device=dev)

    with torch.cuda.device(dev):
        for key in ("znw", "znu", "dnw", "rdnw", "dn", "rdn", "fnm", "fnp"):
            B[key].copy_(d1(vc[key]))
        B["c1h"].fill_(1.0)
        B["c1f"].fill_(1.0)
        B["c3h"].copy_(d1(vc["znu"]))
        B["c3f"].copy_(d1(vc["znw"]))
        for key in ("msfux", "msfuy", "msfvx", "msfvx_inv", "msfvy", "msftx", "msfty", "cosa"):
            B[key].fill_(1.0)
        znw, znu, rdnw, rdn = d1(vc["znw"]), d1(vc["znu"]), d1(vc["rdnw"]), d1(vc["rdn"])

        # terrain and column mass over the whole plane (2-D, FP64)
        ii = torch.arange(ide, dtype=f64, device=dev) + 1.0
        jj = torch.arange(jde, dtype=f64, device=dev) + 1.0
        xc, yc = 0.5 * (nx + 1), 0.5 * (ny + 1)
        rr2 = ((ii[None, :] - xc) * dx) ** 2 + ((jj[:, None] - yc) * dx) ** 2
        ht = torch.zeros((jde, ip), dtype=f64, device=dev)
        ht[:, :ide] = hill_height * torch.exp(-rr2 / hill_halfwidth ** 2)
        _copy_slots(ht, ide, jde, False, False)
        a_ex = g * g / (cp * theta_s * bvf * bvf)
        exner_s = 1.0 + a_ex * (torch.exp(-bvf * bvf * ht / g) - 1.0)
        mub = p0 * exner_s ** (cp / r_d) - ptop
        B["ht"].copy_(ht)
        B["mub"].copy_(mub)

        # base state, moisture, noise and perturbation pressure, a chunk of rows at a time
        for c, j0 in enumerate(range(0, jde, rows_per_chunk)):
            j1 = min(j0 + rows_per_chunk, jde)
            m = mub[j0:j1]
            hc = ht[j0:j1]
            pb = znu[None, :nz, None] * m[:, None, :] + ptop
            pi = (pb / p0) ** (r_d / cp)
            zb = -g / (bvf * bvf) * torch.log(1.0 + (pi - 1.0) / a_ex)
            thb = theta_s * torch.exp(bvf * bvf * zb / g)
            alb = (r_d / p0) * thb * (pb / p0) ** cvpm
            pfd = znw[None, :nz, None] * m[:, None, :] + ptop
            pfu = znw[None, 1:nz + 1, None] * m[:, None, :] + ptop
            lg = torch.log(pfd / pfu)
            phb = torch.zeros((j1 - j0, kde, ip), dtype=f64, device=dev)
            phb[:, 0, :] = g * hc
            phb[:, 1:, :] = g * hc[:, None, :] + torch.cumsum(alb * pb * lg, dim=1)
            t2 = thb - t0
            temp = thb * pi
            es = SVP1 * torch.exp(SVP2 * (temp - SVPT0) / (temp - SVP3))
            qv = torch.clamp(rh * (EP_2 * es / (pb / 1000.0 - es)), 0.0, 0.03)
            # moist hydrostatic perturbation pressure, top down, mu = 0
            p = torch.zeros((j1 - j0, nz, ip), dtype=f64, device=dev)
            for k in range(nz - 1, -1, -1):
                if k == nz - 1:
                    q1 = qv[:, k, :]
                    q2 = 1.0 / (1.0 + q1)
                    p[:, k, :] = -0.5 * (q1 * q2 * m) / rdnw[k] / q2
                else:
                    q1 = 0.5 * (qv[:, k, :] + qv[:, k + 1, :])
                    q2 = 1.0 / (1.0 + q1)
                    p[:, k, :] = p[:, k + 1, :] - (q1 * q2 * m) / q2 / rdn[k + 1]
            alt = (r_d / p0) * (t2 + t0) * (1.0 + rvovrd * qv) * ((p + pb) / p0) ** cvpm
            al = alt - alb
            ph = torch.zeros((j1 - j0, kde, ip), dtype=f64, device=dev)
            ph[:, 1:, :] = torch.cumsum(al * pb * lg, dim=1)
            rows = slice(j0, j1)
            B["pb"][rows, :nz].copy_(pb)
            B["alb"][rows, :nz].copy_(alb)
            B["phb"][rows].copy_(phb)
            B["t_2"][rows, :nz].copy_(t2)
            B["moist"][0, rows, :nz].copy_(qv)
            B["p"][rows, :nz].copy_(p)
            B["al"][rows, :nz].copy_(al)
            B["ph_2"][rows].copy_(ph)
            gen = torch.Generator(device=dev)
            gen.manual_seed(seed + c)
            nu = max(0, min(j1, ny) - j0)
            if nu > 0:
                B["u_2"][j0:j0 + nu, :nz, :ide].copy_(
                    u0 + noise * torch.randn((nu, nz, ide), generator=gen, dtype=torch.float32, device=dev))
            B["v_2"][rows, :nz, :nx].copy_(noise * torch.randn((j1 - j0, nz, nx), generator=gen, dtype=torch.float32, device=dev))
            del pb, pi, zb, thb, alb, pfd, pfu, lg, phb, t2, temp, es, qv, p, alt, al, ph

        for key, xs, ys in (("u_2", True, False), ("v_2", False, True), ("t_2", False, False), ("ph_2", False, False),
                            ("p", False, False), ("al", False, False), ("pb", False, False), ("alb", False, False),
                            ("phb", False, False)):
            _copy_slots(B[key], ide, jde, xs, ys)
        _copy_slots(B["moist"][0], ide, jde, False, False)

        # lower-boundary w from the terrain slope (set_w_surface), interior zero
        cf1, cf2, cf3 = vc["cf1"], vc["cf2"], vc["cf3"]
        u, v = B["u_2"], B["v_2"]
        ub = (cf1 * u[:, 0, :].double() + cf2 * u[:, 1, :].double() + cf3 * u[:, 2, :].double())
        vb = (cf1 * v[:, 0, :].double() + cf2 * v[:, 1, :].double() + cf3 * v[:, 2, :].double())
        jr = torch.arange(jde - 1, device=dev)
        jm, jp = torch.clamp(jr - 1, min=0), torch.clamp(jr + 1, max=jde - 2)
        ir = torch.arange(ide - 1, device=dev)
        im, ipp = torch.clamp(ir - 1, min=0), torch.clamp(ir + 1, max=ide - 2)
        hh = ht[:jde - 1, :ide - 1]
        yp = 0.5 / dx * ((ht[jp][:, :ide - 1] - hh) * vb[jp][:, :ide - 1] + (hh - ht[jm][:, :ide - 1]) * vb[:jde - 1, :ide - 1])
        xp = 0.5 / dx * ((ht[:jde - 1, ipp] - hh) * ub[:jde - 1][:, ipp] + (hh - ht[:jde - 1, im]) * ub[:jde - 1, :ide - 1])
        B["w_2"][:jde - 1, 0, :ide - 1].copy_(yp + xp)
        _copy_slots(B["w_2"], ide, jde, False, False)

        for a, b in (("u_1", "u_2"), ("v_1", "v_2"), ("w_1", "w_2"), ("t_1", "t_2"), ("ph_1", "ph_2"), ("mu_1", "mu_2")):
            B[a].copy_(B[b])
        _fill_boundary_arrays(B, dims)
    torch.cuda.synchronize(dev)
    return dom


def _fill_boundary_arrays(B: dict, d: Dims) -> None:
    """Lateral-boundary arrays = coupled initial edge state, zero tendency
    (FP32, the arithmetic of wrfdc.cases._fill_boundary_arrays)."""
    torch = _torch()
    ide, jde, bw = d.ide, d.jde, d.spec_bdy_width
    mub, mu = B["mub"], B["mu_2"]
    mut = mub + mu
    dev = mub.device
    ia = torch.arange(ide, device=dev)
    ib = torch.clamp(ia - 1, min=0)
    ia_ = torch.where(ia == ide - 1, ide - 2, ia)
    ib_ = torch.where(ia == ide - 1, ide - 2, ib)
    muu = torch.zeros_like(mut)
    muu[:, :ide] = 0.5 * (((mu[:, ia_] + mu[:, ib_]) + mub[:, ia_]) + mub[:, ib_])
    ja = torch.arange(jde, device=dev)
    jb = torch.clamp(ja - 1, min=0)
    ja_ = torch.where(ja == jde - 1, jde - 2, ja)
    jb_ = torch.where(ja == jde - 1, jde - 2, jb)
    muv = torch.zeros_like(mut)
    muv[:jde, :] = 0.5 * (((mu[ja_, :] + mu[jb_, :]) + mub[ja_, :]) + mub[jb_, :])
    pairs = (("u", "u_2", muu, ide, jde - 1), ("v", "v_2", muv, ide - 1, jde), ("t", "t_2", mut, ide - 1, jde - 1),
             ("ph", "ph_2", mut, ide - 1, jde - 1), ("qv", None, mut, ide - 1, jde - 1))
    for name, field, mm, ibe, jbe in pairs:
        fld = B["moist"][0] if field is None else B[field]
        for b in range(1, bw + 1):
            B[f"{name}_bxs"][b - 1].copy_((fld[:, :, b - 1] * mm[:, None, b - 1]).T)
            B[f"{name}_bxe"][b - 1].copy_((fld[:, :, ibe - b] * mm[:, None, ibe - b]).T)
            B[f"{name}_bys"][b - 1].copy_(fld[b - 1] * mm[b - 1][None, :])
            B[f"{name}_bye"][b - 1].copy_(fld[jbe - b] * mm[jbe - b][None, :])
    for b in range(1, bw + 1):
        B["mu_bxs"][b - 1].copy_(mu[:, b - 1])
        B["mu_bxe"][b - 1].copy_(mu[:, ide - 1 - b])
        B["mu_bys"][b - 1].copy_(mu[b - 1, :])
        B["mu_bye"][b - 1].copy_(mu[jde - 1 - b, :])


def footprint_bytes(nx: int, ny: int, nz: int, nmoist: int = 1, nscalar: int = 0) -> int:
    """Device bytes of a Domain's registry at these sizes (scratch excluded)."""
    from .fields import FIELD_SPECS

    d = Dims(nx=nx, ny=ny, nz=nz, nmoist=nmoist, nscalar=nscalar)
    return 4 * sum(math.prod(d.shape(k)) for _, k in FIELD_SPECS)


__all__ = ["device_case", "footprint_bytes"]
