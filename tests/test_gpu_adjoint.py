"""GPU tests of the tangent-linear / adjoint scalar transport (csrc/k_adjoint.h).

* the forward (= tangent-linear) window against an FP64 evaluation with the
  oracle's advect_scalar and the rk_update_scalar formula on the same
  recorded dynamics: relative L2 <= 1e-5;
* dot-product test <M dx, y> = <dx, M^T y> (FP64 sums of the FP32 results);
* over a 6-step window of the 200x200x45 case: the cost
  J = 0.5 ||M q0 - q_obs||^2 with q_obs the forward state times
  (1 + 0.05 N(0,1)); the adjoint gradient against central finite
  differences (eps = 1e-3 relative) and against the FP64 evaluation of
  <r, M d>; the dot-product test;
* torch.autograd through TransportFunction gives the adjoint's gradient.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(gpu, name=None, shape=(40, 36, 12), nsteps=2):
    from wrfdc.adjoint import Trajectory, Transport
    from wrfdc.cases import make_case, named_case
    from wrfdc.device import Domain

    case = named_case(name) if name else make_case(*shape, nmoist=1, nscalar=0, hill_height=500.0)
    dom = Domain(case.config, gpu)
    dom.upload(case.arrays)
    traj = Trajectory(dom, nsteps)
    return case, dom, traj, Transport(traj)


def _oracle_window(case, traj, q0):
    """FP64: per stage, the oracle's advect_scalar (t_2 -> t_tend with the
    checkpointed fluxes) and the rk_update_scalar formula."""
    from oracle import Oracle

    cfg = case.config
    A = {k: v.astype(np.float64) for k, v in case.arrays.items()}
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    sz = cfg.spec_zone
    c1 = A["c1h"][:nz][None, :, None]
    c2 = A["c2h"][:nz][None, :, None]
    inner = np.zeros((ny, nz, nx), bool)
    inner[sz:ny - sz, :, sz:nx - sz] = True
    msfty = A["msfty"][:ny, None, :nx]
    mub = A["mub"][:ny, None, :nx]
    dts = {1: cfg.dt / 3.0, 2: 0.5 * cfg.dt, 3: cfg.dt}
    o = Oracle(cfg, A, "f64")
    q_old = np.asarray(q0, np.float64).copy()
    for n in range(traj.nsteps):
        q_in = q_old
        for rk in (1, 2, 3):
            ru, rv, ww, mu1, mu2 = (t.cpu().numpy().astype(np.float64) for t in traj.stages[3 * n + rk - 1])
            o.arrays["t_2"][:] = q_in
            o.arrays["ru"][:] = ru
            o.arrays["rv"][:] = rv
            o.arrays["ww"][:] = ww
            o.arrays["t_tend"][:] = 0.0
            o.stage("advect_scalar", 2, 0)
            tend = o.arrays["t_tend"][:ny, :nz, :nx]
            oldm = c1 * (mu1[:ny, None, :nx] + mub) + c2
            newm = c1 * (mu2[:ny, None, :nx] + mub) + c2
            t = np.where(inner, tend * msfty, 0.0)
            out = q_in.copy()
            out[:ny, :nz, :nx] = (oldm * q_old[:ny, :nz, :nx] + dts[rk] * t) / newm
            q_in = out
        q_old = q_in
    o.close()
    return q_old


def _mass(a, cfg):
    return a[:cfg.ny, :cfg.nz, :cfg.nx]


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_transport_window_matches_fp64(gpu):
    import torch

    print("[adjoint] forward window 40x36x12, 2 steps vs FP64", flush=True)
    case, dom, traj, tr = _setup(gpu)
    q0 = torch.tensor(case.arrays["moist"][0], device=gpu)
    got = tr.forward(q0).cpu().numpy()
    ref = _oracle_window(case, traj, case.arrays["moist"][0])
    cfg = case.config
    assert _rel(_mass(got, cfg), _mass(ref, cfg)) < 1e-5
    dom.close()


def _dot_test(gpu, tr, shape, seed):
    import torch

    rng = np.random.default_rng(seed)
    x = torch.tensor(rng.standard_normal(shape).astype(np.float32), device=gpu)
    y = torch.tensor(rng.standard_normal(shape).astype(np.float32), device=gpu)
    mx = tr.forward(x).cpu().numpy().astype(np.float64)
    aty = tr.adjoint(y).cpu().numpy().astype(np.float64)
    lhs = float(np.sum(mx * y.cpu().numpy()))
    rhs = float(np.sum(x.cpu().numpy() * aty))
    return lhs, rhs


def test_dot_product_small(gpu):
    print("[adjoint] dot-product test 40x36x12, 2 steps", flush=True)
    case, dom, traj, tr = _setup(gpu)
    lhs, rhs = _dot_test(gpu, tr, case.arrays["moist"][0].shape, 11)
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), abs(rhs)), (lhs, rhs)
    dom.close()


def test_autograd_uses_the_adjoint(gpu):
    import torch

    from wrfdc.adjoint import transport_apply

    print("[adjoint] torch.autograd through the transport", flush=True)
    case, dom, traj, tr = _setup(gpu)
    q0 = torch.tensor(case.arrays["moist"][0], device=gpu, requires_grad=True)
    obs = torch.tensor(case.arrays["moist"][0] * 1.01, device=gpu)
    q_end = transport_apply(tr, q0)
    J = 0.5 * ((q_end - obs) ** 2).sum()
    J.backward()
    want = tr.adjoint((tr.forward(q0.detach()) - obs).contiguous())
    assert torch.equal(q0.grad, want)
    dom.close()


@pytest.mark.timeout(600)
def test_gradient_c3_window(gpu):
    """200x200x45, 6 steps: J = 0.5 ||M q0 - q_obs||^2 with q_obs the forward
    state times (1 + 0.05 N(0,1)); gradient M^T (M q0 - q_obs) by the
    adjoint against central differences (eps = 1e-3 relative) and the FP64
    evaluation; dot-product test."""
    import time

    import torch

    print("[adjoint] 200x200x45 case, 6-step window", flush=True)
    case, dom, traj, tr = _setup(gpu, name="c3", nsteps=6)
    cfg = case.config
    q0 = torch.tensor(case.arrays["moist"][0], device=gpu)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    q_end = tr.forward(q0)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    rng = np.random.default_rng(20261202 + 3000)
    q_obs = (q_end.cpu().numpy().astype(np.float64) * (1.0 + 0.05 * rng.standard_normal(q_end.shape))).astype(np.float32)
    obs = torch.tensor(q_obs, device=gpu)
    r = (q_end - obs).contiguous()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    grad = tr.adjoint(r)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    print(f"[adjoint] forward window {1e3 * (t1 - t0):.1f} ms, adjoint window {1e3 * (t3 - t2):.1f} ms", flush=True)
    g = grad.cpu().numpy().astype(np.float64)

    def J(q):
        d = tr.forward(q).cpu().numpy().astype(np.float64) - q_obs
        return 0.5 * float(np.sum(_mass(d, cfg) ** 2))

    # direction: the gradient's direction plus noise, so <grad, d> is not small
    noise = rng.standard_normal(g.shape)
    d = g / np.linalg.norm(g) + 0.5 * noise / np.linalg.norm(noise)
    x = q0.cpu().numpy().astype(np.float64)
    eps = 1e-3 * np.linalg.norm(_mass(x, cfg)) / np.linalg.norm(_mass(d, cfg))
    dt = torch.tensor(d.astype(np.float32), device=gpu)
    fd = (J(q0 + eps * dt) - J(q0 - eps * dt)) / (2.0 * eps)
    an = float(np.sum(g * dt.cpu().numpy().astype(np.float64)))
    print(f"[adjoint] central difference {fd:.6e}, adjoint {an:.6e}", flush=True)
    assert abs(fd - an) <= 2e-3 * abs(an), (fd, an)
    # FP64: <r64, M64 d> with r64 = M64 q0 - q_obs
    r64 = _oracle_window(case, traj, x) - q_obs
    md64 = _oracle_window(case, traj, dt.cpu().numpy())
    an64 = float(np.sum(_mass(r64, cfg) * _mass(md64, cfg)))
    print(f"[adjoint] FP64 <r, M d> {an64:.6e}", flush=True)
    assert abs(an - an64) <= 1e-4 * abs(an64), (an, an64)
    lhs, rhs = _dot_test(gpu, tr, q0.shape, 12)
    print(f"[adjoint] dot-product test {lhs:.9e} vs {rhs:.9e}", flush=True)
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), abs(rhs))
    dom.close()
