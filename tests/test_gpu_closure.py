"""GPU tests of the learned eddy viscosity (csrc/k_closure.h).

* network forward and gradient kernels (bf16 MFMA, FP32 accumulate) against
  the FP64 restatement oracle/mlp_ref.py on the same bf16-rounded inputs;
  tolerances: output z relative L2 2e-3 (FP32 vs FP64 sums, bf16 ties of
  the hidden layer), gradient 3e-2 per parameter block (the backward pass
  also rounds its layer deltas to bf16 for the MFMA);
* coarse-graining and the sub-grid stresses of a fine domain against NumPy
  block means of the same fine fields;
* training rows (deformation, Smagorinsky km, features) and the network's km
  inside the step against a host evaluation from the domain's own arrays;
* end to end: a 1 km run of this dycore is the truth; the network trained on
  its coarse-grained stresses explains more of the stresses of a later,
  held-out state at 3 km than Smagorinsky does.
"""

import numpy as np
import pytest

from mlp_ref import bf16_round, forward, loss_and_gradient

pytestmark = pytest.mark.gpu

BLOCKS = {"W1": (0, 4096), "b1": (4096, 4160), "W2": (4160, 8256), "b2": (8256, 8320), "w3": (8320, 8384), "b3": (8384, 8385)}


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _rows(n, seed):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, 64), dtype=np.float32)
    X[:, :55] = rng.standard_normal((n, 55))
    X = bf16_round(X).astype(np.float32)
    T = np.zeros((n, 6), dtype=np.float32)
    T[:, 0] = 0.3 * rng.standard_normal(n)
    T[:, 1] = 0.3 * rng.standard_normal(n)
    T[:, 2] = 1e-3 * rng.standard_normal(n)
    T[:, 3] = 1e-3 * rng.standard_normal(n)
    T[:, 4] = 1.0 + rng.random(n)
    T[:, 5] = 100.0
    return X, T


def _net(gpu, seed=1, kref=150.0):
    from wrfdc.closure import ClosureNorm, Network, init_theta

    theta = init_theta(seed)
    theta[8384] = 0.5
    return Network(ClosureNorm(kref=kref), gpu, theta=theta), theta


def test_mlp_forward_matches_fp64(gpu):
    import torch

    print("[closure] forward 1000 rows vs FP64", flush=True)
    net, theta = _net(gpu)
    X, _ = _rows(1000, 2)
    z = net.z(torch.tensor(X, dtype=torch.bfloat16, device=gpu))
    zr, _ = forward(theta.astype(np.float32).astype(np.float64), X)
    assert _rel(z, zr) < 2e-3, _rel(z, zr)


def test_mlp_gradient_matches_fp64(gpu):
    import torch

    print("[closure] loss gradient over 777 rows vs FP64", flush=True)
    net, theta = _net(gpu)
    X, T = _rows(777, 3)
    lossnorm = 1.0 / 777.0
    loss, g = net.gradient(torch.tensor(X, dtype=torch.bfloat16, device=gpu), torch.tensor(T, device=gpu), lossnorm)
    lr, gr = loss_and_gradient(theta.astype(np.float32).astype(np.float64), X, T, lossnorm, 150.0)
    assert abs(loss - lr) <= 1e-3 * abs(lr), (loss, lr)
    for name, (a, b) in BLOCKS.items():
        e = _rel(g[a:b], gr[a:b])
        assert e < 3e-2, (name, e)


# ---------------------------------------------------------------------------
# data from a fine run
# ---------------------------------------------------------------------------
def _pair(gpu, nxc=24, nz=10, ratio=3, fine_steps=2):
    from wrfdc.closure import truth_pair

    return truth_pair(gpu, nxc=nxc, nz=nz, ratio=ratio, fine_steps=fine_steps)


def test_coarse_grain_block_means(gpu):
    from wrfdc.closure import coarse_grain

    print("[closure] coarse-graining 72x72x10 -> 24x24x10", flush=True)
    fine, coarse = _pair(gpu)
    R, n = 3, 24
    coarse_grain(fine, coarse, R)
    names = ["u_2", "v_2", "t_2", "ph_2", "mu_2", "moist"]
    f = {k: v.astype(np.float64) for k, v in fine.download(names).items()}
    c = coarse.download(names)
    # [j][k][i]: x faces I on fine face (I-1)R+1 averaged over the cell's rows
    u = f["u_2"][: n * R, :, 0: n * R + 1: R]
    u = u.reshape(n, R, u.shape[1], n + 1).mean(axis=1)
    np.testing.assert_allclose(c["u_2"][:n, :, : n + 1], u, rtol=1e-6, atol=1e-6)
    v = f["v_2"][0: n * R + 1: R, :, : n * R]
    v = v.reshape(n + 1, v.shape[1], n, R).mean(axis=3)
    np.testing.assert_allclose(c["v_2"][: n + 1, :, :n], v, rtol=1e-6, atol=1e-6)
    for name in ("t_2", "ph_2"):
        a = f[name][: n * R, :, : n * R]
        a = a.reshape(n, R, a.shape[1], n, R).mean(axis=(1, 4))
        np.testing.assert_allclose(c[name][:n, :, :n], a, rtol=1e-6, atol=1e-4)
    mu = f["mu_2"][: n * R, : n * R].reshape(n, R, n, R).mean(axis=(1, 3))
    np.testing.assert_allclose(c["mu_2"][:n, :n], mu, rtol=1e-6, atol=1e-3)
    q = f["moist"][:, : n * R, :, : n * R]
    q = q.reshape(q.shape[0], n, R, q.shape[2], n, R).mean(axis=(2, 5))
    np.testing.assert_allclose(c["moist"][:, :n, :, :n], q, rtol=1e-6, atol=1e-12)
    fine.close()
    coarse.close()


def test_subgrid_stress_and_rows(gpu):
    import torch

    from wrfdc.closure import ClosureNorm, coarse_grain, training_rows

    print("[closure] sub-grid stresses and training rows", flush=True)
    fine, coarse = _pair(gpu)
    R, nxc, nz = 3, 24, 10
    coarse_grain(fine, coarse, R)
    norm = ClosureNorm(d_shift=-8.0, d_scale=2.0, n_shift=1e-4, n_scale=1e-4, dx_feature=1.1, kref=30.0)
    X, T = training_rows(coarse, norm, fine, R)
    X, T = X.float().cpu().numpy(), T.cpu().numpy()
    f = fine.download(["u_2", "v_2"])
    # stresses from the fine winds at the fine mass points
    u, v = f["u_2"], f["v_2"]
    um = np.float32(0.5) * (u[:, :, :-1][: nxc * R, :nz, : nxc * R] + u[:, :, 1:][: nxc * R, :nz, : nxc * R])
    vm = np.float32(0.5) * (v[: nxc * R, :nz, : nxc * R] + v[1: nxc * R + 1, :nz, : nxc * R])
    um, vm = um.astype(np.float64), vm.astype(np.float64)

    def bm(a):
        return a.reshape(nxc, R, nz, nxc, R).mean(axis=(1, 4))

    t11 = bm(um * um) - bm(um) ** 2
    t22 = bm(vm * vm) - bm(vm) ** 2
    t12 = bm(um * vm) - bm(um) * bm(vm)
    rows_t1 = (0.5 * (t11 - t22))[1:nxc - 1, :, 1:nxc - 1].reshape(-1)  # rows (j, k, i) over i, j in [2, nxc-1]
    rows_t2 = t12[1:nxc - 1, :, 1:nxc - 1].reshape(-1)
    np.testing.assert_allclose(T[:, 0], rows_t1, rtol=1e-4, atol=1e-6 * np.abs(rows_t1).max())
    np.testing.assert_allclose(T[:, 1], rows_t2, rtol=1e-4, atol=1e-6 * np.abs(rows_t2).max())
    # deformation columns and Smagorinsky km from the coarse domain's own arrays
    c = coarse.download(["defor11", "defor22", "defor12", "xkmh"])
    sl = (slice(1, nxc - 1), slice(0, nz), slice(1, nxc - 1))
    d11, d22, d12 = c["defor11"], c["defor22"], c["defor12"]
    s1 = np.float32(0.5) * (d11[sl] - d22[sl])
    s2 = np.float32(0.25) * (((d12[1:nxc - 1, :nz, 1:nxc - 1] + d12[2:nxc, :nz, 1:nxc - 1]) + d12[1:nxc - 1, :nz, 2:nxc]) +
                             d12[2:nxc, :nz, 2:nxc])
    np.testing.assert_array_equal(T[:, 2], s1.reshape(-1))
    np.testing.assert_array_equal(T[:, 3], s2.reshape(-1))
    np.testing.assert_array_equal(T[:, 5], c["xkmh"][sl].reshape(-1))
    assert np.all(T[:, 4] >= 1.0)
    # centre log|D| feature (bf16 of the normalised value)
    dd = d11[sl] - d22[sl]
    def2 = np.float32(0.25) * dd * dd + s2 * s2
    feat = (np.log(np.sqrt(def2).astype(np.float64) + 1e-7) - norm.d_shift) / norm.d_scale
    np.testing.assert_allclose(X[:, 13], feat.reshape(-1), rtol=2 ** -7, atol=1e-3)
    # slots: 32*h + s (stencil point s < 27 of log|D| / N^2), 27 the spacing
    assert np.all(X[:, 27] == np.float32(bf16_round(np.array([1.1]))[0]))
    assert np.all(X[:, 28:32] == 0.0) and np.all(X[:, 59:] == 0.0)
    fine.close()
    coarse.close()


def test_closure_km_in_the_step(gpu):
    """The network's km as the step's first-RK-stage diffusion computes it
    (k_closure_km, features formed in the kernel) against the host evaluation:
    features from the domain's deformation and N^2, the FP64 network on them,
    min(kref*softplus(z), 10 l) / divisor."""
    import torch

    from wrfdc.cases import make_case
    from wrfdc.closure import ClosureNorm, Network, init_theta
    from wrfdc.device import Domain

    print("[closure] network km inside first_rk_step_part2, 28x24x12", flush=True)
    case = make_case(28, 24, 12, nmoist=6, nscalar=0, hill_height=500.0)
    dom = Domain(case.config, gpu)
    dom.upload(case.arrays)
    dom.step(1, 0.0)
    norm = ClosureNorm(d_shift=-7.0, d_scale=1.5, n_shift=1e-4, n_scale=1e-4, dx_feature=float(np.log(3.0)), kref=80.0)
    theta = init_theta(4)
    net = Network(norm, gpu, theta=theta)
    net.attach(dom)
    dom.stage("first_rk_step_part2", 1, 0)
    got = dom.download()
    Network.detach(dom)
    dom.close()
    cfg = case.config
    ide, jde, kde = cfg.nx + 1, cfg.ny + 1, cfg.nz + 1
    d11, d22, d12, bn2 = got["defor11"], got["defor22"], got["defor12"], got["bn2"]
    zx, zy, rdzw = got["zx"], got["zy"], got["rdzw"]
    f32 = np.float32

    def dmag(i, k, j):  # 1-based
        dd = d11[j - 1, k - 1, i - 1] - d22[j - 1, k - 1, i - 1]
        tmp = f32(0.25) * (((d12[j - 1, k - 1, i - 1] + d12[j, k - 1, i - 1]) + d12[j - 1, k - 1, i]) + d12[j, k - 1, i])
        return np.sqrt(f32(0.25) * dd * dd + tmp * tmp)

    pts = [(i, k, j) for j in range(2, jde - 1) for k in range(1, kde) for i in range(2, ide - 1)]
    X = np.zeros((len(pts), 64))
    for r, (i, k, j) in enumerate(pts):
        for g in range(27):
            di, dk, dj = g % 3 - 1, (g // 3) % 3 - 1, g // 9 - 1
            ii, kk, jj = min(max(i + di, 2), ide - 2), min(max(k + dk, 1), kde - 1), min(max(j + dj, 2), jde - 2)
            X[r, g] = (np.log(np.float64(dmag(ii, kk, jj)) + 1e-7) - norm.d_shift) / norm.d_scale
            X[r, 32 + g] = (np.float64(bn2[jj - 1, kk - 1, ii - 1]) - norm.n_shift) / norm.n_scale
        X[r, 27] = norm.dx_feature
    z, _ = forward(theta.astype(np.float32).astype(np.float64), X)
    km = np.zeros(len(pts))
    mlen = np.sqrt(cfg.dx * cfg.dy)
    for r, (i, k, j) in enumerate(pts):
        tzx = 0.25 * (abs(zx[j - 1, k - 1, i - 1]) + abs(zx[j - 1, k - 1, i]) + abs(zx[j - 1, k, i - 1]) + abs(zx[j - 1, k, i])) * rdzw[j - 1, k - 1, i - 1] * cfg.dx
        tzy = 0.25 * (abs(zy[j - 1, k - 1, i - 1]) + abs(zy[j, k - 1, i - 1]) + abs(zy[j - 1, k, i - 1]) + abs(zy[j, k, i - 1])) * rdzw[j - 1, k - 1, i - 1] * cfg.dy
        alpha = max(np.sqrt(tzx * tzx + tzy * tzy), 1.0)
        div = alpha * alpha if dmag(i, k, j) > max(10.0 / mlen, 1e-3) else alpha
        sp = z[r] if z[r] > 20 else np.log1p(np.exp(z[r]))
        km[r] = min(norm.kref * sp, 10.0 * mlen) / div
    gk = np.array([got["xkmh"][j - 1, k - 1, i - 1] for (i, k, j) in pts])
    assert _rel(gk, km) < 2e-3, _rel(gk, km)
    gh = np.array([got["xkhh"][j - 1, k - 1, i - 1] for (i, k, j) in pts])
    assert _rel(gh, km / cfg.prandtl) < 2e-3


def test_trained_closure_beats_smagorinsky(gpu):
    """Truth: a 1 km run (96x96x16, dt 6 s) of this dycore.  Training rows:
    its state after 30 steps coarse-grained to 3 km (32x32x16); held-out
    rows: after 60 steps.  The network trained on the first explains more
    of the held-out sub-grid stress than Smagorinsky: flux_error (the
    unexplained fraction, 1 with no model) is lower."""
    from wrfdc.closure import Network, coarse_grain, fit_norm, flux_error, training_rows

    print("[closure] training on a 1 km run, evaluation at 3 km", flush=True)
    R = 3
    fine, coarse = _pair(gpu, nxc=32, nz=16, ratio=R, fine_steps=30)
    coarse_grain(fine, coarse, R)
    norm = fit_norm(coarse)
    Xa, Ta = training_rows(coarse, norm, fine, R)
    fine.step(30, 30 * fine.config.dt)
    coarse_grain(fine, coarse, R)
    Xb, Tb = training_rows(coarse, norm, fine, R)
    net = Network(norm, gpu, seed=0)
    hist = net.fit(Xa, Ta, epochs=30, batch=1024, lr=2e-3, log=lambda s: print("[closure]", s, flush=True))
    Tb_np = Tb.cpu().numpy()
    e_smag = flux_error(Tb_np[:, 5], Tb_np)
    e_mlp = flux_error(net.km(Xb, Tb_np), Tb_np)
    e_zero = flux_error(np.zeros(len(Tb_np)), Tb_np)
    print(f"[closure] held-out flux error: network {e_mlp:.4f}, Smagorinsky {e_smag:.4f}, no model {e_zero:.4f}; "
          f"training loss {hist[0]:.4f} -> {hist[-1]:.4f}", flush=True)
    assert hist[-1] < hist[0]
    assert e_mlp < e_smag
    # the coarse model runs with the trained network
    net.attach(coarse)
    coarse.step(5, 0.0)
    ck = coarse.checksums()
    assert ck[3] == 0
    fine.close()
    coarse.close()
