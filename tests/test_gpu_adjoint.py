"""Discrete adjoint d(dP)/d(gamma) of the converged SIMPLE state on the HIP
path, pinned by brute force: central differences of the ORACLE's converged
pressure drop (each perturbed field re-solved to its fixed point on the CPU).

The oracle's fixed points come from the reference's red-black SOR pressure
solve with a fixed sweep count (pressure_solver.cpp:819-919) and a 1e-14
momentum stop: unlike the Jacobi-PCG, whose |pAp| < 1e-20 stop leaves a
~1e-13 continuity residual, it settles FOU and SOU states to round-off, so
directional differences are limited only by the step (measured on the CPU
at 32 x 16: they converge to the dense oracle adjoint's own 1.2e-7 level
from t = 1e-4 down to 1e-6).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def tight(cfg, scheme):
    cfg.convection_scheme = scheme
    cfg.pressure_solver = 0  # red-black SOR, a fixed 300 sweeps
    cfg.p_tol = 0.0
    cfg.max_pressure_iter = 300
    cfg.momentum_tol = 1e-14
    cfg.max_momentum_iter = 5000
    return cfg


def oracle_fixed_point(orc, T, cfg, g, inlet, iters, init=None):
    o = orc.OracleSimple(T, cfg, g, inlet, prec=64)
    if init is not None:
        for k, a in zip(("u", "v", "p"), init):
            o.set(k, a)
    o.step(iters)
    return o


def dof_list(cfg):
    M, N = cfg.ny + 1, cfg.nx + 1
    Ud = [(i, j) for i in range(1, M) for j in range(1, N - 1)]
    Vd = [(i, j) for i in range(1, M - 1) for j in range(1, N)]
    return [("u", a) for a in Ud] + [("v", a) for a in Vd] + [("p", a) for a in Ud]


def oracle_adjoint_gradient(pkg, orc, T, cfg, g, inlet, base):
    """d(dP)/d(gamma) from the oracle alone: the Jacobian of its assembled
    equations (momentum A x - b at the state, mass residual) and their
    gamma-derivative by central differences, then a dense adjoint solve."""
    dofs = dof_list(cfg)
    idx = {d: k for k, d in enumerate(dofs)}
    w0 = {k: base.get(k) for k in ("u", "v", "p")}

    def vec(rf):
        return np.array([rf[f][a] for f, a in dofs])

    n = len(dofs)
    J = np.zeros((n, n))
    eps = 1e-7
    for c, (f, (i, j)) in enumerate(dofs):
        cols = []
        for sgn in (1, -1):
            a = {k: v.copy() for k, v in w0.items()}
            sc = max(1e-3, abs(a[f][i, j]))
            a[f][i, j] += sgn * eps * sc
            cols.append(vec(base.residual_fields(a["u"], a["v"], a["p"])))
        J[:, c] = (cols[0] - cols[1]) / (2 * eps * sc)
    G = base.get("gamma")
    rows = [i for i in range(1, cfg.ny + 1) if G[i - 1, 0] >= 1 - cfg.dp_fluid_tol]
    gvec = np.zeros(n)
    for i in rows:
        gvec[idx[("p", (i, 1))]] = 1.0 / len(rows)
    lam = np.linalg.solve(J.T, gvec)
    grad = np.zeros((cfg.ny, cfg.nx))
    hg = 1e-6
    for r in range(cfg.ny):
        for c in range(cfg.nx):
            res = []
            for sgn in (1, -1):
                g2 = g.copy()
                g2[r, c] += sgn * hg
                o2 = orc.OracleSimple(T, cfg, g2, inlet, prec=64)
                res.append(vec(o2.residual_fields(w0["u"], w0["v"], w0["p"])))
            grad[r, c] = -lam @ ((res[0] - res[1]) / (2 * hg))
    return grad


@pytest.mark.parametrize("scheme", [0, 1])
def test_dp_gradient_matches_oracle_adjoint(pkg, orc, T, scheme):
    nx, ny = 32, 16
    cfg, g, inlet = pkg.configs.density_field(nx=nx, ny=ny, seed=4, sigma=3.0)
    tight(cfg, scheme)
    base = oracle_fixed_point(orc, T, cfg, g, inlet, 3000)
    s = pkg.SIMPLE(cfg, g, inlet)
    for k in ("u", "v", "p"):
        s.set(k, base.get(k))
    res = s.state_residual()
    print("state residual (U, V, continuity):", res, flush=True)
    dp, grad = s.gradient_dp(rtol=1e-12, max_iter=1000)
    print("adjoint:", s.last_adjoint, flush=True)
    assert s.last_adjoint["converged"] == 1
    assert dp == pytest.approx(base.dp_static(), rel=1e-9)
    ref = oracle_adjoint_gradient(pkg, orc, T, cfg, g, inlet, base)
    # gamma is clamped to [0, 1]: at the bounds the differences are one-sided
    inner = (g > 1e-5) & (g < 1 - 1e-5)
    assert inner.sum() > 0.9 * g.size
    err = np.abs(grad - ref)[inner].max() / np.abs(ref[inner]).max()
    print("max rel err vs oracle adjoint:", err, flush=True)
    assert err < 1e-6
    s.close()


@pytest.mark.parametrize("scheme", [0, 1])
def test_dp_directional_derivative_matches_resolved_oracle(pkg, orc, T, scheme):
    """gradient . d against central differences of the oracle's re-solved
    pressure drop along a random direction d over the interior cells, with
    steps t = 1e-5 and 1e-6: relative difference <= 1e-6 for both."""
    nx, ny = 32, 16
    cfg, g, inlet = pkg.configs.density_field(nx=nx, ny=ny, seed=4, sigma=3.0)
    tight(cfg, scheme)
    base = oracle_fixed_point(orc, T, cfg, g, inlet, 3000)
    w0 = [base.get(k) for k in ("u", "v", "p")]
    s = pkg.SIMPLE(cfg, g, inlet)
    for k, a in zip(("u", "v", "p"), w0):
        s.set(k, a)
    dp, grad = s.gradient_dp(rtol=1e-12, max_iter=1000)
    assert s.last_adjoint["converged"] == 1
    rng = np.random.default_rng(0)
    mask = (g > 0.05) & (g < 0.95)
    mask[:, 0] = False
    mask[0, :] = False
    mask[-1, :] = False
    d = np.where(mask, rng.standard_normal(g.shape), 0.0)
    ad = float((grad * d).sum())
    for t in (1e-5, 1e-6):
        dps = [oracle_fixed_point(orc, T, cfg, g + sgn * t * d, inlet, 1500, init=w0).dp_static() for sgn in (1, -1)]
        fd = (dps[0] - dps[1]) / (2 * t)
        print(f"directional t={t}: adjoint {ad:.12e} fd {fd:.12e} rel {abs(fd - ad) / abs(ad):.2e}", flush=True)
        assert ad == pytest.approx(fd, rel=1e-6)
    s.close()


def test_gradient_of_all_fluid_channel_is_brinkman_free(pkg, orc, T):
    """gamma = 1 everywhere: alpha'(1) = -alpha_max / (1 + q) is finite, the
    flow is symmetric about the channel centre line, so the gradient field
    is mirror-symmetric in y and negative (adding solid raises dP)."""
    cfg, g, _ = pkg.configs.channel_200x50(nx=40, ny=12, brinkman_kmin=1e-9, brinkman_q=1.0)
    tight(cfg, 0)
    base = oracle_fixed_point(orc, T, cfg, g, None, 3000)
    s = pkg.SIMPLE(cfg, g)
    for k in ("u", "v", "p"):
        s.set(k, base.get(k))
    s.step(2)
    dp, grad = s.gradient_dp(rtol=1e-12, max_iter=800)
    assert s.last_adjoint["converged"] == 1
    assert np.all(grad[:, 1:-1] > 0) or np.all(grad[:, 1:-1] < 0)
    np.testing.assert_allclose(grad, grad[::-1], rtol=1e-7, atol=1e-9 * np.abs(grad).max())
    s.close()


def dof_masks(cfg):
    M, N = cfg.ny + 1, cfg.nx + 1
    mu = np.zeros((M + 1, N), bool)
    mu[1:M, 1:N - 1] = True
    mv = np.zeros((M, N + 1), bool)
    mv[1:M - 1, 1:N] = True
    mp = np.zeros((M + 1, N + 1), bool)
    mp[1:M, 1:N - 1] = True
    return {"u": mu, "v": mv, "p": mp}


@pytest.mark.parametrize("scheme", [0, 1])
def test_jacobian_matches_oracle_residual_differences(pkg, orc, T, scheme):
    """J delta from the device Jacobian equals central differences of the
    oracle's assembled equations (momentum A x - b, mass residual) along a
    random direction; J^T is its exact transpose."""
    nx, ny = 32, 16
    cfg, g, inlet = pkg.configs.density_field(nx=nx, ny=ny, seed=4, sigma=3.0)
    tight(cfg, scheme)
    base = oracle_fixed_point(orc, T, cfg, g, inlet, 3000)
    w0 = {k: base.get(k) for k in ("u", "v", "p")}
    s = pkg.SIMPLE(cfg, g, inlet)
    for k, a in w0.items():
        s.set(k, a)
    masks = dof_masks(cfg)
    rng = np.random.default_rng(1)
    scale = {"u": 0.1, "v": 0.1, "p": 100.0}
    d = {k: np.where(masks[k], rng.standard_normal(masks[k].shape) * scale[k], 0.0) for k in masks}
    Jd = s.jacobian_apply(d)
    eps = 1e-7
    rp = base.residual_fields(*(w0[k] + eps * d[k] for k in ("u", "v", "p")))
    rm = base.residual_fields(*(w0[k] - eps * d[k] for k in ("u", "v", "p")))
    for k in ("u", "v", "p"):
        fd = (rp[k] - rm[k]) / (2 * eps)
        m = masks[k]
        err = np.abs(Jd[k][m] - fd[m]).max() / np.abs(fd[m]).max()
        print(k, "max rel err", err, "worst at", np.argwhere(m)[np.argmax(np.abs(Jd[k][m] - fd[m]))], flush=True)
        assert err < 1e-6, k
    x = {k: np.where(masks[k], rng.standard_normal(masks[k].shape), 0.0) for k in masks}
    JTx = s.jacobian_apply(x, transpose=True)
    lhs = sum((JTx[k][masks[k]] * d[k][masks[k]]).sum() for k in masks)
    rhs = sum((x[k][masks[k]] * Jd[k][masks[k]]).sum() for k in masks)
    assert lhs == pytest.approx(rhs, rel=1e-12)
    s.close()


def test_torch_autograd_pressure_drop(pkg):
    import torch

    cfg, g, inlet = pkg.configs.density_field(nx=48, ny=24, seed=2, sigma=3.0)
    gt = torch.tensor(g, requires_grad=True)
    dp = pkg.pressure_drop(gt, cfg, inlet, steps=3000, rtol=1e-10)
    dp.backward()
    s = pkg.SIMPLE(cfg, g, inlet)
    s.step(3000)
    dp2, grad = s.gradient_dp(rtol=1e-10)
    assert float(dp) == pytest.approx(dp2, rel=1e-12)
    np.testing.assert_allclose(gt.grad.numpy(), grad, rtol=1e-8, atol=1e-10 * np.abs(grad).max())
    s.close()


@pytest.mark.parametrize("scheme", [0, 1])
def test_preconditioners_agree_and_lsc_line_needs_fewer_iterations(pkg, scheme):
    """The line / least-squares-commutator preconditioners and the SIMPLE
    block preconditioner solve the same adjoint system: same gradient to the
    solve tolerance, the former in fewer FGMRES iterations."""
    cfg, g, inlet = pkg.configs.density_field(nx=96, ny=48, seed=2, sigma=9.0)
    cfg.convection_scheme = scheme
    cfg.p_tol, cfg.max_pressure_iter = 1e-13, 200
    cfg.momentum_tol, cfg.max_momentum_iter = 1e-13, 200
    s = pkg.SIMPLE(cfg, g, inlet)
    s.step(4000)
    out = {}
    for kind, sweeps in (("simple", 0), ("lsc_line", 4), ("lsc_line", 0), ("lsc_adi", 2)):
        s.adjoint_options(kind, sweeps)
        dp, grad = s.gradient_dp(rtol=1e-9, max_iter=3000)
        print(kind, sweeps, s.last_adjoint, flush=True)
        assert s.last_adjoint["converged"] == 1
        out[(kind, sweeps)] = (dp, grad, s.last_adjoint["iterations"])
    ref = out[("simple", 0)]
    for key, (dp, grad, its) in out.items():
        assert dp == ref[0]
        err = np.abs(grad - ref[1]).max() / np.abs(ref[1]).max()
        assert err < 1e-6, (key, err)
    assert out[("lsc_line", 4)][2] < out[("simple", 0)][2]
    assert out[("lsc_adi", 2)][2] < out[("simple", 0)][2]
    with pytest.raises(pkg._capi.CpsError):
        s.adjoint_options("lsc_line", -1)
    with pytest.raises(pkg._capi.CpsError):
        s.adjoint_options("lsc_adi", 2, omega=1.5)
    s.close()
