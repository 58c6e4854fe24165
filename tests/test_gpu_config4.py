"""The adjoint workload at its own size: d(dP)/d(gamma) of a 512 x 256 SOU
density field (Brinkman, 2.5-D sink, parabolic inlet) at its FP64 fixed
point, and 32 such gradients computed as one batch.

Central differences (h = 1e-6) on 64 seeded cells.  Each perturbed field
is re-solved on the device from the base fixed point (FP64 multigrid PCG,
p_rtol 1e-12, momentum stop 1e-14) to a state residual of ~1e-13, far
below 1e-6 |grad|.  At that level the converged dP still jitters by
~1e-10 Pa from iteration to iteration (round-off of the fixed-point map;
dP ~ 2.6e4 Pa), which alone would put 5e-5 of noise on a difference over
2h = 2e-6.  So after 300 iterations the dP of the next 500 iterations is
averaged (logged on the device by the outer loop, one row per iteration);
its standard error, from 10 batch means, is ~5e-12 Pa.  Cells within 2 of
the domain edges (where the inlet, wall and patch masks act) and cells
with gamma outside (1e-3, 1 - 1e-3) are skipped.  Bars: ||g_adj - g_fd|| <=
1e-5 ||g_fd|| over the 64 cells, and per cell |g_adj - g_fd| <=
1e-5 max(|g_fd|, rms g_fd) + 5 standard errors of that difference.  The
rms floor is the fixed point's own: the FP64 fixed-point equations fix dP
only to a few ulps (2.6e4 Pa x 2.2e-16 = 6e-12 Pa), a path-dependent offset
that averaging over iterations does not remove and that amounts to ~3e-6
over 2h, above 1e-5 |g| for cells whose gradient is well below the rms
(~1.3 here).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cfg(pkg, seed, nx=512, ny=256):
    cfg, gamma, inlet = pkg.configs.density_field(nx=nx, ny=ny, seed=seed)
    cfg.pressure_solver, cfg.precision = pkg._types.CPS_PRESSURE_GMG, 64
    cfg.p_tol, cfg.p_rtol, cfg.max_pressure_iter = 0.0, 1e-12, 60
    cfg.momentum_tol, cfg.max_momentum_iter = 1e-14, 60
    cfg.epsilon, cfg.use_dp_convergence = 0.0, 0
    return cfg, gamma, inlet


def _mean_dp(pkg, cfg, gamma, inlet, w0, settle=300, avg=500):
    c = cfg.copy()
    c.max_iterations = settle + avg
    s = pkg.SIMPLE(c, gamma, inlet)
    for k, a in w0.items():
        s.set(k, a)
    st, hist = s.runIterationsLogged(poll_every=100)
    res = s.state_residual()
    s.close()
    x = hist["dp_full_static"][settle:]
    bm = x.reshape(10, -1).mean(axis=1)
    return x.mean(), bm.std(ddof=1) / np.sqrt(10), float(res.max())


def test_sou_512x256_gradient_matches_central_differences_on_64_cells(pkg):
    cfg, gamma, inlet = _cfg(pkg, seed=11)
    s = pkg.SIMPLE(cfg, gamma, inlet)
    s.step(1000)
    res0 = s.state_residual()
    print("base state residual", res0, flush=True)
    assert res0.max() < 1e-11
    w0 = {k: s.get(k) for k in ("u", "v", "p")}
    dp0, grad = s.gradient_dp(rtol=3e-10, max_iter=3000)
    print("adjoint", s.last_adjoint, flush=True)
    assert s.last_adjoint["converged"] == 1
    s.close()
    ok = (gamma > 1e-3) & (gamma < 1 - 1e-3)
    ok[:, :2] = ok[:, -2:] = False
    ok[:2, :] = ok[-2:, :] = False
    cand = np.argwhere(ok)
    cells = cand[np.random.default_rng(64).choice(len(cand), 64, replace=False)]
    h = 1e-6
    ga, gf, se = [], [], []
    for n, (i, j) in enumerate(cells):
        out = []
        for sgn in (1, -1):
            g2 = gamma.copy()
            g2[i, j] += sgn * h
            out.append(_mean_dp(pkg, cfg, g2, inlet, w0))
        fd = (out[0][0] - out[1][0]) / (2 * h)
        e = np.hypot(out[0][1], out[1][1]) / (2 * h)
        ga.append(grad[i, j])
        gf.append(fd)
        se.append(e)
        print(f"cell {n} ({i},{j}) gamma {gamma[i, j]:.4f} adjoint {grad[i, j]:.10e} fd {fd:.10e} "
              f"rel {abs(fd - grad[i, j]) / abs(fd):.2e} fd-se {e:.1e} residual {max(out[0][2], out[1][2]):.1e}",
              flush=True)
        assert max(out[0][2], out[1][2]) < 1e-11
    ga, gf, se = map(np.array, (ga, gf, se))
    rel = np.linalg.norm(ga - gf) / np.linalg.norm(gf)
    print(f"64 cells: ||adj - fd|| / ||fd|| = {rel:.3e}, max per-cell rel {np.max(np.abs(ga - gf) / np.abs(gf)):.2e}",
          flush=True)
    assert rel <= 1e-5
    floor = np.sqrt(np.mean(gf ** 2))
    assert np.all(np.abs(ga - gf) <= 1e-5 * np.maximum(np.abs(gf), floor) + 5 * se)


def test_32_batched_gradients_match_single_field_gradients(pkg):
    """32 seeded 512 x 256 fields stepped as one batch (one size class: each
    kernel launch covers all 32 fields), their 32 gradients computed
    concurrently, each equal to the gradient of a standalone solver set to
    that field's state (same kernels: bit for bit)."""
    fl = []
    for k in range(32):
        cfg, gamma, inlet = pkg.configs.density_field(nx=512, ny=256, seed=k)
        cfg.pressure_solver, cfg.precision = pkg._types.CPS_PRESSURE_GMG, 64
        cfg.p_tol, cfg.p_rtol, cfg.max_pressure_iter, cfg.max_refine = 0.0, 1e-6, 30, 1
        cfg.momentum_tol, cfg.max_momentum_iter = 1e-13, 30
        fl.append((cfg, gamma, inlet))
    print("case 32 batched gradients 512x256", flush=True)
    b = pkg.Batch(fl)
    b.plan("class", unroll=10)
    assert b.plan_info()["classes"] == 1
    b.step(300)
    b.synchronize()
    res = b.gradients_dp(rtol=1e-8, workers=8)
    for k, (cfg, gamma, inlet) in enumerate(fl):
        dp, grad, st = res[k]
        assert st["converged"] == 1, k
        s = pkg.SIMPLE(cfg, gamma, inlet)
        for name in ("u", "v", "p"):
            s.set(name, b.get(k, name))
        dp1, grad1 = s.gradient_dp(rtol=1e-8)
        s.close()
        assert dp == dp1, k
        np.testing.assert_array_equal(grad, grad1, err_msg=f"field {k}")
    b.close()
