"""The coupled optimiser objective f0_smooth(gamma) (coupled.py: SIMPLE fixed
point -> cell-centred velocity -> 3-D thermal solve -> smooth f0) and its
device gradient (thermal adjoint, velocity sensitivities, flow adjoint with
that right side), against central differences of the same pipeline run
entirely through the oracle: the oracle's SIMPLE restatement taken to its
fixed point (red-black SOR, 300 sweeps; momentum stop 1e-14) and the
oracle's thermal assembly solved directly (sparse LU).

Three checks along seeded directions over the interior cells:
  * the whole pipeline (flow re-solved):             total gradient, 1e-6;
  * velocity frozen at the base state:               thermal part, 1e-7;
  * the thermal solve alone along a velocity change: dJ/du_c, dJ/dv_c, 1e-7.
The flow part is a sizeable share of the total (printed), so a gradient
that ignored the flow, as the reference's does, fails the first check."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import rel_l2  # noqa: F401

pytestmark = pytest.mark.gpu

NX, NY = 32, 16


def _problem(pkg):
    cfg, g, inlet = pkg.configs.density_field(nx=NX, ny=NY, seed=4, sigma=3.0)
    stack = pkg.configs.thermal_stack(nx=NX, ny=NY, n_base=2, n_channel=2, n_lid=1, seed=2)
    return cfg, g, inlet, stack


def _oracle_flow(pkg, orc, cfg, g, inlet, iters, init=None):
    c = cfg.copy()
    c.pressure_solver, c.p_tol, c.max_pressure_iter = 0, 0.0, 300
    c.momentum_tol, c.max_momentum_iter = 1e-14, 5000
    o = orc.OracleSimple(pkg._types, c, g, inlet, prec=64)
    if init is not None:
        for k, a in init.items():
            o.set(k, a)
    o.step(iters)
    return o


def _oracle_T(pkg, orc, stack, g, uc, vc):
    th = orc.OracleThermal(pkg._types, stack, g, uc, vc)
    rp, ci, val, b = th.csr()
    A = sp.csr_matrix((val, ci, rp), shape=(len(b), len(b)))
    return spla.spsolve(A.tocsc(), b).reshape(len(stack["dz"]), NY, NX)


def test_coupled_f0_gradient_matches_oracle_pipeline(pkg, orc):
    cfg, g, inlet, stack = _problem(pkg)
    kw = dict(w_peak=1.0, w_uniformity=0.5, p=8.0)
    dcfg = cfg.copy()
    dcfg.pressure_solver, dcfg.precision = pkg._types.CPS_PRESSURE_GMG, 64
    dcfg.p_tol, dcfg.p_rtol, dcfg.max_pressure_iter = 0.0, 1e-12, 60
    dcfg.momentum_tol, dcfg.max_momentum_iter = 1e-14, 60
    print("case coupled f0 32x16", flush=True)
    obj = pkg.CoupledObjective(dcfg, inlet, stack, flow_steps=1500, **kw)
    J = obj.evaluate(g)
    grad, parts = obj.gradient()
    print("device J", J, "flow adjoint", parts["flow_adjoint"], flush=True)
    assert parts["flow_adjoint"]["converged"] == 1

    base = _oracle_flow(pkg, orc, cfg, g, inlet, 3000)
    w0 = {k: base.get(k) for k in ("u", "v", "p")}
    uc0, vc0 = pkg.configs.cell_centre_velocity(w0["u"], w0["v"])
    T0 = _oracle_T(pkg, orc, stack, g, uc0, vc0)
    Jr = pkg.f0_smooth(T0, stack, **kw)[0]
    print("oracle J", Jr, flush=True)
    assert J == pytest.approx(Jr, rel=1e-10)

    rng = np.random.default_rng(7)
    mask = (g > 0.05) & (g < 0.95)
    mask[:, 0] = mask[0, :] = mask[-1, :] = False
    d = np.where(mask, rng.standard_normal(g.shape), 0.0)

    def J_pipeline(gg, frozen=False):
        if frozen:
            uc, vc = uc0, vc0
        else:
            o = _oracle_flow(pkg, orc, cfg, gg, inlet, 1500, init=w0)
            uc, vc = pkg.configs.cell_centre_velocity(o.get("u"), o.get("v"))
        return pkg.f0_smooth(_oracle_T(pkg, orc, stack, gg, uc, vc), stack, **kw)[0]

    t = 1e-5
    fd = (J_pipeline(g + t * d) - J_pipeline(g - t * d)) / (2 * t)
    ad, ad_th, ad_flow = (grad * d).sum(), (parts["thermal"] * d).sum(), (parts["flow"] * d).sum()
    print(f"total: adjoint {ad:.12e} fd {fd:.12e} rel {abs(ad - fd) / abs(fd):.2e}; "
          f"flow share {abs(ad_flow) / abs(ad):.3f}", flush=True)
    assert ad == pytest.approx(fd, rel=1e-6)
    assert abs(ad_th - fd) > 1e-3 * abs(fd)  # the flow term is not negligible
    fdf = (J_pipeline(g + t * d, True) - J_pipeline(g - t * d, True)) / (2 * t)
    print(f"frozen velocity: thermal part {ad_th:.12e} fd {fdf:.12e}", flush=True)
    assert ad_th == pytest.approx(fdf, rel=1e-7)
    # velocity sensitivities alone: J(T(gamma, u_c + s e_u, v_c + s e_v))
    eu, ev = rng.standard_normal(uc0.shape), rng.standard_normal(vc0.shape)
    s = 1e-6
    Jp = pkg.f0_smooth(_oracle_T(pkg, orc, stack, g, uc0 + s * eu, vc0 + s * ev), stack, **kw)[0]
    Jm = pkg.f0_smooth(_oracle_T(pkg, orc, stack, g, uc0 - s * eu, vc0 - s * ev), stack, **kw)[0]
    fdv = (Jp - Jm) / (2 * s)
    adv = (parts["dJ_du_centre"] * eu).sum() + (parts["dJ_dv_centre"] * ev).sum()
    print(f"velocity: adjoint {adv:.12e} fd {fdv:.12e}", flush=True)
    assert adv == pytest.approx(fdv, rel=1e-7)
    obj.close()


def test_coupled_f0_torch_autograd(pkg):
    import torch

    cfg, g, inlet, stack = _problem(pkg)
    cfg.pressure_solver, cfg.precision = pkg._types.CPS_PRESSURE_GMG, 64
    cfg.p_tol, cfg.p_rtol, cfg.max_pressure_iter = 0.0, 1e-12, 60
    cfg.momentum_tol, cfg.max_momentum_iter = 1e-14, 60
    obj = pkg.CoupledObjective(cfg, inlet, stack, flow_steps=1500, w_uniformity=0.5)
    gt = torch.tensor(g, dtype=torch.float64, requires_grad=True)
    J = pkg.coupled_f0(gt, obj)
    J.backward()
    grad, _ = obj.gradient()
    assert float(J.detach()) == pytest.approx(obj.J, rel=1e-14)
    np.testing.assert_allclose(gt.grad.numpy(), grad, rtol=1e-9, atol=1e-12 * np.abs(grad).max())
    obj.close()
