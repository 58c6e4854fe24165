"""3-D conjugate-heat-transfer solve on the HIP path against the oracle.

The oracle assembles the reference operator (thermal_solver.cpp:437-530) on
the CPU; scipy's sparse LU of that matrix is the exact solution the
reference's AMGCL solve approximates to rtol 1e-6.  Parity gates:
  - operator: the device stencil equals the oracle CSR entry by entry
    (rel 1e-13: both are FP64 evaluations of the same expressions; the
    compilers may contract different multiply-adds);
  - solve: ||b - A T|| / ||b|| measured with the ORACLE's matrix is below
    rtol, and T matches the direct solve;
  - metrics: the numpy restatement of thermal_output.cpp on the same T;
  - gradient: brute-force central differences of the objective through the
    oracle's direct solve (h = 1e-5).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def synthetic_case(pkg, seed, nx, ny, n_base=2, n_channel=3, n_lid=1, h=5e-5, qk=1.0, k_mode=1, solid_frac=0.3):
    rng = np.random.default_rng(seed)
    stack = pkg.configs.thermal_stack(nx=nx, ny=ny, n_base=n_base, n_channel=n_channel, n_lid=n_lid, h=h, qk=qk,
                                      seed=seed)
    stack["k_mode"] = k_mode
    gamma = rng.uniform(0.05, 1.0, (ny, nx))
    gamma[rng.uniform(size=(ny, nx)) < solid_frac] = 0.0
    u = 0.05 + 0.3 * rng.uniform(size=(ny, nx))
    v = 0.05 * rng.standard_normal((ny, nx))
    return stack, gamma, u, v


def csr_to_stencil(rp, ci, val, nx, ny, nz):
    n = nx * ny * nz
    rows = np.repeat(np.arange(n), np.diff(rp))
    off = ci - rows
    out = np.zeros((7, n))
    for slot, o in enumerate((0, -1, 1, -nx, nx, -nx * ny, nx * ny)):  # d, w, e, s, n, b, t
        sel = off == o
        out[slot, rows[sel]] = val[sel]
    return out.reshape(7, nz, ny, nx)


@pytest.mark.parametrize("seed,nx,ny,k_mode", [(0, 24, 12, 1), (1, 37, 21, 2), (2, 70, 9, 0)])
def test_operator_matches_oracle(pkg, orc, seed, nx, ny, k_mode):
    stack, gamma, u, v = synthetic_case(pkg, seed, nx, ny, k_mode=k_mode)
    s = pkg.ThermalSolver3D(stack, gamma, u, v)
    o = orc.OracleThermal(pkg._types, stack, gamma, u, v)
    rp, ci, val, b = o.csr()
    ref = csr_to_stencil(rp, ci, val, nx, ny, s.nz)
    coef, bg = s.operator()
    scale = np.abs(ref[0]).max()
    np.testing.assert_allclose(coef, ref, rtol=1e-13, atol=1e-13 * scale)
    np.testing.assert_allclose(bg.ravel(), b, rtol=1e-13, atol=1e-13 * np.abs(b).max())
    s.close()


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("seed,nx,ny", [(3, 32, 16), (4, 45, 27)])
def test_solve_matches_direct(pkg, orc, precision, seed, nx, ny):
    stack, gamma, u, v = synthetic_case(pkg, seed, nx, ny)
    s = pkg.ThermalSolver3D(stack, gamma, u, v, precision=precision, rtol=1e-12, max_iter=2000, max_refine=12)
    st = s.solve()
    o = orc.OracleThermal(pkg._types, stack, gamma, u, v)
    Tg = s.T
    Td = o.direct()
    assert st["converged"] == 1, st
    assert o.rel_residual(Tg) < 2e-12, (o.rel_residual(Tg), st)
    rise = np.abs(Td - stack["t_in"]).max()
    assert np.abs(Tg - Td).max() / rise < 1e-8
    s.close()


@pytest.mark.parametrize("seed,nx,ny,layers", [(4, 45, 27, (2, 3, 1)), (7, 512, 256, (4, 4, 4))])
def test_resident_coarse_kernel_is_bit_identical(pkg, seed, nx, ny, layers, monkeypatch):
    """The FP32 V-cycle's one-workgroup coarse end runs with every coarse
    level resident in LDS; CPS_THERMAL_COARSE_GLOBAL=1 keeps the levels in
    global memory.  The mixed solve's temperature field and its iteration
    and refinement counts agree bit for bit."""
    stack, gamma, u, v = synthetic_case(pkg, seed, nx, ny, n_base=layers[0], n_channel=layers[1], n_lid=layers[2],
                                        h=2e-5)
    out = []
    for glob in ("1", "0"):
        monkeypatch.setenv("CPS_THERMAL_COARSE_GLOBAL", glob)
        print(f"case thermal coarse {nx}x{ny} layers={layers} global={glob}", flush=True)
        s = pkg.ThermalSolver3D(stack, gamma, u, v, precision=32)
        st = s.solve()
        out.append((st["iterations"], st["refinements"], s.T))
        s.close()
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2])


def test_reference_tolerance_full_stack(pkg, orc):
    """512 x 256 x 12 stack (4 Cu base, 4 channel, 4 Cu lid) at the
    reference stop rule rtol = 1e-6: the oracle's matrix confirms the
    residual for both precisions, and the metrics match the restatement."""
    nx, ny = 512, 256
    stack, gamma, u, v = synthetic_case(pkg, 7, nx, ny, n_base=4, n_channel=4, n_lid=4, h=2e-5)
    o = orc.OracleThermal(pkg._types, stack, gamma, u, v)
    out = {}
    for prec in (64, 32):
        s = pkg.ThermalSolver3D(stack, gamma, u, v, precision=prec)
        st = s.solve()
        print(f"precision {prec}: {st}", flush=True)
        assert st["converged"] == 1
        T = s.T
        assert o.rel_residual(T) < 1.0001e-6
        m = s.metrics(8.0)
        ref = orc.thermal_metrics(stack, gamma, u, T, 8.0)
        loc = ("grad_max_i", "grad_max_j", "grad_max_k")
        for k, val in ref.items():
            if k not in loc:
                assert m[k] == pytest.approx(val, rel=1e-11, abs=1e-9), k
        # the maximum's cell: |grad T| there is the maximum (a cell within
        # rounding of it may win the tie)
        G = orc.grad_magnitude(T, stack["dx"], stack["dy"], np.asarray(stack["dz"]))
        assert G[m["grad_max_k"], m["grad_max_j"], m["grad_max_i"]] == pytest.approx(ref["grad_max"], rel=1e-13)
        out[prec] = T
        s.close()
    rise = np.abs(out[64] - stack["t_in"]).max()
    assert np.abs(out[32] - out[64]).max() / rise < 1e-3


@pytest.mark.parametrize("precision", [64, 32])
def test_adjoint_solve_matches_direct(pkg, orc, precision):
    """A^T lambda = rhs against the direct transposed solve; the mixed form
    runs FP32 V-cycles inside FP64 BiCGStab and FP64 refinement."""
    stack, gamma, u, v = synthetic_case(pkg, 5, 30, 14)
    s = pkg.ThermalSolver3D(stack, gamma, u, v, precision=precision, rtol=1e-12, max_iter=2000, max_refine=12)
    o = orc.OracleThermal(pkg._types, stack, gamma, u, v)
    rhs = np.random.default_rng(0).standard_normal((s.nz, s.ny, s.nx))
    lam = s.solve_adjoint(rhs)
    ref = o.direct(transpose=True, rhs=rhs)
    assert np.abs(lam - ref).max() / np.abs(ref).max() < 1e-8
    s.close()


@pytest.mark.parametrize("objective,k_mode,precision", [("t_avg_base", 1, 64), ("pnorm", 1, 64), ("t_avg_base", 2, 64),
                                                        ("pnorm", 0, 64), ("pnorm", 1, 32)])
def test_gradient_matches_central_differences(pkg, orc, objective, k_mode, precision):
    nx, ny = 12, 8
    stack, gamma, u, v = synthetic_case(pkg, 11, nx, ny, k_mode=k_mode, qk=3.0)
    s = pkg.ThermalSolver3D(stack, gamma, u, v, precision=precision, rtol=1e-13, max_iter=2000, max_refine=12)
    s.solve()
    J, grad = s.gradient(objective, 8.0)
    o0 = orc.OracleThermal(pkg._types, stack, gamma, u, v)
    assert J == pytest.approx(orc.thermal_objective(stack, gamma, u, o0.direct(), objective, 8.0), rel=1e-10)
    rng = np.random.default_rng(1)
    cand = np.argwhere((gamma > 0.06) & (gamma < 0.95))
    cells = cand[rng.choice(len(cand), 8, replace=False)]
    h = 1e-5
    gscale = np.abs(grad).max()
    for (j, i) in cells:
        Js = []
        for sgn in (1, -1):
            g2 = gamma.copy()
            g2[j, i] += sgn * h
            o = orc.OracleThermal(pkg._types, stack, g2, u, v)
            Js.append(orc.thermal_objective(stack, g2, u, o.direct(), objective, 8.0))
        fd = (Js[0] - Js[1]) / (2 * h)
        assert grad[j, i] == pytest.approx(fd, rel=1e-5, abs=1e-6 * gscale), (j, i, grad[j, i], fd)
    s.close()


def test_torch_autograd_objective(pkg):
    import torch

    stack, gamma, u, v = synthetic_case(pkg, 13, 20, 10)
    s = pkg.ThermalSolver3D(stack, gamma, u, v, rtol=1e-12, max_iter=2000)
    g = torch.tensor(gamma, dtype=torch.float64, requires_grad=True)
    J = pkg.thermal_objective(g, s, "pnorm", 8.0)
    J.backward()
    J2, grad = s.gradient("pnorm", 8.0)
    assert float(J) == pytest.approx(J2, rel=1e-14)
    np.testing.assert_allclose(g.grad.numpy(), grad, rtol=1e-12, atol=0)
    s.close()


def test_thermal_errors_are_reported(pkg):
    stack, gamma, u, v = synthetic_case(pkg, 0, 8, 4)
    with pytest.raises(ValueError):
        pkg.ThermalSolver3D(stack, gamma[:, :3], u, v)
    s = pkg.ThermalSolver3D(stack, gamma, u, v)
    with pytest.raises(pkg._capi.CpsError):
        s.metrics()  # no solve yet
    s.close()


def test_surrogate_adjoint_matches_reference_vectors(pkg):
    """Device version of AdjointMethod/thermal_adjoint.py against the
    reference's own outputs (tests/golden/surrogate_adjoint.npz)."""
    import os

    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "surrogate_adjoint.npz"))
    out = pkg.surrogate_adjoint(z["ext_k"], P=z["ext_P"], T=z["ext_T"], gamma=z["ext_gamma"], rtol=1e-13)
    assert out["stats"]["converged"] == 1
    np.testing.assert_allclose(out["sens"], z["ext_sens"], rtol=1e-8, atol=1e-12 * np.abs(z["ext_sens"]).max())
    g = np.clip(z["ext_gamma"], 0, 1)
    dk = -(400.0 - 0.6) * 2.0 / (1.0 + g) ** 2
    np.testing.assert_allclose(out["dJ_dgamma"], z["ext_sens"] * dk, rtol=1e-8,
                               atol=1e-12 * np.abs(z["ext_sens"] * dk).max())
    out = pkg.surrogate_adjoint(z["def_k"], P=z["def_P"], rtol=1e-13)
    np.testing.assert_allclose(out["sens"], z["def_sens"], rtol=1e-8, atol=1e-12 * np.abs(z["def_sens"]).max())


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("seed,nx,ny", [(5, 32, 16), (6, 45, 27)])
def test_csr_dropin_solves_the_oracle_matrix(pkg, orc, precision, seed, nx, ny):
    """cps_thermal_solve_csr on the oracle's assembled CSR (the
    thermal_solver.cpp:534-552 AMGCL call): from T = T_in to rtol 1e-12 it
    matches the sparse direct solve, and at the reference's rtol 1e-6 the
    oracle's matrix confirms the residual; a matrix with an entry outside
    the 7-point pattern is rejected."""
    stack, gamma, u, v = synthetic_case(pkg, seed, nx, ny)
    o = orc.OracleThermal(pkg._types, stack, gamma, u, v)
    rp, ci, val, b = o.csr()
    nz = len(stack["dz"])
    T0 = np.full(nx * ny * nz, stack["t_in"])
    print(f"case thermal csr nx={nx} ny={ny} precision={precision}", flush=True)
    T, (its, rr) = pkg.thermal.solve_csr(nx, ny, nz, rp, ci, val, b, T0, rtol=1e-12, max_iter=3000,
                                         precision=precision)
    Td = o.direct().ravel()
    rise = np.abs(Td - stack["t_in"]).max()
    print(f"  iterations {its} rel residual {rr:.2e}", flush=True)
    assert rr <= 1e-12 and o.rel_residual(T.reshape(nz, ny, nx)) < 2e-12
    assert np.abs(T - Td).max() / rise < 1e-8
    T6, (its6, rr6) = pkg.thermal.solve_csr(nx, ny, nz, rp, ci, val, b, T0, rtol=1e-6, max_iter=1000,
                                            precision=precision)
    assert rr6 <= 1e-6 and o.rel_residual(T6.reshape(nz, ny, nx)) < 1.0001e-6
    bad_ci = ci.copy()
    r0 = 3
    bad_ci[rp[r0]] = (r0 + 2) % (nx * ny * nz)  # not a grid neighbour of row 3
    with pytest.raises(pkg._capi.CpsError):
        pkg.thermal.solve_csr(nx, ny, nz, rp, bad_ci, val, b, T0)
