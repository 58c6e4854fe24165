"""Edge cases of the HIP path against the oracle: the smallest grids under
each pressure backend, a one-layer and a one-column thermal stack, and a
one-field batch.

Tolerances: SOR and Jacobi-PCG follow the oracle's arithmetic step by step
(fields to round-off, rel. L2 <= 1e-10); the GMG backend solves the same
pressure system to 1e-13 absolute, so its fields agree with the oracle's
PCG to that solve tolerance."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

SIZES = [(2, 2), (3, 2), (2, 5), (5, 3), (7, 4)]


@pytest.mark.parametrize("solver", [0, 1, 5])
@pytest.mark.parametrize("nx,ny", SIZES)
def test_smallest_grids_match_oracle(pkg, orc, solver, nx, ny):
    cfg, g, inlet = pkg.configs.density_field(nx=nx, ny=ny, seed=nx * 10 + ny, sigma=1.0)
    cfg.pressure_solver = solver
    # SOR stops on max |change| < pTol: at 1e-13 that test sits at round-off
    # and the stop sweep could differ by a sweep or two; 1e-9 is far above it
    cfg.p_tol, cfg.p_rtol, cfg.max_pressure_iter = (1e-9 if solver == 0 else 1e-13), 0.0, 500
    print(f"case nx={nx} ny={ny} pressure_solver={solver}", flush=True)
    gpu = pkg.SIMPLE(cfg, g, inlet)
    ocfg = cfg.copy()
    ocfg.pressure_solver = 0 if solver == 0 else 1
    ref = orc.OracleSimple(pkg._types, ocfg, g, inlet, prec=64)
    sg, sr = gpu.step(5), ref.step(5)
    tol = 1e-10 if solver != 5 else 1e-7
    for name in ("u", "v", "p"):
        assert rel_l2(gpu.get(name), ref.get(name)) <= tol, name
    assert gpu.full_static_drop() == pytest.approx(ref.dp_static(), rel=tol, abs=1e-12)
    if solver != 5:
        assert abs(sg["p_iters"] - sr["p_iters"]) <= 1
    gpu.close()


def _stack_case(pkg, nx, ny, n_base, n_channel, n_lid, seed):
    rng = np.random.default_rng(seed)
    stack = pkg.configs.thermal_stack(nx=nx, ny=ny, n_base=n_base, n_channel=n_channel, n_lid=n_lid, seed=seed)
    gamma = rng.uniform(0.0, 1.0, (ny, nx))
    u = 0.05 + 0.3 * rng.uniform(size=(ny, nx))
    v = 0.05 * rng.standard_normal((ny, nx))
    return stack, gamma, u, v


@pytest.mark.parametrize("nx,ny,n_base,n_channel,n_lid", [(6, 4, 0, 1, 0), (1, 1, 2, 2, 1), (1, 5, 1, 2, 1),
                                                          (9, 1, 1, 1, 1)])
def test_thermal_degenerate_stacks_match_direct_solve(pkg, orc, nx, ny, n_base, n_channel, n_lid):
    """One layer (the z lines are single cells), one column (the multigrid
    hierarchy is the coarsest level alone), one-row and one-column grids.
    The metrics (an axis of one cell has no gradient; no base layers leave
    the base maxima at -1e30) match the restatement on the solved field."""
    stack, gamma, u, v = _stack_case(pkg, nx, ny, n_base, n_channel, n_lid, seed=nx + 7 * ny)
    print(f"thermal case nx={nx} ny={ny} nz={len(stack['dz'])}", flush=True)
    for prec in (64, 32):
        s = pkg.ThermalSolver3D(stack, gamma, u, v, precision=prec, rtol=1e-12, max_iter=500, max_refine=12)
        st = s.solve()
        assert st["converged"] == 1, st
        o = orc.OracleThermal(pkg._types, stack, gamma, u, v)
        Td = o.direct()
        assert o.rel_residual(s.T) < 2e-12
        rise = max(np.abs(Td - stack["t_in"]).max(), 1e-12)
        assert np.abs(s.T - Td).max() / rise < 1e-8
        T = s.T
        m, ref = s.metrics(8.0), orc.thermal_metrics(stack, gamma, u, T, 8.0)
        for k, val in ref.items():
            if not k.startswith("grad_max_") or k == "grad_max_base":
                assert m[k] == pytest.approx(val, rel=1e-11, abs=1e-9), k
        G = orc.grad_magnitude(T, stack["dx"], stack["dy"], np.asarray(stack["dz"]))
        assert G[m["grad_max_k"], m["grad_max_j"], m["grad_max_i"]] == pytest.approx(ref["grad_max"], rel=1e-13)
        s.close()


def test_one_field_batch_equals_single_solver(pkg):
    cfg, g, inlet = pkg.configs.density_field(nx=40, ny=20, seed=3, sigma=3.0)
    T = pkg._types
    cfg.pressure_solver, cfg.precision = T.CPS_PRESSURE_GMG, T.CPS_PREC_MIXED
    cfg.p_tol, cfg.p_rtol, cfg.max_pressure_iter, cfg.max_refine = 0.0, 1e-3, 6, 1
    cfg.max_momentum_iter = 8
    b = pkg.Batch([(cfg, g, inlet)])
    b.tune(streams=1)
    b.step(7)
    b.synchronize()
    n = b.stats(0)["iteration"]  # includes the batch's own eager set-up step
    s = pkg.SIMPLE(cfg, g, inlet)
    s.step(n)
    for name in ("u", "v", "p"):
        assert np.array_equal(b.get(0, name), s.get(name)), name
    b.close()
    s.close()
