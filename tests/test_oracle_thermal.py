"""CPU checks that pin the thermal oracle (assembly of thermal_solver.cpp:
437-530 and the post-processing of thermal_output.cpp) with textbook
solutions: series resistances in z and x, pure upwind advection (energy
balance cell by cell), the missing inflow term at j = 0, dense solves and
closed forms of the metrics.  No GPU is used.
"""
import numpy as np
import pytest


def stack_of(nx, ny, dz, channel, h=1e-4, k_s=390.0, k_f=0.6, rho=997.0, cp=4180.0, t_in=30.0, qk=1.0, q=1e5,
             k_mode=1):
    dz = np.asarray(dz, dtype=np.float64)
    ch = np.asarray(channel, dtype=np.int32)
    ht = float(dz[ch == 1].sum()) if ch.any() else 1.0
    return dict(nx=nx, ny=ny, dx=h, dy=h, dz=dz, channel=ch, ht=ht, k_s=k_s, k_f=k_f, rho=rho, cp=cp,
                t_in=t_in, qk=qk, k_mode=k_mode, q0=np.full((ny, nx), q))


def ramp(stack, g):
    ks, kf, q = stack["k_s"], stack["k_f"], stack["qk"]
    return kf + (ks - kf) * (1 - g) / (1 + q * g)


def test_ramp_special_values():
    st = stack_of(1, 1, [1e-4], [1])
    assert ramp(st, 0.0) == st["k_s"] and ramp(st, 1.0) == pytest.approx(st["k_f"], rel=1e-15)
    # q_k = 1, gamma = 1/2: k_f + (k_s - k_f) / 3
    assert ramp(st, 0.5) == pytest.approx(st["k_f"] + (st["k_s"] - st["k_f"]) / 3, rel=1e-15)


def test_series_resistance_column(pkg, orc, T):
    """One x-y cell, solid layers under a fluid channel cell at rest: the
    bottom flux Q = q dx dy crosses each layer interface through the series
    resistance dz_k / (2 k_k A) + dz_k+1 / (2 k_k+1 A) and leaves through the
    inlet Dirichlet face, conductance 2 k_f dy dz / dx (thermal_solver.cpp:
    458-466, 509-522).  Centimetre cells keep the reference's diag += 1e-12
    (a leak to 0 degrees) below 1e-9 of the smallest conductance."""
    dz = [5e-3, 8e-3, 3e-3, 1e-2]
    st = stack_of(1, 1, dz, [0, 0, 0, 1], q=3e5, h=1e-2)
    g = np.ones((1, 1))
    z = np.zeros((1, 1))
    o = orc.OracleThermal(T, st, g, z, z)
    Tn = o.direct().ravel()
    A = st["dx"] * st["dy"]
    Q = st["q0"][0, 0] * A
    k = [st["k_s"]] * 3 + [st["k_f"]]
    expect = np.zeros(4)
    expect[3] = st["t_in"] + Q / (2 * st["k_f"] * st["dy"] * dz[3] / st["dx"])
    for kk in (2, 1, 0):
        R = dz[kk] / (2 * k[kk] * A) + dz[kk + 1] / (2 * k[kk + 1] * A)
        expect[kk] = expect[kk + 1] + Q * R
    np.testing.assert_allclose(Tn, expect, rtol=1e-8)


@pytest.mark.parametrize("k_mode", [0, 1, 2])
def test_series_resistance_row_harmonic_faces(pkg, orc, T, k_mode):
    """One row of channel cells at rest, heated from below, Dirichlet at
    the inlet face only: the heat of cells i+1.. crosses face i+1/2 through
    two half-cell resistances in series, which is what the harmonic face
    conductance harm(k_i, k_i+1) A / dx is (thermal_solver.cpp:236, 446-456).
    k(gamma) is checked at gamma in {0, 1} where every law gives k_s / k_f."""
    nx = 9
    st = stack_of(nx, 1, [2e-2], [1], q=2e5, k_mode=k_mode, h=1e-3)
    rng = np.random.default_rng(k_mode)
    g = rng.choice([0.0, 1.0], size=(1, nx))
    g[0, 0] = 1.0  # inlet cell is fluid: it carries the Dirichlet face
    z = np.zeros((1, nx))
    o = orc.OracleThermal(T, st, g, z, z)
    Tn = o.direct().ravel()
    kc = np.where(g[0] > 0.5, st["k_f"], st["k_s"])
    h, dz = st["dx"], st["dz"][0]
    Ax = st["dy"] * dz
    Qc = st["q0"][0, 0] * st["dx"] * st["dy"]
    expect = np.zeros(nx)
    expect[0] = st["t_in"] + nx * Qc / (2 * kc[0] * Ax / h)
    for i in range(nx - 1):
        R = h / (2 * kc[i] * Ax) + h / (2 * kc[i + 1] * Ax)
        expect[i + 1] = expect[i] + Qc * (nx - 1 - i) * R
    np.testing.assert_allclose(Tn, expect, rtol=1e-8)


def test_upwind_advection_energy_balance_along_x(pkg, orc, T):
    """Advection-dominated row (k_f -> 1e-12): each cell's heat is carried
    downstream, T_i = T_in + (i + 1) Q / (rho cp u A) -- inflow enthalpy at
    i = 0 and the outflow term at i = Nx-1 (thermal_solver.cpp:458-490)."""
    nx, U = 7, 0.02
    st = stack_of(nx, 1, [1e-4], [1], k_f=1e-12, q=1e5)
    g = np.ones((1, nx))
    o = orc.OracleThermal(T, st, g, np.full((1, nx), U), np.zeros((1, nx)))
    Tn = o.direct().ravel()
    F = st["rho"] * st["cp"] * U * 1.0 * st["dy"] * st["dz"][0]  # pois = 1 at the middle of a 1-layer channel
    Qc = st["q0"][0, 0] * st["dx"] * st["dy"]
    np.testing.assert_allclose(Tn, st["t_in"] + Qc / F * np.arange(1, nx + 1), rtol=1e-8)
    m = orc.thermal_metrics(st, g, np.full((1, nx), U), Tn.reshape(1, 1, nx))
    assert m["q_out"] == pytest.approx(m["q_in"], rel=1e-6)  # diag += 1e-12 leaks ~4e-8


def test_no_inflow_term_on_the_south_boundary(pkg, orc, T):
    """Flow in +y: the reference has an outflow term at j = Ny-1 but no
    inflow term at j = 0 (thermal_solver.cpp:491-507), so the column is
    heated from 0 degrees: T_j = (j + 1) Q / (rho cp v A), independent of
    T_in (the inlet face of x = 0 is conduction-only and k_f -> 1e-12)."""
    ny, V = 6, 0.01
    st = stack_of(1, ny, [1e-4], [1], k_f=1e-12, q=1e5, t_in=55.0)
    g = np.ones((ny, 1))
    o = orc.OracleThermal(T, st, g, np.zeros((ny, 1)), np.full((ny, 1), V))
    Tn = o.direct().ravel()
    F = st["rho"] * st["cp"] * V * st["dx"] * st["dz"][0]
    Qc = st["q0"][0, 0] * st["dx"] * st["dy"]
    np.testing.assert_allclose(Tn, Qc / F * np.arange(1, ny + 1), rtol=1e-6)


def test_small_stack_dense_solve(pkg, orc, T):
    """4 x 3 x 4 mixed stack: the oracle's BiCGStab and a dense solve of the
    oracle matrix agree; the matrix has the 7-point pattern with an M-matrix
    sign structure (positive diagonal, non-positive off-diagonals)."""
    st = stack_of(4, 3, [5e-5, 5e-5, 3e-5, 3e-5], [0, 0, 1, 1])
    rng = np.random.default_rng(0)
    g = rng.uniform(0, 1, (3, 4))
    o = orc.OracleThermal(T, st, g, 0.1 * rng.uniform(size=(3, 4)), 0.05 * rng.standard_normal((3, 4)))
    A, b = o.matrix()
    D = A.toarray()
    assert np.all(np.diag(D) > 0) and np.all(D - np.diag(np.diag(D)) <= 0)
    assert max(np.diff(A.indptr)) <= 7
    Tb = o.solve(rtol=1e-13, maxit=5000).ravel()
    np.testing.assert_allclose(Tb, np.linalg.solve(D, b), rtol=1e-10)


def test_gradient_metrics_closed_forms(pkg, orc):
    """Gradient pass of compute_thermal_metrics (thermal_output.cpp:63-120,
    270-321).  A linear field 3x + 4y + 12z on a uniform grid: |grad T| = 13
    at every cell, faces included, so the maximum sits at the first cell.
    x^2: central differences are exact inside (2x), the faces are first
    differences (dx at i = 0, (2 nx - 3) dx at i = nx - 1, the maximum).  In
    z, T = k over dz = (1, 2, 4): 1 / dz0 at k = 0, 2 / (dz0 + dz1) at k = 1,
    1 / dz2 at k = 2."""
    nx, ny = 6, 5
    st = stack_of(nx, ny, [0.25] * 4, [0, 0, 1, 1], h=0.5)
    g, u = np.ones((ny, nx)), np.zeros((ny, nx))
    k, j, i = np.indices((4, ny, nx)).astype(np.float64)
    m = orc.thermal_metrics(st, g, u, 3.0 * i * 0.5 + 4.0 * j * 0.5 + 12.0 * k * 0.25)
    for key in ("grad_max", "grad_avg", "grad_max_base", "grad_avg_base"):
        assert m[key] == pytest.approx(13.0, rel=1e-15), key
    assert (m["grad_max_i"], m["grad_max_j"], m["grad_max_k"]) == (0, 0, 0)
    m = orc.thermal_metrics(st, g, u, (i * 0.5) ** 2)
    assert m["grad_max"] == pytest.approx((2 * nx - 3) * 0.5, rel=1e-15)
    assert (m["grad_max_i"], m["grad_max_j"], m["grad_max_k"]) == (nx - 1, 0, 0)
    x = np.arange(nx) * 0.5
    per_col = np.r_[0.5, 2 * x[1:-1], (2 * nx - 3) * 0.5]
    assert m["grad_avg"] == pytest.approx(per_col.mean(), rel=1e-14)
    st = stack_of(2, 2, [1.0, 2.0, 4.0], [0, 1, 1], h=1.0)
    k3 = np.indices((3, 2, 2))[0].astype(np.float64)
    G = orc.grad_magnitude(k3, 1.0, 1.0, st["dz"])
    np.testing.assert_allclose(G[:, 0, 0], [1.0, 2.0 / 3.0, 0.25], rtol=1e-15)
    m = orc.thermal_metrics(st, np.ones((2, 2)), np.zeros((2, 2)), k3)
    assert m["grad_max_base"] == 1.0 and m["grad_avg_base"] == 1.0
    assert m["grad_avg"] == pytest.approx((1.0 * 1 + 2.0 / 3.0 * 2 + 0.25 * 4) / 7, rel=1e-15)


def test_metrics_closed_forms(pkg, orc):
    """Uniform T: every mean is T, std 0, uniformity 1, p-norm T.  Layer 0
    split in two values a, b: p-norm = ((a^p + b^p) / 2)^(1/p)."""
    st = stack_of(4, 2, [5e-5, 5e-5, 5e-5], [0, 1, 1])
    g = np.array([[1.0, 0.0, 1.0, 0.5], [0.0, 1.0, 0.2, 0.0]])
    u = np.zeros((2, 4))
    Tn = np.full((3, 2, 4), 47.5)
    m = orc.thermal_metrics(st, g, u, Tn, 8.0)
    for k in ("t_avg_base", "t_avg_fluid", "t_avg_solid", "t_avg_global", "tb_max", "tb_pnorm"):
        assert m[k] == pytest.approx(47.5, rel=1e-14)
    for k in ("t_std_base", "t_std_global", "t_range_global"):
        assert m[k] == 0.0
    assert m["t_uniformity_base"] == 1.0
    Tn[0, :, :2] = 40.0
    Tn[0, :, 2:] = 60.0
    m = orc.thermal_metrics(st, g, u, Tn, 8.0)
    assert m["tb_pnorm"] == pytest.approx(((40.0 ** 8 + 60.0 ** 8) / 2) ** (1 / 8), rel=1e-14)
    assert m["t_avg_base"] == pytest.approx(50.0, rel=1e-14)
    assert m["t_std_base"] == pytest.approx(10.0, rel=1e-14)
    assert m["t_range_base"] == 20.0
