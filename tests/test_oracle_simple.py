"""CPU checks that pin the SIMPLE oracle to the reference's behaviour.

The reference ships no tests or golden vectors, so the oracle is pinned by
what the mathematics fixes: closed forms, special cases, invariants, dense
solves and known answers of the reference's own solver loops.  Each check is
chosen so that a plausible slip (wrong index, dropped term, swapped sign)
fails it.  No GPU is used.
"""
import numpy as np
import pytest
import scipy.sparse as sp


def small_cfg(pkg, nx=12, ny=8, **kw):
    cfg = pkg.default_config(nx=nx, ny=ny, hx=1e-4, hy=1e-4, u_in=0.1)
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def csr(sysd):
    n = len(sysd["rhs"])
    return sp.csr_matrix((sysd["values"], sysd["col_idx"], sysd["row_ptr"]), shape=(n, n))


# ---------------------------------------------------------------- a1-a2 ----
def test_alpha_special_values(pkg, orc, T):
    """alpha = alpha_max (1 - g) / (1 + q g + 1e-10) (SIMPLE.cpp:751-777):
    all fluid -> 0 exactly; all solid -> alpha_max; g = 0.5, q = 1 ->
    alpha_max / 3 (the reference's own remark, SIMPLE.h:228)."""
    kmin = 1e-9
    amax = 8.9e-4 / kmin
    for gval, q, expect in ((1.0, 1.0, 0.0), (0.0, 1.0, amax), (0.5, 1.0, amax / 3)):
        cfg = small_cfg(pkg, brinkman_kmin=kmin, brinkman_q=q)
        o = orc.OracleSimple(T, cfg, np.full((cfg.ny, cfg.nx), gval))
        a = o.get("alpha")
        assert a == pytest.approx(np.full_like(a, expect), rel=1e-9, abs=0)


def test_gamma_padding_copies_last_row_and_column(pkg, orc, T):
    """The physical field fills [0..M-2] x [0..N-2]; row M-1 and column N-1
    repeat their neighbours (SIMPLE.cpp:707-729)."""
    cfg = small_cfg(pkg)
    g = np.random.default_rng(0).uniform(size=(cfg.ny, cfg.nx))
    o = orc.OracleSimple(T, cfg, g)
    G = o.get("gamma")
    assert G.shape == (cfg.ny + 1, cfg.nx + 1)
    np.testing.assert_array_equal(G[:-1, :-1], g)
    np.testing.assert_array_equal(G[-1, :-1], g[-1])
    np.testing.assert_array_equal(G[:-1, -1], g[:, -1])
    assert G[-1, -1] == g[-1, -1]


# ---------------------------------------------------------------- a3, a6 ---
def test_face_averages_midpoint_and_out_of_range(pkg, orc, T):
    """Face values are midpoints of the two adjacent cells (exact for a
    field linear in the index); outside the DOF ranges alpha -> 0 and
    gamma -> 1 (SIMPLE.h:642-686, 740-778).  gamma faces feed only mask
    decisions and are formed in float, as the reference's float gamma is."""
    cfg = small_cfg(pkg, nx=10, ny=6)
    M, N = cfg.ny + 1, cfg.nx + 1
    jj = np.arange(cfg.nx) / 16.0
    ii = np.arange(cfg.ny)[:, None] / 32.0
    g = 0.25 + ii + jj[None, :] * 0.5  # linear, exactly representable
    g = g / (2 * g.max())
    o = orc.OracleSimple(T, cfg, g)
    G = o.get("gamma")
    for i in range(1, M):
        for j in range(1, N - 1):  # U faces sit between cells (i-1, j-1) and (i-1, j)
            assert o.face("gammaAtU", i, j) == float(np.float32(0.5) * (np.float32(G[i - 1, j - 1]) + np.float32(G[i - 1, j])))
    for i in range(1, M - 1):
        for j in range(1, N):  # V faces between cells (i-1, j-1) and (i, j-1)
            assert o.face("gammaAtV", i, j) == float(np.float32(0.5) * (np.float32(G[i - 1, j - 1]) + np.float32(G[i, j - 1])))
    for (i, j) in ((0, 3), (M, 3), (2, 0), (2, N - 1)):
        assert o.face("alphaAtU", i, j) == 0.0 and o.face("gammaAtU", i, j) == 1.0
    for (i, j) in ((0, 3), (M - 1, 3), (2, 0), (2, N)):
        assert o.face("alphaAtV", i, j) == 0.0 and o.face("gammaAtV", i, j) == 1.0


def test_wall_half_cell_flags(pkg, orc, T):
    """U rows next to the walls carry the half-cell wall term only where
    the face is fluid (gammaAtU >= 0.999f), i = 1 south and i = M-1 north;
    V-west / V-east are dead (boundaries.cpp:148-206)."""
    cfg = small_cfg(pkg, nx=10, ny=6)
    M, N = cfg.ny + 1, cfg.nx + 1
    g = np.ones((cfg.ny, cfg.nx))
    g[0, 4] = 0.5  # the face (1, 4) and (1, 5) averages drop below 0.999
    o = orc.OracleSimple(T, cfg, g)
    o.apply_bc()
    assert o.face("southU", 1, 2) == 1 and o.face("southU", 1, 4) == 0 and o.face("southU", 2, 2) == 0
    assert o.face("northU", M - 1, 2) == 1 and o.face("northU", M - 2, 2) == 0
    for j in range(1, N):
        assert o.face("westV", 2, j) == 0
        assert o.face("eastV", 2, j) == 0


# ---------------------------------------------------------------- a7 -------
def test_sou_vanishes_on_uniform_field(pkg, orc, T):
    cfg = small_cfg(pkg, convection_scheme=1)
    o = orc.OracleSimple(T, cfg, np.ones((cfg.ny, cfg.nx)))
    u = np.full(o.shape("u"), 2.0)
    o.set("u", u)
    v = np.full(o.shape("v"), 2.0)
    o.set("v", v)
    for (i, j) in ((3, 4), (1, 1), (5, 9)):
        for F in ((1.0, 0.5, 0.3, -0.2), (-1.0, -0.5, -0.3, 0.2)):
            assert o.sou(True, i, j, *F) == 0.0
            assert o.sou(False, i, j, *F) == 0.0


@pytest.mark.parametrize("signs", [(1, 1), (-1, -1), (1, -1), (-1, 1)])
def test_sou_exact_on_linear_field(pkg, orc, T, signs):
    """phi_SOU = 1.5 phi_up - 0.5 phi_upup is exact for a linear field, so
    each face contributes F * (face value - upwind value) = F * (+-a/2):
    Sdc = -Fe (se a/2) + Fw (sw a/2) with s = +1 for F >= 0 (upwind cell
    behind the face) and -1 otherwise (convection.cpp:53-186)."""
    cfg = small_cfg(pkg, nx=14, ny=9, convection_scheme=1)
    o = orc.OracleSimple(T, cfg, np.ones((cfg.ny, cfg.nx)))
    a = 0.25
    uj = np.broadcast_to(a * np.arange(o.shape("u")[1])[None, :], o.shape("u")).copy()
    o.set("u", uj)
    se, sw = signs
    Fe, Fw = 0.7 * se, 0.4 * sw
    i, j = 4, 6  # interior: all stencil guards open
    expect = -Fe * (se * a / 2) + Fw * (sw * a / 2)
    assert o.sou(True, i, j, Fe, Fw, 0.0, 0.0) == pytest.approx(expect, abs=1e-15)
    # linear in i, north / south faces
    ui = np.broadcast_to(a * np.arange(o.shape("u")[0])[:, None], o.shape("u")).copy()
    o.set("u", ui)
    Fn, Fs = 0.7 * se, 0.4 * sw
    expect = -Fn * (se * a / 2) + Fs * (sw * a / 2)
    assert o.sou(True, i, j, 0.0, 0.0, Fn, Fs) == pytest.approx(expect, abs=1e-15)


def test_sou_boundary_guards(pkg, orc, T):
    """Next to the inlet (j = 1) the far-upwind cell j - 2 does not exist:
    the east / west corrections with F >= 0 are skipped (j >= 2 guard)."""
    cfg = small_cfg(pkg, nx=14, ny=9, convection_scheme=1)
    o = orc.OracleSimple(T, cfg, np.ones((cfg.ny, cfg.nx)))
    uj = np.broadcast_to(0.25 * np.arange(o.shape("u")[1])[None, :], o.shape("u")).copy()
    o.set("u", uj)
    assert o.sou(True, 4, 1, 0.7, 0.4, 0.0, 0.0) == 0.0
    assert o.sou(True, 4, 2, 0.7, 0.4, 0.0, 0.0) == pytest.approx((-0.7 + 0.4) * 0.125, abs=1e-15)


# ---------------------------------------------------------------- a8 -------
def test_momentum_operator_at_rest(pkg, orc, T):
    """Zero velocity, all fluid, no sinks: the U operator is the symmetric
    5-point diffusion mu * (4, -1, -1, -1, -1) (hx = hy) relaxed by uvAlpha,
    with the half-cell wall term mu on the diagonal of rows i = 1 and M-1
    (momentum_solver.cpp:195-314)."""
    cfg = small_cfg(pkg, nx=10, ny=6, u_in=0.0)
    mu = cfg.eta
    o = orc.OracleSimple(T, cfg, np.ones((cfg.ny, cfg.nx)))
    sysd = o.assemble("u")
    A = csr(sysd).toarray()
    assert np.abs(A - A.T).max() <= 1e-15 * np.abs(A).max()
    M = cfg.ny + 1
    off = A - np.diag(np.diag(A))
    dofs = sysd["ij"]
    for r, (i, j) in enumerate(dofs):
        wall = 1 if i in (1, M - 1) else 0
        assert A[r, r] * cfg.uv_alpha == pytest.approx(mu * (4 + wall), rel=1e-12)
    assert set(np.unique(np.round(off[off != 0] / mu, 12))) == {-1.0}


def test_momentum_diagonal_dominance_with_flow(pkg, orc, T):
    """With convection, Brinkman and the 2.5-D sink: aP0 >= sum |a_nb|, i.e.
    uvAlpha * diag >= sum of |off-diagonals| in every row."""
    cfg = small_cfg(pkg, nx=16, ny=9, enable_25d=1, ht=1e-4, convection_scheme=1)
    g = pkg.configs.projected_random_field(cfg.ny, cfg.nx, seed=3, sigma=2.0)
    o = orc.OracleSimple(T, cfg, g)
    o.step(3)
    for comp in ("u", "v"):
        A = csr(o.assemble(comp))
        d = A.diagonal()
        offsum = np.asarray(abs(A).sum(axis=1)).ravel() - np.abs(d)
        assert np.all(d * cfg.uv_alpha >= offsum * (1 - 1e-12))


# ---------------------------------------------------------------- a9, a14 --
def test_jacobi_known_answers(pkg, orc, T):
    """solveMomentumJacobi (jacobi_solver.cpp:20-75) tests the pre-update
    residual and counts the sweep that met it: a diagonal system takes 2
    sweeps from x = 0 (exact after the first), 1 from the exact solution."""
    n = 7
    A = sp.diags(np.arange(2.0, 2.0 + n)).tocsr()
    b = np.linspace(1, 2, n)
    x, it = orc.jacobi_csr(T, n, A.indptr, A.indices, A.data, b, np.zeros(n), 1e-6, 100)
    assert it == 2
    np.testing.assert_allclose(x, b / A.diagonal(), rtol=1e-15)
    x, it = orc.jacobi_csr(T, n, A.indptr, A.indices, A.data, b, b / A.diagonal(), 1e-6, 100)
    assert it == 1


def test_jacobi_converges_to_dense_solve(pkg, orc, T):
    rng = np.random.default_rng(2)
    n = 40
    B = sp.random(n, n, density=0.1, random_state=3, format="csr")
    B.data = -np.abs(B.data)
    A = (B + sp.diags(np.asarray(abs(B).sum(axis=1)).ravel() * 1.5 + 1.0)).tocsr()
    A.sort_indices()
    b = rng.standard_normal(n)
    x, it = orc.jacobi_csr(T, n, A.indptr, A.indices, A.data, b, np.zeros(n), 1e-13, 5000)
    assert 2 < it < 5000
    np.testing.assert_allclose(x, np.linalg.solve(A.toarray(), b), rtol=1e-11, atol=1e-12)


def test_pcg_iteration_count_is_distinct_eigenvalues_minus_one(pkg, orc, T):
    """Jacobi-PCG on c I + U W U^T with a constant diagonal converges in as
    many updates as A has distinct eigenvalues; solvePCG_CSR returns the
    loop index of the converging update (pressure_solver.cpp:121-184)."""
    H = np.array([[1, 1, 1, 1], [1, -1, 1, -1], [1, 1, -1, -1], [1, -1, -1, 1]], float)
    H = np.kron(H, H) / 4  # orthogonal, |entries| = 1/4 -> constant diagonal of U U^T
    b = np.random.default_rng(0).standard_normal(16)
    for weights in ([], [5.0], [5.0, 9.0], [5.0, 9.0, 20.0]):
        U = H[:, : len(weights)]
        D = 2 * np.eye(16) + (U * weights) @ U.T
        A = sp.csr_matrix(D)
        x, it = orc.pcg_csr(T, 16, A.indptr, A.indices, A.data, b, np.zeros(16), 1e-9, 100)
        assert it == len(weights)
        np.testing.assert_allclose(x, np.linalg.solve(D, b), rtol=1e-10)


def test_pcg_absolute_tolerance(pkg, orc, T):
    """The stop test is ||r||_2 < tol (absolute, after the update)."""
    n = 30
    A = sp.diags([-1.0, 2.5, -1.0], [-1, 0, 1], shape=(n, n), format="csr")
    b = np.ones(n)
    for tol in (1e-2, 1e-6):
        x, it = orc.pcg_csr(T, n, A.indptr, A.indices, A.data, b, np.zeros(n), tol, 1000)
        assert np.linalg.norm(b - A @ x) < tol
        if it > 0:
            x2, _ = orc.pcg_csr(T, n, A.indptr, A.indices, A.data, b, np.zeros(n), tol, it)
            assert np.linalg.norm(b - A @ x2) >= tol  # one update fewer is not enough


# ---------------------------------------------------------------- a12-a13 --
def neumann_lap(n):
    L = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(n, n)).tolil()
    L[0, 0] = 1.0
    L[n - 1, n - 1] = 1.0
    return L.tocsr()


def test_pressure_matrix_is_scaled_laplacian(pkg, orc, T):
    """Constant dE = dN = d: the pressure operator on DOFs i = 1..M-1,
    j = 1..N-2 is rho d h times the 5-point Laplacian with Neumann N / S / W
    and a Dirichlet (dropped) east neighbour at j = N-1
    (pressure_solver.cpp:425-458, 649-729), built here as a Kronecker sum."""
    cfg = small_cfg(pkg, nx=9, ny=5)
    M, N = cfg.ny + 1, cfg.nx + 1
    o = orc.OracleSimple(T, cfg, np.ones((cfg.ny, cfg.nx)))
    d = 3e-3
    o.set("dE", np.full(o.shape("dE"), d))
    o.set("dN", np.full(o.shape("dN"), d))
    sysd = o.assemble("p")
    A = csr(sysd)
    ni, nj = M - 1, N - 2
    Ly = neumann_lap(ni)
    Lx = neumann_lap(nj).tolil()
    Lx[nj - 1, nj - 1] = 2.0  # east Dirichlet
    ref = cfg.rho * d * cfg.hy * sp.kron(Ly, sp.identity(nj)) + cfg.rho * d * cfg.hy * sp.kron(sp.identity(ni), Lx)
    ij = sysd["ij"]
    order = np.lexsort((ij[:, 1], ij[:, 0]))  # row-major (i outer, j inner)
    assert np.all(order == np.arange(len(order)))
    assert np.abs(A - ref).max() <= 1e-12 * np.abs(ref).max()


def test_pressure_matrix_symmetry_and_row_sums(pkg, orc, T):
    cfg = small_cfg(pkg, nx=11, ny=7, convection_scheme=1)
    g = pkg.configs.projected_random_field(cfg.ny, cfg.nx, seed=1, sigma=2.0)
    o = orc.OracleSimple(T, cfg, g)
    o.step(2)
    sysd = o.assemble("p")
    A = csr(sysd)
    assert abs(A - A.T).max() <= 1e-14 * abs(A).max()
    rs = np.asarray(A.sum(axis=1)).ravel()
    N = cfg.nx + 1
    dE = o.get("dE")
    for r, (i, j) in enumerate(sysd["ij"]):
        expect = cfg.rho * dE[i, j] * cfg.hy if j == N - 2 else 0.0
        assert rs[r] == pytest.approx(expect, abs=1e-12 * abs(A).max())


def test_mass_residual_telescopes_to_boundary_flux(pkg, orc, T):
    """Summing the pressure right side over the DOFs telescopes the discrete
    divergence to the boundary fluxes (pressure_solver.cpp:470-496)."""
    cfg = small_cfg(pkg, nx=10, ny=6)
    M, N = cfg.ny + 1, cfg.nx + 1
    o = orc.OracleSimple(T, cfg, np.ones((cfg.ny, cfg.nx)))
    rng = np.random.default_rng(4)
    us = rng.standard_normal(o.shape("uStar"))
    vs = rng.standard_normal(o.shape("vStar"))
    o.set("uStar", us)
    o.set("vStar", vs)
    b = o.assemble("p")["rhs"]
    expect = cfg.rho * cfg.hy * (us[1:M, 0] - us[1:M, N - 2]).sum() + \
        cfg.rho * cfg.hx * (vs[0, 1:N - 1] - vs[M - 1, 1:N - 1]).sum()
    assert b.sum() == pytest.approx(expect, rel=1e-12)


# ---------------------------------------------------------------- whole ----
def test_developed_channel_poiseuille_band_and_mass_balance(pkg, orc, T):
    """Plane channel, L / H = 10, Re_H = 13 (entrance length ~0.8 mm << L =
    12 mm): the converged pressure drop is within 10 % above the developed
    value 12 mu U L / H^2 (entrance and half-cell wall discretisation add a
    few per cent), and inflow equals outflow."""
    nx, ny, h, U = 120, 12, 1e-4, 0.01
    cfg = small_cfg(pkg, nx=nx, ny=ny, u_in=U, epsilon=1e-9, max_iterations=20000, p_tol=1e-12,
                    max_pressure_iter=1000, momentum_tol=1e-10, max_momentum_iter=2000, use_dp_convergence=0)
    o = orc.OracleSimple(T, cfg, np.ones((ny, nx)))
    r = o.run()
    assert r["error"] <= 1e-9 and r["dp_converged"] == 0
    dp = o.dp_static()
    ana = 12 * cfg.eta * U * (nx * h) / (ny * h) ** 2
    assert 1.0 < dp / ana < 1.10, dp / ana
    flux = o.patch(True)["mass_flux"]
    assert o.patch(False)["mass_flux"] == pytest.approx(flux, rel=1e-8)
    # every x-plane carries the same mass flux: sum_i u(i, j) is the same for
    # each column at a converged state, so any lerp of two columns is too
    for x in (-1.0, 0.3e-3, 5.55e-3, 11.9e-3, 1.0):
        assert o.plane(x)["mass_flux"] == pytest.approx(flux, rel=1e-8), x
    assert r["dp_full_static"] == pytest.approx(dp, rel=1e-12)


def _plane_field(pkg, orc, T):
    """4 x 3 cells, h = 1e-4, cell (1, 2) not pure fluid; p(i, j) = 10 j + i,
    u(i, j) = 0.01 (j + 1) + 0.001 i, v(i, j) = 0.002 j - 0.001 i."""
    cfg = small_cfg(pkg, nx=4, ny=3)
    g = np.ones((3, 4))
    g[1, 2] = 0.5
    o = orc.OracleSimple(T, cfg, g)
    for name, f in (("p", lambda i, j: 10.0 * j + i), ("u", lambda i, j: 0.01 * (j + 1) + 0.001 * i),
                    ("v", lambda i, j: 0.002 * j - 0.001 * i)):
        i, j = np.indices(o.shape(name))
        o.set(name, f(i, j).astype(np.float64))
    return cfg, o


def _expected_plane(cfg, rows):
    """Metrics from hand-interpolated (p, u, v_tangential) per sampled row."""
    area = cfg.hy * len(rows)
    dyn = [0.5 * cfg.rho * (u * u + vt * vt) for _, u, vt in rows]
    fl = [abs(cfg.rho * u * cfg.hy) for _, u, _ in rows]
    tot = sum((p + d) * f for (p, _, _), d, f in zip(rows, dyn, fl)) / sum(fl)
    return dict(area=area, mass_flux=sum(fl), avg_static=sum(p for p, _, _ in rows) / len(rows),
                avg_dynamic=sum(dyn) / len(rows), avg_total=tot, valid=1)


def test_plane_at_x_known_answer(pkg, orc, T):
    """samplePlaneAtX (postprocessing.cpp:62-243) on a hand-set field.
    x = 1.75e-4 lies a quarter of the way from cell centre 1 (1.5e-4) to
    cell centre 2: p lerps columns 2, 3, u and v columns 1, 2 with t = 0.25,
    and row i = 2 is skipped (its cell in column 2 has gamma 0.5).  x beyond
    either end clamps to the first / last cell centre."""
    cfg, o = _plane_field(pkg, orc, T)
    # rows i = 1, 3: p = 22.5 + i, u = 0.0225 + 0.001 i,
    # v_t = mean over rows i-1, min(i, 3) of 0.0025 - 0.001 r
    want = _expected_plane(cfg, [(23.5, 0.0235, 0.002), (25.5, 0.0255, 0.0)])
    got = o.plane(1.75e-4)
    for k, v in want.items():
        assert got[k] == pytest.approx(v, rel=1e-12, abs=1e-18), k
    # x = 1 m -> last centre: cells 2, 3 (row 2 skipped), t = 1: p col 4, u / v col 3
    want = _expected_plane(cfg, [(41.0, 0.041, 0.5 * (0.006 + 0.005)), (43.0, 0.043, 0.5 * (0.004 + 0.003))])
    got = o.plane(1.0)
    for k, v in want.items():
        assert got[k] == pytest.approx(v, rel=1e-12, abs=1e-18), k
    # x = -1 m -> first centre: cells 0, 1 (no row skipped), t = 0: p col 1, u / v col 0
    want = _expected_plane(cfg, [(11.0, 0.011, -0.0005), (12.0, 0.012, -0.0015), (13.0, 0.013, -0.0025)])
    got = o.plane(-1.0)
    for k, v in want.items():
        assert got[k] == pytest.approx(v, rel=1e-12, abs=1e-18), k


def _window_cfg(pkg, **kw):
    """60 x 20 channel whose core dP settles within a 15-iteration window a
    few dozen iterations after a 20-iteration inlet ramp."""
    cfg, g, inl = pkg.configs.channel_200x50(nx=60, ny=20, max_pressure_iter=600, p_tol=1e-14)
    opts = dict(use_dp_convergence=1, inlet_ramp=1, ramp_steps=20, dp_window=15, dp_percent=1.0,
                x_core_in=0.002, x_core_out=0.004, epsilon=1e-12, max_iterations=150)
    opts.update(kw)
    for k, v in opts.items():
        setattr(cfg, k, v)
    return cfg, g, inl


def _first_settled(h, start, win, pct, slope_win=0, slope_tol=None):
    """First iteration k >= start + win - 1 at which the last win core dPs
    have std and |newest - mean| below pct % of max(|mean|, 1) and (with a
    slope window) the least-squares slope of the last slope_win dPs is at
    most slope_tol in magnitude."""
    for k in range(start + win - 1, len(h)):
        w = h[k - win + 1:k + 1]
        m = w.mean()
        den = max(abs(m), 1.0)
        ok = w.std() / den * 100 < pct and abs(w[-1] - m) / den * 100 < pct
        if ok and slope_win and k + 1 >= slope_win:
            s = np.polyfit(np.arange(slope_win), h[k - slope_win + 1:k + 1], 1)[0]
            ok = abs(s) <= slope_tol
        if ok:
            return k
    return None


def test_dp_window_stop_rule(pkg, orc, T):
    """runIterations' dP-window stop (SIMPLE.cpp:441-491): the run ends on
    the first iteration whose window (entries from iteration rampSteps on)
    has settled, checked against the recorded core-dP history."""
    cfg, g, inl = _window_cfg(pkg)
    r = orc.OracleSimple(T, cfg, g, inl).run(history_len=200)
    h = r["dp_history"]
    k = _first_settled(h, cfg.ramp_steps, cfg.dp_window, cfg.dp_percent)
    assert r["dp_converged"] == 1 and r["iterations"] == k + 1 < cfg.max_iterations
    assert r["dp_core_total"] == pytest.approx(h[-1], rel=1e-15)
    # the window ignores everything before rampSteps: a later ramp end
    # delays the stop to the first settled window after it
    cfg2, _, _ = _window_cfg(pkg, inlet_ramp=0, ramp_steps=k + 5)
    r2 = orc.OracleSimple(T, cfg2, g, inl).run(history_len=200)
    k2 = _first_settled(r2["dp_history"], cfg2.ramp_steps, cfg2.dp_window, cfg2.dp_percent)
    assert k2 is not None and r2["iterations"] == k2 + 1 and k2 >= cfg2.ramp_steps + cfg2.dp_window - 1
    # window off: the run goes to max_iterations
    cfg3, _, _ = _window_cfg(pkg, use_dp_convergence=0, max_iterations=k + 10)
    r3 = orc.OracleSimple(T, cfg3, g, inl).run()
    assert r3["iterations"] == k + 10 and r3["dp_converged"] == 0


def test_dp_slope_gate(pkg, orc, T):
    """usePressureDropSlopeGate (SIMPLE.cpp:347-366, 478-483): the window
    also needs |least-squares slope| of the last dpSlopeWindowIters core dPs
    <= dpSlopeMaxDelta / dpSlopeWindowIters; the slope is checked against
    numpy.polyfit on the recorded history."""
    cfg0, g, inl = _window_cfg(pkg)
    k0 = orc.OracleSimple(T, cfg0, g, inl).run()["iterations"] - 1
    # a loose gate changes nothing
    cfg, _, _ = _window_cfg(pkg, use_dp_slope_gate=1, dp_slope_window=10, dp_slope_max_delta=1e6)
    assert orc.OracleSimple(T, cfg, g, inl).run()["iterations"] == k0 + 1
    # a tight gate delays the stop to where the slope has flattened
    cfg, _, _ = _window_cfg(pkg, use_dp_slope_gate=1, dp_slope_window=10, dp_slope_max_delta=2e-3,
                            max_iterations=400)
    r = orc.OracleSimple(T, cfg, g, inl).run(history_len=400)
    k = _first_settled(r["dp_history"], cfg.ramp_steps, cfg.dp_window, cfg.dp_percent, 10, 2e-4)
    assert r["dp_converged"] == 1 and k > k0 and r["iterations"] == k + 1
    # a zero tolerance never passes
    cfg, _, _ = _window_cfg(pkg, use_dp_slope_gate=1, dp_slope_window=10, dp_slope_max_delta=0.0, max_iterations=k0 + 20)
    r = orc.OracleSimple(T, cfg, g, inl).run()
    assert r["dp_converged"] == 0 and r["iterations"] == k0 + 20


def test_inlet_ramp(pkg, orc, T):
    """enableInletRamp (SIMPLE.cpp:381-386): iteration k < rampSteps uses
    inlet speed (k + 1) / rampSteps * targetVel."""
    cfg, g, inl = _window_cfg(pkg, use_dp_convergence=0, max_iterations=3, ramp_steps=10)
    o = orc.OracleSimple(T, cfg, g, inl)
    o.run()
    u = o.get("u")
    np.testing.assert_allclose(u[:, 0], 0.3 * cfg.u_in, rtol=1e-15)
    # with the ramp off the inlet is at full speed from the first iteration
    cfg, g, inl = _window_cfg(pkg, use_dp_convergence=0, max_iterations=3, ramp_steps=10, inlet_ramp=0)
    o = orc.OracleSimple(T, cfg, g, inl)
    o.run()
    np.testing.assert_allclose(o.get("u")[:, 0], cfg.u_in, rtol=1e-15)


def test_boundary_patch_known_answer(pkg, orc, T):
    """sampleBoundaryPatch on a hand-set field (postprocessing.cpp:245-327):
    inlet uses p(i, 0), u(i, 0); outlet p(i, N) = 0 and u(i, N-1); rows whose
    cell is not pure fluid are skipped."""
    cfg = small_cfg(pkg, nx=4, ny=3)
    g = np.ones((3, 4))
    g[2, 0] = 0.5  # inlet row of cell (2, 0) is not pure fluid
    o = orc.OracleSimple(T, cfg, g)
    M, N = 4, 5
    p = np.zeros(o.shape("p"))
    p[:, 0] = [0.0, 3.0, 5.0, 100.0, 0.0]
    o.set("p", p)
    u = np.zeros(o.shape("u"))
    u[1:M, 0] = 0.2
    u[1:M, N - 1] = 0.3
    o.set("u", u)
    inlet = o.patch(True)
    assert inlet["valid"] == 1
    assert inlet["avg_static"] == pytest.approx(4.0)  # rows 1, 2 (row 3's cell is g = 0.5)
    assert inlet["area"] == pytest.approx(2 * cfg.hy)
    outlet = o.patch(False)
    assert outlet["avg_static"] == 0.0


# ---------------------------------------------------------------- a15 ------
def _sor_state(pkg, orc, T, maxit, tol, fluid=False):
    """A small density field (or an all-fluid channel) one SIMPLE step in,
    with the SOR pressure backend (pressureSolverType 0) capped at maxit
    iterations."""
    rng = np.random.default_rng(7)
    nx, ny = 14, 9
    g = np.clip(rng.uniform(-0.3, 1.3, (ny, nx)), 0.0, 1.0)
    g[:, 0] = 1.0
    if fluid:
        g[:] = 1.0
    cfg = small_cfg(pkg, nx=nx, ny=ny, pressure_solver=0, max_pressure_iter=maxit, p_tol=tol)
    o = orc.OracleSimple(T, cfg, g)
    st = o.step(1)
    sysd = o.assemble("p")  # the system of that step (uStar, vStar, dE, dN are unchanged by the corrections)
    return cfg, o, st, csr(sysd), sysd["rhs"], sysd["ij"]


def _red_black_sor(A, b, ij, omega, x, sweeps=1):
    """Textbook SOR on A x = b with red-black ordering by (i + j) % 2:
    x_k += omega ((b_k - sum_{m != k} A_km x_m) / A_kk - x_k); cells of one
    colour do not couple in a 5-point operator, so each colour is one Jacobi
    half-step on the latest values.  Returns x and the last max |change|."""
    d = A.diagonal()
    colour = (ij[:, 0] + ij[:, 1]) % 2
    for _ in range(sweeps):
        mx = 0.0
        for c in (0, 1):
            sel = colour == c
            gs = (b - (A @ x - d * x)) / d
            new = x + omega * (gs - x)
            mx = max(mx, float(np.abs(new - x)[sel].max()))
            x = np.where(sel, new, x)
    return x, mx


def test_oracle_sor_is_red_black_sor_of_the_pressure_matrix(pkg, orc, T):
    """The SOR backend (pressure_solver.cpp:819-919) is red-black SOR on the
    same pressure matrix as the PCG path (pinned above by the Laplacian and
    symmetry tests), from p' = 0, with omega = 2 / (1 + sqrt(1 - cos^2(pi /
    max(M, N)))) clamped to [1, 1.95]; column N-1 stays 0."""
    cfg, o, st, A, b, ij = _sor_state(pkg, orc, T, maxit=1, tol=0.0)
    M, N = cfg.ny + 1, cfg.nx + 1
    sr = np.cos(np.pi / max(M, N))
    omega = min(1.95, max(1.0, 2.0 / (1.0 + np.sqrt(1.0 - sr * sr))))
    x, _ = _red_black_sor(A, b, ij, omega, np.zeros(len(b)))
    ps = o.get("pStar")
    np.testing.assert_allclose(ps[ij[:, 0], ij[:, 1]], x, rtol=1e-12, atol=1e-12 * np.abs(x).max())
    assert st["p_iters"] == 1
    assert np.all(ps[1:M, N - 1] == 0.0)
    # a wrong omega (plain Gauss-Seidel) would not match
    x1, _ = _red_black_sor(A, b, ij, 1.0, np.zeros(len(b)))
    assert np.abs(x1 - x).max() > 1e-3 * np.abs(x).max()


def test_oracle_sor_converges_to_the_direct_solve_and_counts_sweeps(pkg, orc, T):
    """All-fluid channel (p' well inside the +-1e4 clamp): the converged
    sweep is the direct solve, and the count is that of the first sweep whose
    max |change| is below pTol."""
    import scipy.sparse.linalg as sla

    cfg, o, st, A, b, ij = _sor_state(pkg, orc, T, maxit=20000, tol=1e-13, fluid=True)
    x = sla.spsolve(A.tocsc(), b)
    assert np.abs(x).max() < 1e3
    ps = o.get("pStar")
    assert np.abs(ps[ij[:, 0], ij[:, 1]] - x).max() < 1e-9 * np.abs(x).max()
    tol = 1e-4 * np.abs(x).max()
    cfg, o, st, A, b, ij = _sor_state(pkg, orc, T, maxit=20000, tol=tol, fluid=True)
    M, N = cfg.ny + 1, cfg.nx + 1
    sr = np.cos(np.pi / max(M, N))
    omega = min(1.95, max(1.0, 2.0 / (1.0 + np.sqrt(1.0 - sr * sr))))
    xx, k = np.zeros(len(b)), 0
    while True:
        xx, mx = _red_black_sor(A, b, ij, omega, xx)
        k += 1
        if mx < tol:
            break
    assert abs(st["p_iters"] - k) <= 1, (st["p_iters"], k)


def test_oracle_sor_clamps_pressure_correction(pkg, orc, T):
    """Brinkman field whose direct solve exceeds 1e4: the SOR iterate is
    clamped to +-1e4 (pressure_solver.cpp:853) and the sweep stalls there."""
    import scipy.sparse.linalg as sla

    cfg, o, st, A, b, ij = _sor_state(pkg, orc, T, maxit=3000, tol=1e-12)
    x = sla.spsolve(A.tocsc(), b)
    assert np.abs(x).max() > 1e5
    ps = o.get("pStar")[ij[:, 0], ij[:, 1]]
    assert np.abs(ps).max() == 1e4
