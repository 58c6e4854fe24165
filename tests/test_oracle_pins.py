"""Closed-form pins of the density-field terms of the momentum equations, the
Brinkman field and the inlet mask in the SIMPLE oracle.

Each pin states what the reference's formulas fix for a constructed state
and checks the oracle against it, with numbers computed here from the
configuration alone (no oracle internals):

  pin_diagonal_at_rest        u = v = p = 0, uniform gamma < 1, 2.5-D on:
                              aP0 = 2 De + 2 Dn + alpha(gamma) vol
                              + mult 12 mu / Ht^2 vol, dE = hy / aP0,
                              dN = hx / aP0  (momentum_solver.cpp:185-190,
                              221-264; SIMPLE.cpp:751-777)
  pin_convection_factor       uniform u: the west link of a U row is
                              De + cs rho hy u with cs = 1.2 under 2.5-D,
                              1 without (momentum_solver.cpp:221-236)
  pin_pressure_source         u = v = 0, linear p: b = (p_P - p_E) hy for U,
                              (p_P - p_N) hx for V (momentum_solver.cpp:268-276)
  pin_relaxation_and_folds    uniform u (boundary values included), SOU on:
                              A 1 U0 - b = (sink + alpha) vol U0 in every row,
                              which holds only with the relaxation source
                              (1 - a)/a aP0 u and every non-DOF neighbour
                              folded into b (:277, :292-303)
  pin_inlet_mask              hand fixture of boundaries.cpp:82-100: the
                              gamma > 0.5 test on the cell row of each u row,
                              corner rows i = 0 and i = M included
  pin_plug_flow               uniform gamma channel with Brinkman and 2.5-D
                              sinks at its fixed point: the pressure gradient
                              and the velocity profile of the developed flow
                              equal the solution of the 1-D discrete problem
                              (sink + alpha) vol u_i + Dn (2 u_i - u_i+1 - u_i-1)
                              = G hy, sum u_i = (M-1) U_in

test_mutants_are_caught builds the oracle with each planted mistake of its
ORACLE_MUTANT switch (6 mu for 12 mu, the 2.5-D convection factor dropped,
alpha or the sink without vol, the relaxation term without 1/a, the U
pressure source on hx, boundary values not folded, >= for > in the inlet
test, a wrong corner row, q ignored in alpha, dE on hx) and checks that at
least one pin rejects every one of them.
"""
import os
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import scipy.sparse as sp

MU = 8.9e-4


def _cfg(pkg, **kw):
    base = dict(nx=12, ny=8, hx=1e-4, hy=0.5e-4, u_in=0.1, eta=MU, rho=997.0, uv_alpha=0.7)
    base.update(kw)
    return pkg.default_config(**base)


def _alpha(cfg, g):
    amax = cfg.eta / max(cfg.brinkman_kmin, 1e-20)
    return amax * (1.0 - g) / (1.0 + cfg.brinkman_q * g + 1e-10)


def _csr(sysd):
    n = len(sysd["rhs"])
    return sp.csr_matrix((sysd["values"], sysd["col_idx"], sysd["row_ptr"]), shape=(n, n))


def _link(A, sysd, r, di, dj, shape_cols):
    """Matrix entry of row r for the neighbour (i + di, j + dj), or None."""
    ij = sysd["ij"]
    i, j = ij[r]
    lin = {(a, b): k for k, (a, b) in enumerate(map(tuple, ij))}
    k = lin.get((i + di, j + dj))
    return None if k is None else A[r, k]


# ---------------------------------------------------------------------------
def pin_diagonal_at_rest(pkg, make):
    for mult in (1.0, 2.0):
        cfg = _cfg(pkg, enable_25d=1, ht=1e-4, sink_mult_25d=mult, brinkman_q=10.0, brinkman_kmin=MU / 1e8)
        g0 = 0.6
        o = make(cfg, np.full((cfg.ny, cfg.nx), g0))
        vol = cfg.hx * cfg.hy
        De, Dn = cfg.eta * cfg.hy / cfg.hx, cfg.eta * cfg.hx / cfg.hy
        aP0 = 2 * De + 2 * Dn + _alpha(cfg, g0) * vol + mult * 12.0 * cfg.eta / cfg.ht ** 2 * vol
        for comp, dname, h in (("u", "dE", cfg.hy), ("v", "dN", cfg.hx)):
            sysd = o.assemble(comp)
            d = _csr(sysd).diagonal() * cfg.uv_alpha
            assert d == pytest.approx(np.full_like(d, aP0), rel=1e-12), comp
            dd = o.get(dname)[sysd["ij"][:, 0], sysd["ij"][:, 1]]
            assert dd == pytest.approx(np.full_like(dd, h / aP0), rel=1e-12), dname


def pin_convection_factor(pkg, make):
    U0, V0 = 0.05, 0.04
    for on, cs in ((0, 1.0), (1, 1.2)):
        cfg = _cfg(pkg, enable_25d=on, ht=1e-4, conv_factor_25d=1.2)
        M, N = cfg.ny + 1, cfg.nx + 1
        De, Dn = cfg.eta * cfg.hy / cfg.hx, cfg.eta * cfg.hx / cfg.hy
        o = make(cfg, np.ones((cfg.ny, cfg.nx)))
        o.set("u", np.full(o.shape("u"), U0))
        sysd = o.assemble("u")
        A = _csr(sysd).toarray()
        for r, (i, j) in enumerate(sysd["ij"]):
            if 2 <= i <= M - 2 and 2 <= j <= N - 3:
                assert -_link(A, sysd, r, 0, -1, N) - De == pytest.approx(cs * cfg.rho * cfg.hy * U0, rel=1e-12)
                assert -_link(A, sysd, r, 0, 1, N) == pytest.approx(De, rel=1e-12)
                break
        o.set("u", np.zeros(o.shape("u")))
        o.set("v", np.full(o.shape("v"), V0))
        sysd = o.assemble("v")
        A = _csr(sysd).toarray()
        for r, (i, j) in enumerate(sysd["ij"]):
            if 2 <= i <= M - 3 and 2 <= j <= N - 2:
                assert -_link(A, sysd, r, -1, 0, N + 1) - Dn == pytest.approx(cs * cfg.rho * cfg.hx * V0, rel=1e-12)
                assert -_link(A, sysd, r, 1, 0, N + 1) == pytest.approx(Dn, rel=1e-12)
                break


def pin_pressure_source(pkg, make):
    cfg = _cfg(pkg, convection_scheme=1, enable_25d=1, ht=1e-4)
    a, b = 3.0, -2.0
    o = make(cfg, np.full((cfg.ny, cfg.nx), 0.7))
    ii, jj = np.mgrid[0:cfg.ny + 2, 0:cfg.nx + 2]
    o.set("p", a * jj + b * ii)
    su = o.assemble("u")
    assert su["rhs"] == pytest.approx(np.full(len(su["rhs"]), -a * cfg.hy), rel=1e-12)
    sv = o.assemble("v")
    assert sv["rhs"] == pytest.approx(np.full(len(sv["rhs"]), -b * cfg.hx), rel=1e-12)


def pin_relaxation_and_folds(pkg, make):
    U0 = 0.05
    cfg = _cfg(pkg, convection_scheme=1, enable_25d=1, ht=1e-4, brinkman_q=10.0, brinkman_kmin=MU / 1e7)
    g0 = 0.7
    vol = cfg.hx * cfg.hy
    sink = (12.0 * cfg.eta / cfg.ht ** 2 + _alpha(cfg, g0)) * vol
    o = make(cfg, np.full((cfg.ny, cfg.nx), g0))
    for comp, other in (("u", "v"), ("v", "u")):
        o.set(comp, np.full(o.shape(comp), U0))
        o.set(other, np.zeros(o.shape(other)))
        sysd = o.assemble(comp)
        r = _csr(sysd) @ np.full(len(sysd["rhs"]), U0) - sysd["rhs"]
        assert r == pytest.approx(np.full_like(r, sink * U0), rel=1e-9), comp


def pin_inlet_mask(pkg, make):
    cfg = _cfg(pkg, nx=4, ny=5, u_in=0.1)
    g = np.ones((5, 4))
    g[:, 0] = [0.9, 0.2, 0.5, 0.51, 0.0]
    # u rows 0..M (M = 6) read cell rows 0, 0, 1, 2, 3, 4, M-2 = 4
    expect = np.array([1, 1, 0, 0, 1, 0, 0], dtype=float)
    o = make(cfg, g)
    o.apply_bc()
    np.testing.assert_array_equal(o.get("u")[:, 0], expect * cfg.u_in)
    prof = np.linspace(0.3, 0.9, 7)
    o = make(cfg, g, prof)
    o.apply_bc()
    np.testing.assert_array_equal(o.get("u")[:, 0], expect * prof)


def pin_plug_flow(pkg, make):
    cfg = pkg.default_config(nx=60, ny=12, hx=1e-4, hy=1e-4, u_in=0.1, eta=MU, convection_scheme=0, enable_25d=1,
                             ht=1e-4, brinkman_q=1.0, brinkman_kmin=MU / 4e6, pressure_solver=1, p_tol=1e-14,
                             max_pressure_iter=500, momentum_tol=1e-14, max_momentum_iter=400)
    g0 = 0.6
    o = make(cfg, np.full((cfg.ny, cfg.nx), g0))
    o.step(150)
    M = cfg.ny + 1
    vol = cfg.hx * cfg.hy
    Dn = cfg.eta * cfg.hx / cfg.hy
    s = (12.0 * cfg.eta / cfg.ht ** 2 + _alpha(cfg, g0)) * vol
    K = np.diag(np.full(M - 1, s + 2 * Dn)) - Dn * np.eye(M - 1, k=1) - Dn * np.eye(M - 1, k=-1)
    phi = np.linalg.solve(K, np.full(M - 1, cfg.hy))
    G = (M - 1) * cfg.u_in / phi.sum()
    p, u = o.get("p"), o.get("u")
    for j in (25, 30, 35):
        assert p[1:M, j] - p[1:M, j + 1] == pytest.approx(np.full(M - 1, G), rel=1e-7)
        assert u[1:M, j] == pytest.approx(G * phi, rel=1e-7)


def float_mask_field(ny=6, nx=8):
    """gamma whose mask decisions differ between float and double arithmetic:
    bottom-row pairs straddling 0.999f (the float average rounds up to it),
    inlet cells just above 0.5 that round to 0.5f, and outlet-column cells
    just above 1 - 1e-6 that round below 1.0f - 1e-6f."""
    g = np.ones((ny, nx))
    g[0, :] = [0.9989999999, 0.9989999999, 0.99899, 0.99899, 0.9990001, 0.998, 1.0, 0.9989999999][:nx]
    g[1:, 0] = [0.50000000001, 0.6, 0.5, 0.50000001, 0.4][: ny - 1]
    g[1:, nx - 1] = [1.0, 0.99999899, 0.9999990, 0.9999985, 1.0][: ny - 1]
    return g


def float_mask_expect(g, tol=1e-6):
    """The reference's float decisions (numpy float32 arithmetic)."""
    f = g.astype(np.float32)
    south = np.float32(0.5) * (f[0, :-1] + f[0, 1:]) >= np.float32(0.999)  # u rows i = 1, j = 1..N-2
    inlet = f[:, 0] > np.float32(0.5)
    pure = ~(f < np.float32(1.0) - np.float32(tol))
    return south, inlet, pure


def pin_float_masks(pkg, make):
    g = float_mask_field()
    ny, nx = g.shape
    south, inlet, pure = float_mask_expect(g)
    # the fixture is only meaningful where double arithmetic decides otherwise
    assert (south != (0.5 * (g[0, :-1] + g[0, 1:]) >= float(np.float32(0.999)))).any()
    assert (inlet != (g[:, 0] > 0.5)).any()
    assert (pure[:, nx - 1] != (g[:, nx - 1] >= 1.0 - 1e-6)).any()
    cfg = _cfg(pkg, nx=nx, ny=ny, u_in=0.1, dp_fluid_tol=1e-6)
    o = make(cfg, g)
    flags = np.array([o.face("southU", 1, j) for j in range(1, nx)]) != 0
    np.testing.assert_array_equal(flags, south)
    o.apply_bc()
    M = ny + 1
    rows = [0] + list(range(0, M - 1)) + [M - 2]  # cell row of u rows 0..M
    np.testing.assert_array_equal(o.get("u")[:, 0], np.where(inlet[rows], cfg.u_in, 0.0))
    # outlet patch: rows i = 1..M-1 whose cell (i-1, nx-1) is pure fluid
    assert o.patch(False)["area"] == pytest.approx(cfg.hy * pure[:, nx - 1].sum(), rel=1e-12)


PINS = [pin_diagonal_at_rest, pin_convection_factor, pin_pressure_source, pin_relaxation_and_folds,
        pin_inlet_mask, pin_plug_flow, pin_float_masks]


def _maker(orc, T, lib_path=None):
    def make(cfg, gamma, inlet=None):
        return orc.OracleSimple(T, cfg, gamma, inlet, prec=64, lib_path=lib_path)
    return make


@pytest.mark.parametrize("pin", PINS, ids=[p.__name__ for p in PINS])
def test_pin(pkg, orc, T, pin):
    pin(pkg, _maker(orc, T))


# which pins each planted mistake must fail (at least the first one listed)
MUTANTS = {
    1: "pin_diagonal_at_rest",       # 6 mu / Ht^2 sink
    2: "pin_convection_factor",      # 2.5-D convection factor dropped
    3: "pin_diagonal_at_rest",       # alpha without vol
    4: "pin_diagonal_at_rest",       # sink without vol
    5: "pin_relaxation_and_folds",   # (1 - a) aP0 u
    6: "pin_pressure_source",        # U source on hx
    7: "pin_relaxation_and_folds",   # boundary values not folded
    8: "pin_inlet_mask",             # >= 0.5
    9: "pin_inlet_mask",             # corner row i = 0 reads cell row 1
    10: "pin_diagonal_at_rest",      # q ignored in alpha
    11: "pin_diagonal_at_rest",      # dE = hx / aP0
}


def test_mutants_are_caught(pkg, orc, T):
    names = {p.__name__: p for p in PINS}
    with tempfile.TemporaryDirectory() as d:
        with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
            paths = dict(zip(MUTANTS, ex.map(lambda k: orc.build_mutant(k, d), MUTANTS)))
        for k, pin_name in MUTANTS.items():
            make = _maker(orc, T, paths[k])
            with pytest.raises(AssertionError):
                names[pin_name](pkg, make)
            print(f"mutant {k} rejected by {pin_name}")
