"""ExportFiles text formats (no GPU): the readers on files written the way
GeometryGenerator_V6.py:507-518 writes them, the reference's reader quirks
(loadTopology's fluid default, the dx / dy order of thermal_params.txt, the
power-map shape test), and the writers' layouts and number formats
(saveMatrix's default float format, the fixed-point metric files, the
'T_range_global' line without a space that run_solvers.py's parser skips).
"""
import os

import numpy as np
import pytest


@pytest.fixture
def xf(pkg):
    return pkg.exportfiles


def _write_generator_files(d, geometry, U=0.5, dx=2e-5, dy=2e-5, n_in=3, n_out=2, ht=1e-4, q=20.0):
    """fluid_params.txt / geometry_fluid.txt / thermal_params.txt /
    geometry_thermal.txt in GeometryGenerator_V6.py's formats."""
    ny, nx = geometry.shape
    with open(os.path.join(d, "fluid_params.txt"), "w") as f:
        f.write(f"{ny} {nx} {dy:.9f} {dx:.9f} {U} {n_in} {n_out} {ht:.9f}")
    np.savetxt(os.path.join(d, "geometry_fluid.txt"), geometry, fmt="%.6f", delimiter="\t")
    geo_t = geometry[:, n_in:nx - n_out]
    with open(os.path.join(d, "thermal_params.txt"), "w") as f:
        f.write(f"{ny} {geo_t.shape[1]} {dy:.9f} {dx:.9f} {q * 10000.0} {ht:.9f}")
    np.savetxt(os.path.join(d, "geometry_thermal.txt"), geo_t, fmt="%.6f", delimiter="\t")
    return geo_t


def test_fluid_case_from_generator_files(xf, tmp_path):
    rng = np.random.default_rng(0)
    geo = np.round(rng.uniform(-0.2, 1.2, (6, 11)), 6)
    _write_generator_files(tmp_path, geo)
    cfg, gamma, meta = xf.fluid_case(str(tmp_path), enable_25d=1)
    assert (cfg.nx, cfg.ny, cfg.u_in) == (11, 6, 0.5)
    assert cfg.hx == pytest.approx(2e-5, rel=1e-12) and cfg.ht == pytest.approx(1e-4, rel=1e-12)
    assert (meta["n_in_buffer"], meta["n_out_buffer"]) == (3, 2)
    assert cfg.enable_25d == 1
    # gamma = 1 - clamp(cellType): the values outside [0, 1] saturate
    np.testing.assert_array_equal(gamma, 1.0 - np.clip(geo, 0.0, 1.0))
    assert gamma.min() == 0.0 and gamma.max() == 1.0
    # other knobs keep the SIMPLE.h defaults
    dflt = xf.T.default_config()
    assert (cfg.uv_alpha, cfg.p_alpha, cfg.max_iterations, cfg.use_dp_convergence) == (
        dflt.uv_alpha, dflt.p_alpha, dflt.max_iterations, dflt.use_dp_convergence)


def test_fluid_params_optional_ht_and_2p5d_rule(xf, tmp_path):
    """Ht is optional (SIMPLE.cpp:55-59); 2.5-D with Ht <= 0 is switched off
    (SIMPLE.cpp:66-72); extra columns are ignored."""
    with open(tmp_path / "fluid_params.txt", "w") as f:
        f.write("4 5 0.0001 0.0002 0.1 0 0")
    cfg, gamma, meta = xf.fluid_case(str(tmp_path), enable_25d=1)
    assert cfg.ht == 0.0 and cfg.enable_25d == 0
    assert (cfg.hy, cfg.hx) == (0.0001, 0.0002)
    np.testing.assert_array_equal(gamma, 1.0)  # no geometry file: all fluid
    with open(tmp_path / "fluid_params.txt", "w") as f:
        f.write("4 5 0.0001 0.0002 0.1 0 0 0.0003 997 1e-6")
    cfg, _, _ = xf.fluid_case(str(tmp_path))
    assert cfg.ht == 0.0003 and cfg.rho == 997.0
    with open(tmp_path / "fluid_params.txt", "w") as f:
        f.write("4 0 0.0001 0.0002 0.1 0 0")
    with pytest.raises(ValueError):
        xf.fluid_case(str(tmp_path))


def test_geometry_short_and_long_files(xf, tmp_path):
    """loadTopology (SIMPLE.cpp:651-667): values fill rows in order; missing
    trailing entries stay fluid; extra entries are ignored."""
    p = tmp_path / "g.txt"
    p.write_text("1 1 1\n0.5 1")
    ct = xf.read_celltype(str(p), 2, 3)
    np.testing.assert_array_equal(ct, [[1, 1, 1], [0.5, 1, 0]])
    p.write_text("1 0 1 0 1 0 1 1 1")
    np.testing.assert_array_equal(xf.read_celltype(str(p), 2, 3), [[1, 0, 1], [0, 1, 0]])


def test_thermal_case_params_order_power_map_and_stack(xf, tmp_path):
    """thermal_params.txt is read as 'rows cols dx dy q Ht' although the
    generator writes dy before dx; power_map.txt (W/cm^2, '#' comments and
    blank lines skipped) replaces the uniform flux only when its shape
    matches; the stack is Params' 2 base + 7 channel layers."""
    geo = np.zeros((4, 9))
    geo[0] = geo[-1] = 1.0
    geo_t = _write_generator_files(tmp_path, geo, dx=3e-5, dy=2e-5, n_in=2, n_out=1, ht=7e-4, q=20.0)
    stack, gamma, u, v = xf.thermal_case(str(tmp_path), qk=3.0, k_mode=2)
    assert (stack["nx"], stack["ny"]) == (6, 4)
    assert stack["dx"] == pytest.approx(2e-5) and stack["dy"] == pytest.approx(3e-5)  # swapped, as read
    np.testing.assert_allclose(stack["q0"], 20.0 * 1e4)
    np.testing.assert_array_equal(gamma, 1.0 - geo_t)
    np.testing.assert_allclose(stack["dz"], [2.5e-4] * 2 + [1e-4] * 7)
    assert list(stack["channel"]) == [0, 0] + [1] * 7 and stack["qk"] == 3.0 and stack["k_mode"] == 2
    assert not u.any() and not v.any()  # no velocity files yet
    pm = np.arange(24, dtype=float).reshape(4, 6)
    with open(tmp_path / "power_map.txt", "w") as f:
        f.write("# W/cm^2\n\n")
        for row in pm:
            f.write(" ".join(str(x) for x in row) + "\n")
    stack, _, _, _ = xf.thermal_case(str(tmp_path))
    np.testing.assert_allclose(stack["q0"], pm * 1e4)
    np.savetxt(tmp_path / "power_map.txt", pm[:, :5])
    stack, _, _, _ = xf.thermal_case(str(tmp_path))
    np.testing.assert_allclose(stack["q0"], 20.0 * 1e4)


def test_save_matrix_default_float_format(xf, tmp_path):
    """saveMatrix writes floats as operator<< does: 6 significant digits,
    shortest of fixed / exponent form, tab-separated, one row per line."""
    p = tmp_path / "m.txt"
    xf.save_matrix(str(p), [[1.0, 0.5, 1e-7, 123456789.0], [-0.0, 2.5e-5, 100000.0, 1234567.0]])
    assert p.read_text() == "1\t0.5\t1e-07\t1.23457e+08\n-0\t2.5e-05\t100000\t1.23457e+06\n"
    np.testing.assert_allclose(xf.read_matrix(str(p)), [[1, 0.5, 1e-7, 1.23457e8], [0, 2.5e-5, 1e5, 1.23457e6]])


def test_cell_fields_and_pressure_gradient(xf):
    """Cell centres average the two faces of each cell (output.cpp:98-120);
    |grad p| of a linear p is exact everywhere, edges included (one-sided
    differences of a linear field are exact)."""
    ny, nx = 4, 6
    M, N = ny + 1, nx + 1
    i, j = np.indices((M + 1, N))
    u = 1.0 * j + 10.0 * i
    i, j = np.indices((M, N + 1))
    v = 2.0 * i + 0.5 * j
    i, j = np.indices((M + 1, N + 1))
    p = 3.0 * j - 4.0 * i
    uc, vc, pc = xf.cell_fields(u, v, p)
    ci, cj = np.indices((ny, nx))
    np.testing.assert_array_equal(uc, cj + 0.5 + 10.0 * (ci + 1))
    np.testing.assert_array_equal(vc, 2.0 * ci + 1.0 + 0.5 * (cj + 1))
    np.testing.assert_array_equal(pc, 3.0 * (cj + 1) - 4.0 * (ci + 1))
    g = xf.pressure_gradient(pc, 0.5, 0.25)
    np.testing.assert_allclose(g, np.hypot(3.0 / 0.5, 4.0 / 0.25), rtol=1e-15)


class _FakeFluid:
    """Fields of a solver for the writers (no device needed)."""

    def __init__(self, pkg, ny, nx):
        self.cfg = pkg.default_config(nx=nx, ny=ny, hx=1e-4, hy=2e-4)
        rng = np.random.default_rng(1)
        self.f = {n: rng.standard_normal(pkg.field_shape(n, nx, ny)) for n in ("u", "v", "p")}

    def get(self, n):
        return self.f[n]


def test_fluid_outputs_layout(pkg, xf, tmp_path):
    """saveAll's text files: staggered shapes, full-domain cell fields and
    the heat-sink crop (N_in_buffer columns skipped, N_out_buffer dropped)."""
    ny, nx = 5, 12
    s = _FakeFluid(pkg, ny, nx)
    n_th = xf.save_fluid_outputs(str(tmp_path), s, dict(n_in_buffer=3, n_out_buffer=2))
    assert n_th == 7
    shape = lambda n: np.array(xf.read_matrix(str(tmp_path / n))).shape  # noqa: E731
    assert shape("u.txt") == (ny + 2, nx + 1) and shape("v.txt") == (ny + 1, nx + 2) and shape("p.txt") == (ny + 2, nx + 2)
    assert shape("u_full.txt") == (ny, nx) and shape("pressure_thermal.txt") == (ny, 7)
    uc, _, pc = xf.cell_fields(s.f["u"], s.f["v"], s.f["p"])
    np.testing.assert_allclose(xf.read_matrix(str(tmp_path / "u_thermal.txt")), uc[:, 3:10], rtol=1e-5, atol=1e-6)
    gp = xf.pressure_gradient(pc, 1e-4, 2e-4)
    np.testing.assert_allclose(xf.read_matrix(str(tmp_path / "p_gradient.txt")), gp[:, 3:10], rtol=1e-5)


def test_thermal_metrics_file_and_reference_parser(xf, tmp_path):
    """thermal_metrics.txt line for line (4 decimals, fixed), and the
    run_solvers.py parser's view of it: 'T_range_global' has no space
    before its value and is skipped; grad_max_loc reads as its i index."""
    keys = [k for k, _ in xf.T.ThermalMetrics._fields_ if not k.startswith("grad_max_") or k == "grad_max_base"]
    m = {k: 10.0 + n + 0.123456 for n, k in enumerate(keys)}
    m.update(grad_max_i=7, grad_max_j=8, grad_max_k=9)
    xf.save_thermal_metrics(str(tmp_path), m)
    lines = (tmp_path / "thermal_metrics.txt").read_text().splitlines()
    assert lines[0] == "# Thermal Metrics - Volume-Weighted Statistics"
    assert lines[4] == "T_avg_base    %.4f" % m["t_avg_base"]
    assert lines[17] == "T_range_global%.4f" % m["t_range_global"]
    assert lines[-1] == "grad_max_loc      7 8 9"
    parsed = xf.parse_thermal_metrics(str(tmp_path / "thermal_metrics.txt"))
    assert "T_range_global" not in parsed and "T_range_global%.4f" % m["t_range_global"] not in parsed
    assert parsed["T_avg_base"] == round(m["t_avg_base"], 4) and parsed["grad_max_loc"] == 7.0
    assert parsed["T_uniformity_global"] == round(m["t_uniformity_global"], 4)
    assert len(parsed) == 16  # 17 value lines, T_range_global unreadable


def test_optimization_outputs_and_k_field(xf, tmp_path):
    """objective.txt / dJ_dgamma.txt (10 decimals, the p line '10') and the
    RAMP k field at the clamped gamma (k_s at gamma 0, k_f at gamma 1)."""
    xf.save_optimization_outputs(str(tmp_path), dict(tb_max=81.25, tb_pnorm=79.5, imbalance_pct=-0.01),
                                 np.array([[1.5, -2.0]]))
    assert (tmp_path / "objective.txt").read_text() == "81.2500000000\n79.5000000000\n10\n-0.0100000000\n"
    assert (tmp_path / "dJ_dgamma.txt").read_text() == "1.5000000000\t-2.0000000000\n"
    k = xf.k_field(np.array([[-0.5, 0.0, 0.5, 1.0, 2.0]]), 400.0, 0.6, 1.0)
    np.testing.assert_allclose(k, [[400.0, 400.0, 0.6 + 399.4 * 0.5 / 1.5, 0.6, 0.6]], rtol=1e-15)


def test_iteration_logs_layout(xf, tmp_path):
    hist = {k: np.array([1.0, 0.5]) * (n + 1) for n, k in enumerate(
        ("resid_mass", "resid_u", "resid_v", "dp_core_total", "dp_full_total", "dp_core_static", "dp_full_static"))}
    xf.save_iteration_logs(str(tmp_path), hist)
    r = (tmp_path / "residuals.txt").read_text().splitlines()
    assert r[0].startswith("Iter MassRMS U_RMS V_RMS") and r[2] == "1 0.5 1 1.5 0 0 0 0 2 2.5 0 0"
    d = (tmp_path / "pressure_drop_history.txt").read_text().splitlines()
    assert d[1] == "0 4 5 6 7"
