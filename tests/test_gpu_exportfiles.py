"""run_solvers on an ExportFiles directory written in GeometryGenerator's
formats: the fluid run and the thermal solve on the device, their text
outputs read back and compared with the FP64 oracle on the same inputs."""
import os

import numpy as np
import pytest

from test_cpu_exportfiles import _write_generator_files

pytestmark = pytest.mark.gpu


def test_run_solvers_on_generator_files(pkg, orc, tmp_path):
    xf = pkg.exportfiles
    ny, nx, n_in, n_out = 20, 40, 5, 5
    geo = np.zeros((ny, nx))
    geo[0] = geo[-1] = 1.0
    geo[6:9, 15:22] = 1.0  # a fin in the heat-sink region
    _write_generator_files(tmp_path, geo, U=0.05, dx=1e-4, dy=1e-4, n_in=n_in, n_out=n_out, ht=1e-4, q=20.0)
    d = str(tmp_path)
    over = dict(p_tol=1e-12, max_pressure_iter=500, epsilon=1e-9, max_iterations=120)
    print(f"export case nx={nx} ny={ny}", flush=True)
    fstats, metrics = xf.run_solvers(d, fluid_overrides=over)

    # fluid: the oracle run on the same files
    cfg, gamma, meta = xf.fluid_case(d, **over)
    ref = orc.OracleSimple(pkg._types, cfg, gamma, None, prec=64)
    r = ref.run(history_len=cfg.max_iterations)
    assert fstats["iteration"] == r["iterations"]
    hist = np.loadtxt(os.path.join(d, "pressure_drop_history.txt"), skiprows=1)
    assert hist.shape == (r["iterations"], 5)
    np.testing.assert_allclose(hist[:, 1], r["dp_history"], rtol=2e-5, atol=1e-9)
    res = np.loadtxt(os.path.join(d, "residuals.txt"), skiprows=1)
    assert res.shape == (r["iterations"], 12) and res[-1, 1] == pytest.approx(r["resid_mass"], rel=2e-5)
    uc, vc, pc = xf.cell_fields(ref.get("u"), ref.get("v"), ref.get("p"))
    sl = slice(n_in, nx - n_out)
    for name, want in (("u_thermal.txt", uc[:, sl]), ("v_thermal.txt", vc[:, sl]), ("pressure_thermal.txt", pc[:, sl])):
        got = np.array(xf.read_matrix(os.path.join(d, name)))
        np.testing.assert_allclose(got, want, rtol=2e-5, atol=1e-6 * np.abs(want).max(), err_msg=name)

    # thermal: the oracle's direct solve of the case the files describe
    stack, tg, tu, tv = xf.thermal_case(d)
    Td = orc.OracleThermal(pkg._types, stack, tg, tu, tv).direct()
    Tb = np.array(xf.read_matrix(os.path.join(d, "T_bottom.txt")))
    rise = np.abs(Td - stack["t_in"]).max()
    assert np.abs(Tb - Td[0]).max() <= 1e-5 * rise + 1e-4
    want = orc.thermal_metrics(stack, tg, tu, Td, 10.0)
    parsed = xf.parse_thermal_metrics(os.path.join(d, "thermal_metrics.txt"))
    for key in ("T_avg_base", "T_avg_fluid", "T_avg_global", "T_std_base", "T_range_base"):
        assert parsed[key] == pytest.approx(want[key.lower()], abs=1e-5 * rise + 1e-4), key
    assert parsed["grad_max"] == pytest.approx(want["grad_max"], rel=1e-4)
    assert metrics["t_avg_base"] == pytest.approx(want["t_avg_base"], abs=1e-5 * rise)
    k = np.array(xf.read_matrix(os.path.join(d, "k_thermal.txt")))
    np.testing.assert_allclose(k, xf.k_field(tg, 400.0, 0.6, 1.0), atol=1e-6)

    # the optimiser's adjoint step on the written files, against the
    # oracle's surrogate adjoint of the same (rounded) inputs
    sens, dJ = xf.run_adjoint(d, q_k=3.0)
    Tb, P = np.loadtxt(os.path.join(d, "T_bottom.txt")), np.loadtxt(os.path.join(d, "pressure_thermal.txt"))
    _, _, sref = orc.surrogate_adjoint(k, P, Tb)
    np.testing.assert_allclose(sens, sref, rtol=1e-8, atol=1e-10 * np.abs(sref).max())
    g = 1.0 - np.clip(np.loadtxt(os.path.join(d, "geometry_thermal.txt")), 0.0, 1.0)
    dk = -(400.0 - 0.6) * 4.0 / (1.0 + 3.0 * g) ** 2
    np.testing.assert_allclose(np.loadtxt(os.path.join(d, "dJ_dgamma.txt")), sref * dk, rtol=1e-8,
                               atol=1e-10 * np.abs(sref * dk).max())
    assert xf.objective(parsed, 1.0, 2.0) == pytest.approx(
        parsed["T_avg_base"] + 0.5 * parsed["T_range_base"] + 2.0 * (1 - parsed["T_uniformity_base"]), rel=1e-15)
