"""Seeded synthetic inputs: closed-form properties and the golden digests
written by tests/golden/make_golden.py."""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs.json")


def test_golden_input_digests(pkg):
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(os.path.dirname(GOLDEN),
                                                                               "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    assert mg.digests() == json.load(open(GOLDEN))


def test_projected_field_range_and_endpoints(pkg):
    """The tanh projection maps the min-max normalised field onto [0, 1]
    with the endpoints fixed: 0 -> 0 and 1 -> 1 exactly."""
    g = pkg.configs.projected_random_field(64, 96, seed=3, sigma=4.0)
    assert g.shape == (64, 96)
    assert g.min() == 0.0 and g.max() == 1.0
    # beta = 8 pushes most cells towards 0 or 1
    assert np.mean((g < 0.1) | (g > 0.9)) > 0.3
    np.testing.assert_array_equal(g, pkg.configs.projected_random_field(64, 96, seed=3, sigma=4.0))


def test_parabolic_inlet(pkg):
    prof = pkg.configs.parabolic_inlet(512, 0.5)
    assert prof.shape == (514,)
    r = np.nonzero(prof)[0] - 1  # physical rows
    assert r.min() == 64 and r.max() == 448
    assert prof.max() == 0.5 and prof[256 + 1] == 0.5  # centre row: xi = 1/2
    np.testing.assert_allclose(prof[65:450], prof[65:450][::-1], rtol=0, atol=1e-15)  # symmetric
    xi =(np.arange(64, 449) + 0.5 - 64) / 385
    np.testing.assert_allclose(prof[65:450], 0.5 * 4 * xi * (1 - xi), rtol=1e-15)


def test_power_map_background_and_peaks(pkg):
    q = pkg.configs.power_map(128, 64, seed=0)
    assert q.min() >= 20.0e4  # 20 W/cm^2 background in W/m^2
    assert q.max() <= (20.0 + 8 * 300.0) * 1e4


def test_thermal_stack_layers(pkg):
    st = pkg.configs.thermal_stack()
    assert list(st["channel"]) == [0] * 4 + [1] * 4 + [0] * 4
    np.testing.assert_allclose(st["dz"][4:8].sum(), st["ht"], rtol=1e-15)
    assert st["q0"].shape == (256, 512)


def test_batch_sweep(pkg):
    b = pkg.configs.batch_sweep()
    assert len(b) == 48
    assert {(c.nx, c.ny) for c, _, _ in b} <= set(pkg.configs.BATCH_SIZES)
    for c, g, inlet in b[:4]:
        assert g.shape == (c.ny, c.nx) and inlet.shape == (c.ny + 2,)
        assert c.max_iterations == 500


@pytest.mark.parametrize("nx,ny", [(64, 32)])
def test_density_field_knobs(pkg, nx, ny):
    cfg, g, inlet = pkg.configs.density_field(nx=nx, ny=ny, seed=0)
    mu = cfg.eta
    assert cfg.brinkman_kmin == pytest.approx(mu / 1e8) and cfg.brinkman_q == 10.0
    assert (cfg.enable_25d, cfg.ht, cfg.convection_scheme) == (1, 1e-4, 1)
    # alpha_max (1 - g) / (1 + 10 g) = 1e8 * 0.1 (1 - g) / (0.1 + g)
    amax = mu / cfg.brinkman_kmin
    assert amax == pytest.approx(1e8)
