"""The smooth optimiser objective of coupled.py on the CPU: its gradient
with respect to T against central differences, its limits (the p-norm range
tends to max - min, the uniformity term is 1 - T_uniformity_base of
thermal_output.cpp:259-267), and the transpose of the cell-centre
velocity average."""
import numpy as np
import pytest


def _stack(pkg):
    return pkg.configs.thermal_stack(nx=6, ny=5, n_base=2, n_channel=2, n_lid=1, seed=1)


def test_f0_smooth_gradient_matches_central_differences(pkg):
    st = _stack(pkg)
    rng = np.random.default_rng(0)
    Tf = 40.0 + 20.0 * rng.uniform(size=(5, 5, 6))
    for wu in (0.0, 0.7):
        J, g, parts = pkg.f0_smooth(Tf, st, w_peak=1.0, w_uniformity=wu, p=8.0)
        assert np.all(g[2:] == 0.0)  # only the two base layers enter
        d = rng.standard_normal(Tf.shape)
        h = 1e-5
        fd = (pkg.f0_smooth(Tf + h * d, st, 1.0, wu, 8.0)[0] - pkg.f0_smooth(Tf - h * d, st, 1.0, wu, 8.0)[0]) / (2 * h)
        assert (g * d).sum() == pytest.approx(fd, rel=1e-8)


def test_f0_smooth_limits(pkg):
    st = _stack(pkg)
    rng = np.random.default_rng(1)
    Tf = 40.0 + 20.0 * rng.uniform(size=(5, 5, 6))
    base = Tf[:2]
    _, _, parts = pkg.f0_smooth(Tf, st, p=4000.0)
    assert parts["range_smooth"] == pytest.approx(base.max() - base.min(), rel=2e-2)
    assert parts["t_avg_base"] == pytest.approx(base.mean(), rel=1e-12)  # equal layer thicknesses
    J, _, parts = pkg.f0_smooth(Tf, st, w_peak=0.0, w_uniformity=1.0)
    assert J == pytest.approx(base.std() / base.mean(), rel=1e-12)


def test_staggered_velocity_vjp_is_the_transpose_of_the_average(pkg):
    nx, ny = 7, 4
    rng = np.random.default_rng(2)
    u = rng.standard_normal(pkg.field_shape("u", nx, ny))
    v = rng.standard_normal(pkg.field_shape("v", nx, ny))
    a, b = rng.standard_normal((ny, nx)), rng.standard_normal((ny, nx))
    uc, vc = pkg.configs.cell_centre_velocity(u, v)
    du, dv = pkg.coupled.staggered_velocity_vjp(a, b, nx, ny)
    assert (a * uc).sum() + (b * vc).sum() == pytest.approx((du * u).sum() + (dv * v).sum(), rel=1e-12)
