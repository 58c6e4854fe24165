"""The 2-D conduction surrogate adjoint (AdjointMethod/thermal_adjoint.py)
restated in the oracle, pinned by the reference's own outputs
(tests/golden/surrogate_adjoint.npz, written by make_surrogate_golden.py)
and, in its self-consistent mode, by central differences of mean(T)."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "surrogate_adjoint.npz")


def test_external_T_mode_matches_reference_vectors(orc):
    z = np.load(GOLD)
    _, _, sens = orc.surrogate_adjoint(z["ext_k"], z["ext_P"], z["ext_T"])
    np.testing.assert_allclose(sens, z["ext_sens"], rtol=1e-11, atol=1e-14 * np.abs(z["ext_sens"]).max())


def test_default_case_matches_reference_vectors(orc):
    z = np.load(GOLD)
    _, _, sens = orc.surrogate_adjoint(z["def_k"], z["def_P"])
    np.testing.assert_allclose(sens, z["def_sens"], rtol=1e-10, atol=1e-13 * np.abs(z["def_sens"]).max())


def test_self_consistent_mode_is_the_derivative_of_mean_T(orc):
    rng = np.random.default_rng(3)
    k = 0.6 + 400 * rng.uniform(size=(7, 9))
    P = 1e2 * rng.uniform(size=(7, 9))
    _, _, sens = orc.surrogate_adjoint(k, P)
    h = 1e-4
    for (i, j) in ((0, 0), (3, 4), (6, 8), (2, 7)):
        q = []
        for s in (1, -1):
            k2 = k.copy()
            k2[i, j] += s * h * k[i, j]
            q.append(orc.surrogate_adjoint(k2, P)[0].mean())
        fd = (q[0] - q[1]) / (2 * h * k[i, j])
        assert sens[i, j] == pytest.approx(fd, rel=1e-6)
