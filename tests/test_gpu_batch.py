"""Batched fields through one captured hipGraph per batch iteration.

Every field of a batch must equal the same field stepped eagerly on its own
(identical kernels in identical order per field, so bit for bit), after
each of several graph replays separated by eager work, for every stream
count the autotuner may pick.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def fields(pkg, precision):
    T = pkg._types
    out = []
    for k, (nx, ny) in enumerate([(40, 20), (64, 33), (97, 41)]):
        cfg, g, inlet = pkg.configs.density_field(nx=nx, ny=ny, seed=10 + k, sigma=3.0)
        cfg.pressure_solver = T.CPS_PRESSURE_GMG
        cfg.precision = precision
        cfg.p_tol, cfg.p_rtol, cfg.max_pressure_iter, cfg.max_refine = 0.0, 1e-3, 6, 1
        cfg.max_momentum_iter = 8
        out.append((cfg, g, inlet))
    return out


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("streams", [1, 3])
def test_batch_replays_match_eager_fields(pkg, precision, streams):
    fl = fields(pkg, precision)
    b = pkg.Batch(fl)
    b.tune(streams=streams, reps=2)
    assert b.streams == streams
    for replay in range(3):
        b.step(1 + replay)
        b.synchronize()
        for k, (cfg, g, inlet) in enumerate(fl):
            st = b.stats(k)
            s = pkg.SIMPLE(cfg, g, inlet)
            s.step(st["iteration"])  # eager work between the replays
            for name in ("u", "v", "p"):
                np.testing.assert_array_equal(b.get(k, name), s.get(name), err_msg=f"{name} field {k} replay {replay}")
            assert b.pressure_drop(k) == s.full_static_drop()
            s.close()
    b.close()


def test_batch_autotune_and_mixed_sizes(pkg):
    fl = fields(pkg, 32)
    b = pkg.Batch(fl + fl[:1])
    ns, us = b.tune(streams=0, reps=2)
    assert ns in (1, 2, 4) and us > 0
    b.step(5)
    b.synchronize()
    for k in range(b.n):
        assert np.isfinite(b.stats(k)["error"])
    b.close()


def test_batch_rejects_uncapturable_caps(pkg):
    cfg, g, inlet = pkg.configs.density_field(nx=32, ny=16, seed=0)
    cfg.max_pressure_iter = 100
    with pytest.raises(pkg._capi.CpsError):
        pkg.Batch([(cfg, g, inlet)])


def test_batch_gradients_threaded_match_single_field(pkg):
    """Gradients of several fields computed concurrently (host threads,
    per-field streams) equal the single-field adjoint at the same state."""
    fl = fields(pkg, 64)
    b = pkg.Batch(fl)
    b.tune(streams=2, reps=2)
    b.step(300)
    b.synchronize()
    threaded = b.gradients_dp(rtol=1e-10, workers=3)
    for k, (cfg, g, inlet) in enumerate(fl):
        dp, grad, st = threaded[k]
        assert st["converged"] == 1
        s = pkg.SIMPLE(cfg, g, inlet)
        for name in ("u", "v", "p"):
            s.set(name, b.get(k, name))
        dp1, grad1 = s.gradient_dp(rtol=1e-10)
        assert dp == dp1
        np.testing.assert_array_equal(grad, grad1)
        s.close()
    b.close()


def test_batch_48_fields_at_real_sizes_replays_match_eager(pkg):
    """The design sweep at its own sizes: 48 fields drawn from 256 x 128 ..
    1024 x 512 (sigma U[4, 12], per-field seeds), FP32-inner / FP64-refined
    with the headline knobs, 16 streams.  After two rounds of graph replays,
    one field of each size equals the same field stepped eagerly on its own,
    bit for bit, and every field stays finite."""
    T = pkg._types
    fl = []
    for cfg, g, inlet in pkg.configs.batch_sweep(n_fields=48):
        cfg.pressure_solver, cfg.precision = T.CPS_PRESSURE_GMG, T.CPS_PREC_MIXED
        cfg.p_tol, cfg.p_rtol, cfg.max_pressure_iter, cfg.max_refine = 0.0, 1e-3, 6, 1
        cfg.max_momentum_iter = 8
        fl.append((cfg, g, inlet))
    pick = {}
    for k, (cfg, _, _) in enumerate(fl):
        pick.setdefault((cfg.nx, cfg.ny), k)
    assert len(pick) == 4
    print("case batch 48 fields, sizes", sorted(pick), flush=True)
    b = pkg.Batch(fl)
    b.tune(streams=16, reps=2)
    for rounds in (3, 7):
        b.step(rounds)
        b.synchronize()
        for (nx, ny), k in sorted(pick.items()):
            cfg, g, inlet = fl[k]
            st = b.stats(k)
            s = pkg.SIMPLE(cfg, g, inlet)
            s.step(st["iteration"])
            for name in ("u", "v", "p"):
                np.testing.assert_array_equal(b.get(k, name), s.get(name), err_msg=f"{name} field {k} {nx}x{ny}")
            assert b.pressure_drop(k) == s.full_static_drop()
            s.close()
    for k in range(b.n):
        assert np.isfinite(b.stats(k)["error"]), k
    b.close()
