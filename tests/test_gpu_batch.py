"""Batched fields through captured hipGraphs.

Every field of a batch must equal the same field stepped eagerly on its own
(identical kernels doing identical arithmetic per field, so bit for bit),
after each of several graph replays separated by eager work, under every
plan the autotuner may pick: the class plan (one launch per kernel covering
all fields of a size class, gridDim.z = fields, buffers at a fixed stride),
with and without unrolled graphs, and the per-field plan on 1 or 3 streams.
The batch mixes sizes and holds several fields of the same size, so the
class plan launches over 1, 2 and 3 fields at once.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


SIZES = [(40, 20), (64, 33), (97, 41), (64, 33), (64, 33), (40, 20)]


def fields(pkg, precision, sizes=SIZES):
    T = pkg._types
    out = []
    for k, (nx, ny) in enumerate(sizes):
        cfg, g, inlet = pkg.configs.density_field(nx=nx, ny=ny, seed=10 + k, sigma=3.0)
        cfg.pressure_solver = T.CPS_PRESSURE_GMG
        cfg.precision = precision
        cfg.p_tol, cfg.p_rtol, cfg.max_pressure_iter, cfg.max_refine = 0.0, 1e-3, 6, 1
        cfg.max_momentum_iter = 8
        out.append((cfg, g, inlet))
    return out


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("plan", [("class", 0, 1), ("class", 0, 3), ("class", 2, 4), ("per_field", 1, 1),
                                  ("per_field", 3, 1)])
def test_batch_replays_match_eager_fields(pkg, precision, plan):
    fl = fields(pkg, precision)
    print("case batch replays", precision, plan, flush=True)
    b = pkg.Batch(fl)
    b.plan(*plan)
    info = b.plan_info()
    assert info["plan"] == plan[0] and info["unroll"] == plan[2] and info["classes"] == 3
    # class plan: classes of 2, 3 and 1 fields, each in min(chains, fields) chains
    chains = sum(min(max(plan[1], 1), m) for m in (2, 3, 1))
    assert b.streams == (chains if plan[0] == "class" else plan[1])
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


def test_batch_autotune_and_json_cache(pkg, tmp_path):
    """The tuner times every plan and keeps one; its pick is stored under the
    batch's signature and a second batch of the same shape captures it from
    the cache without timing, and steps to the same fields."""
    import json

    fl = fields(pkg, 32)
    cache = str(tmp_path / "batch_tune.json")
    b = pkg.Batch(fl)
    ns, us = b.tune(streams=0, reps=2, cache=cache)
    info = b.plan_info()
    print("tuned", info, flush=True)
    assert ns >= 1 and us > 0 and info["classes"] == 3
    entry = json.load(open(cache))["entries"][b.signature()]
    assert entry["plan"] == info["plan"] and entry["unroll"] == info["unroll"] and entry["param"] == info["param"]
    b2 = pkg.Batch(fl)
    ns2, us2 = b2.tune(streams=0, cache=cache)
    info2 = b2.plan_info()
    assert (info2["plan"], info2["param"], info2["unroll"], info2["streams"]) == \
        (info["plan"], info["param"], info["unroll"], info["streams"])
    assert us2 == pytest.approx(entry["us_per_step"])
    it = b.stats(0)["iteration"]
    b2.step(it - b2.stats(0)["iteration"])
    b2.synchronize()
    for k in range(b.n):
        assert np.isfinite(b.stats(k)["error"])
        np.testing.assert_array_equal(b.get(k, "p"), b2.get(k, "p"), err_msg=f"field {k}")
    b.close()
    b2.close()


def test_batch_rejects_uncapturable_caps(pkg):
    cfg, g, inlet = pkg.configs.density_field(nx=32, ny=16, seed=0)
    cfg.max_pressure_iter = 100
    with pytest.raises(pkg._capi.CpsError):
        pkg.Batch([(cfg, g, inlet)])


def test_batch_gradients_threaded_match_single_field(pkg):
    """Gradients of several fields computed concurrently (host threads,
    per-field streams) after class-plan stepping equal the single-field
    adjoint at the same state."""
    fl = fields(pkg, 64)
    b = pkg.Batch(fl)
    b.plan("class", unroll=4)
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
    with the headline knobs, class plan (four classes of about 12 fields,
    graphs of 4 iterations).  After two rounds of graph replays,
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
    b.plan("class", unroll=4)
    assert b.plan_info()["classes"] == 4
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
