#!/usr/bin/env python3
"""SIMPLE iteration throughput on one MI355X.

Workload: the 2-D density-field solve at 1024x512 (dx = dy = 2e-5 m,
projected random gamma with seed 0, Brinkman q = 0.1 / alpha_max = 1e8,
2.5-D sink with Ht = 1e-4 m, SOU, parabolic inlet 0.5 m/s).  A step is one
SIMPLE iteration (momentum assembly + Jacobi, pressure-correction solve,
corrections) with every field resident in HBM.

Two forms run the same algorithm and caps (GMG-PCG pressure, p_rtol 1e-3,
at most 6 PCG iterations, at most 8 Jacobi sweeps): FP32 inner iterations inside FP64
state / residuals / refinement ("value"), and plain FP64.  The pressure drop
of the two after warmup + steps iterations is compared.

The kernel-level roofline line times the fine-grid multigrid down kernel
(the previous PCG update fused with the pressure smoother, residual and
restriction) with HIP events carried by one launch in 8 on the solver
stream, over the timed region.  The CPU baseline is the oracle's FP32 restatement of the
reference CPU path (Jacobi-PCG pressure, 30 iterations, pTol 1e-6; Jacobi
momentum) on a bounded sample of the same workload.

More legs share the line: grads/s (d(dP)/d(gamma) of a 512x256 density
field at its SIMPLE fixed point by the discrete adjoint), a 48-field batch
of mixed sizes stepped by per-stream hipGraph launches, and the
512x256x12 conjugate-heat-transfer solve.

--copy-calibration runs only a device-to-device copy of 1 GiB (4 times),
the reference kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters
to bytes; --leg NAME runs one leg alone (for profiling it);
--profile-table OUT.md digests rocprofv3 CSV output (a --kernel-trace
directory and --pmc FETCH_SIZE / WRITE_SIZE directories of the same command
and of --copy-calibration) into a per-kernel table of durations, HBM bytes
and GB/s.

usage: python bench.py [--gpus 1] [--steps 1000] [--warmup 20]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SIMPLE iter/s @1024x512 FP32-refined vs FP64 at dP err<=1e-6; HBM %; grads/s"
HBM_PEAK_GBS = 8000.0


def perf_config(pkg, nx, ny, precision):
    cfg, gamma, inlet = pkg.configs.density_field(nx=nx, ny=ny, seed=0)
    cfg.pressure_solver = pkg._types.CPS_PRESSURE_GMG
    cfg.precision = precision
    cfg.p_tol = 0.0
    cfg.p_rtol = 1e-3
    cfg.max_pressure_iter = 6
    cfg.max_refine = 1
    cfg.max_momentum_iter = 8
    return cfg, gamma, inlet


def run_form(pkg, nx, ny, precision, steps, warmup, kernel=None):
    cfg, gamma, inlet = perf_config(pkg, nx, ny, precision)
    s = pkg.SIMPLE(cfg, gamma, inlet)
    if warmup:
        s.step(warmup)
    s.synchronize()
    if kernel:
        s.time_kernel(kernel)
    t0 = time.perf_counter()
    st = s.step(steps)
    s.synchronize()
    dt = time.perf_counter() - t0
    kt = s.kernel_timing() if kernel else None
    out = dict(it_s=steps / dt, seconds=dt, stats=st, dp=s.full_static_drop(), kernel=kt,
               fields={k: s.get(k) for k in ("u", "v", "p")})
    s.close()
    return out


def rel_l2(a, b):
    import numpy as np

    nb = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / (nb if nb > 0 else 1.0)


def time_to_dp_leg(pkg, nx=1024, ny=512, n_max=4000, tol=1e-6, ref_steps=3000):
    """Iterations and seconds until the pressure drop stays within tol of
    its converged value, for both forms with the headline knobs.  The
    converged value comes from the same field solved tightly in FP64 (GMG
    p_rtol 1e-12, momentum stop 1e-14, ref_steps iterations).  Each form runs
    n_max iterations with the outer loop's device log (one row per
    iteration); the answer is the first iteration after which every logged
    dP is within tol, and its time is that many timed iterations."""
    import numpy as np

    T = pkg._types
    cfg, gamma, inlet = pkg.configs.density_field(nx=nx, ny=ny, seed=0)
    cfg.pressure_solver, cfg.precision = T.CPS_PRESSURE_GMG, T.CPS_PREC_FP64
    cfg.p_tol, cfg.p_rtol, cfg.max_pressure_iter = 0.0, 1e-12, 60
    cfg.momentum_tol, cfg.max_momentum_iter = 1e-14, 60
    ref = pkg.SIMPLE(cfg, gamma, inlet)
    ref.step(ref_steps)
    dp_inf = ref.full_static_drop()
    ref_res = max(ref.state_residual())
    ref.close()
    out = {"dp_converged_pa": dp_inf, "reference": f"FP64 GMG p_rtol 1e-12, momentum tol 1e-14, {ref_steps} "
                                                   f"iterations, state residual {ref_res:.1e}", "tol": tol}
    for name, prec in (("mixed", 32), ("fp64", 64)):
        c, g, inl = perf_config(pkg, nx, ny, prec)
        c.max_iterations, c.epsilon, c.use_dp_convergence = n_max, 0.0, 0
        s = pkg.SIMPLE(c, g, inl)
        _, hist = s.runIterationsLogged(poll_every=64)
        err = np.abs(hist["dp_full_static"] - dp_inf) / abs(dp_inf)
        bad = np.nonzero(err > tol)[0]
        k = int(bad[-1]) + 1 if len(bad) else 0
        s.close()
        s = pkg.SIMPLE(c, g, inl)
        s.synchronize()
        t0 = time.perf_counter()
        if k > 0:
            s.step(k)
        s.synchronize()
        dt = time.perf_counter() - t0
        s.close()
        out[name] = {"iterations": k if k < n_max else None, "seconds": round(dt, 4) if k < n_max else None,
                     "dp_rel_err_final": float(err[-1])}
    return out


def oracle_parity(pkg, nx, ny, iters):
    """The plain FP64 GPU form with the reference's own algorithm and knobs
    (Jacobi-PCG 30 iterations, Jacobi momentum) against the FP64 oracle over
    the same iterations: relative L2 of u, v, p and the dP error."""
    import __graft_entry__ as entry

    orc = entry.load_oracle()
    cfg, gamma, inlet = pkg.configs.density_field(nx=nx, ny=ny, seed=0)
    o = orc.OracleSimple(pkg._types, cfg, gamma, inlet, prec=64)
    o.step(iters)
    s = pkg.SIMPLE(cfg, gamma, inlet)
    s.step(iters)
    out = {k: rel_l2(s.get(k), o.get(k)) for k in ("u", "v", "p")}
    out["dp"] = abs(s.full_static_drop() - o.dp_static()) / abs(o.dp_static())
    s.close()
    return {"iterations": iters, "rel_l2": {k: float(f"{v:.3g}") for k, v in out.items()}}


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "kernels_simple_1024x512.md")


def measured_traffic():
    """HBM bytes per launch of the fine-grid k_mg_down<float> measured by the
    two rocprofv3 --pmc runs recorded in the profile summary (FETCH_SIZE and
    WRITE_SIZE, scaled by a 1 GiB copy), or None."""
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def grads_leg(pkg, nx=512, ny=256, forward_steps=4000, reps=3, rtol=1e-8):
    """d(dP)/d(gamma) of a 512 x 256 density field: SIMPLE (FP64 GMG) driven
    to its fixed point, then `reps` full gradients (Jacobian build, adjoint
    FGMRES, gradient assembly) timed end to end."""
    T = pkg._types
    cfg, gamma, inlet = pkg.configs.density_field(nx=nx, ny=ny, seed=0)
    cfg.pressure_solver = T.CPS_PRESSURE_GMG
    cfg.p_tol, cfg.p_rtol, cfg.max_pressure_iter = 1e-13, 1e-8, 60
    cfg.momentum_tol, cfg.max_momentum_iter = 1e-13, 60
    s = pkg.SIMPLE(cfg, gamma, inlet)
    t0 = time.perf_counter()
    st = s.step(forward_steps)
    s.synchronize()
    t_fwd = time.perf_counter() - t0
    s.gradient_dp(rtol=rtol, max_iter=4000)  # warm-up (allocations)
    t0 = time.perf_counter()
    for _ in range(reps):
        dp, grad = s.gradient_dp(rtol=rtol, max_iter=4000)
    dt = (time.perf_counter() - t0) / reps
    out = dict(grads_per_s=1.0 / dt, seconds_per_grad=dt, adjoint=s.last_adjoint, dp=dp,
               forward_steps=forward_steps, forward_seconds=t_fwd, state_residual=list(s.state_residual()),
               resid_u=st["resid_u"])
    s.close()
    return out


def batch_leg(pkg, steps=500, n_fields=48, plan=None, precision=None, start_at=0):
    """48 density fields, sizes drawn from {256x128, 512x256, 768x384,
    1024x512}, sigma ~ U[4, 12], per-field seeds, the headline knobs
    (FP32-inner GMG-PCG, p_rtol 1e-3, <= 6 PCG iterations, <= 8 Jacobi
    sweeps): plan (class plan, one launch per kernel covering a size class,
    or per-field chains on several streams) and graph unroll picked by the
    autotuner, `steps` iterations timed.  precision: CPS_PREC_MIXED
    (default) or CPS_PREC_FP64 (the same algorithm in plain FP64).
    start_at: with a given plan, the fields are first stepped to at least
    this iteration (a step costs less as the fields converge, so two legs
    compare only from the same start; the tuner's own timing leaves the
    fields past their 128th iteration)."""
    T = pkg._types
    precision = T.CPS_PREC_MIXED if precision is None else precision
    fl = []
    for cfg, gamma, inlet in pkg.configs.batch_sweep(n_fields=n_fields):
        cfg.pressure_solver, cfg.precision = T.CPS_PRESSURE_GMG, precision
        cfg.p_tol, cfg.p_rtol, cfg.max_pressure_iter, cfg.max_refine = 0.0, 1e-3, 6, 1
        cfg.max_momentum_iter = 8
        fl.append((cfg, gamma, inlet))
    b = pkg.Batch(fl)
    if plan:  # "class:UNROLL[:CHAINS]" or "per_field:STREAMS" instead of the tuner's pick
        kind, *n = plan.split(":")
        if kind == "class":
            b.plan(kind, streams=int(n[1]) if len(n) > 1 else 1, unroll=int(n[0]))
        else:
            b.plan(kind, streams=int(n[0]))
        u = int(n[0]) if kind == "class" else 1
        b.step(u * max(1, -(-start_at // u)))  # warm-up
    else:
        b.tune(streams=0, reps=3)
    ns = b.streams
    plan = b.plan_info()
    b.synchronize()
    start = b.stats(0)["iteration"]
    t0 = time.perf_counter()
    b.step(steps)
    b.synchronize()
    dt = time.perf_counter() - t0
    cells = sum(c.nx * c.ny for c, _, _ in fl)
    finite = all(math.isfinite(b.stats(k)["error"]) for k in range(b.n))
    b.close()
    return dict(fields=n_fields, steps=steps, seconds=dt, streams=ns, plan=plan["plan"], plan_param=plan["param"],
                unroll=plan["unroll"], start_iteration=start,
                size_classes=plan["classes"], batch_it_s=steps / dt,
                field_it_s=n_fields * steps / dt, cell_updates_per_s=cells * steps / dt, total_cells=cells,
                all_finite=finite)


def grads_batch_leg(pkg, n_fields=32, nx=512, ny=256, forward_steps=1500, rtol=1e-8, workers=8):
    """32 seeded 512 x 256 density fields brought near their fixed points as
    one batch (one size class: each kernel launch covers the 32 fields,
    graphs of 10 iterations), then all 32 d(dP)/d(gamma) computed at
    once: one host thread per field in flight, each adjoint on its field's
    stream.  Timed: the 32 gradients."""
    T = pkg._types
    fl = []
    for k in range(n_fields):
        cfg, gamma, inlet = pkg.configs.density_field(nx=nx, ny=ny, seed=k)
        cfg.pressure_solver, cfg.precision = T.CPS_PRESSURE_GMG, T.CPS_PREC_FP64
        cfg.p_tol, cfg.p_rtol, cfg.max_pressure_iter, cfg.max_refine = 0.0, 1e-6, 30, 1
        cfg.momentum_tol, cfg.max_momentum_iter = 1e-13, 30
        fl.append((cfg, gamma, inlet))
    b = pkg.Batch(fl)
    b.plan("class", unroll=10)
    t0 = time.perf_counter()
    b.step(forward_steps)
    b.synchronize()
    t_fwd = time.perf_counter() - t0
    b.gradient_dp(0, rtol=rtol)  # warm-up (one field's adjoint allocations)
    t0 = time.perf_counter()
    res = b.gradients_dp(rtol=rtol, workers=workers)
    dt = time.perf_counter() - t0
    its = [r[2]["iterations"] for r in res]
    ok = all(r[2]["converged"] for r in res)
    its_max = max(its)
    b.close()
    return dict(fields=n_fields, workers=workers, seconds=dt, grads_per_s=n_fields / dt, forward_steps=forward_steps,
                forward_seconds=t_fwd, fgmres_iterations_mean=sum(its) / len(its), fgmres_iterations_max=its_max,
                all_converged=ok)


def thermal_cpu_baseline(pkg, st, gamma, u, v, maxit=4000):
    """The oracle's CPU restatement of the thermal solve on the same system:
    assembly (thermal_solver.cpp:437-530) plus Jacobi-preconditioned
    BiCGStab (OpenMP) to rtol 1e-6, at most maxit iterations.  AMGCL, the
    reference's solver, is not available here; this is the stated stand-in."""
    import __graft_entry__ as entry

    orc = entry.load_oracle()
    t0 = time.perf_counter()
    o = orc.OracleThermal(pkg._types, st, gamma, u, v)
    t_asm = time.perf_counter() - t0
    o.solve(rtol=1e-6, maxit=maxit)
    return {"seconds": t_asm + o.last_seconds, "assembly_seconds": t_asm, "iterations": abs(o.last_iters),
            "converged": o.last_iters >= 0, "rel_residual": o.last_relres,
            "cores": int(os.environ.get("OMP_NUM_THREADS", "0")) or None,
            "kind": "port", "solver": "Jacobi-BiCGStab (oracle), rtol 1e-6, <= %d iterations" % maxit}


def thermal_leg(pkg, reps=3, flow_steps=3000, cpu=True):
    """512 x 256 x 12 conjugate-heat-transfer solve (4 Cu base, 4 channel,
    4 Cu lid layers) to the reference's rtol 1e-6, FP64 and FP32-refined.
    The channel velocity is the cell-centred field of a 512 x 256 density-
    field SIMPLE solve (seed 0, FP64 GMG, flow_steps iterations) on the same
    gamma; seeded power map (20 W/cm^2 + 8 hotspots).  Each form also times
    the gradient of the layer-0 p-norm (p = 8): adjoint solve + sensitivity."""
    T = pkg._types
    st = pkg.configs.thermal_stack()
    cfg, gamma, inlet = pkg.configs.density_field(nx=st["nx"], ny=st["ny"], seed=0)
    cfg.pressure_solver = T.CPS_PRESSURE_GMG
    cfg.p_tol, cfg.p_rtol, cfg.max_pressure_iter = 1e-13, 1e-8, 60
    cfg.momentum_tol, cfg.max_momentum_iter = 1e-13, 60
    flow = pkg.SIMPLE(cfg, gamma, inlet)
    fst = flow.step(flow_steps)
    u, v = pkg.configs.cell_centre_velocity(flow.get("u"), flow.get("v"))
    flow.close()
    out = {"flow_steps": flow_steps, "flow_resid_u": fst["resid_u"]}
    for name, prec in (("fp64", T.CPS_PREC_FP64), ("mixed", T.CPS_PREC_MIXED)):
        s = pkg.ThermalSolver3D(st, gamma, u, v, precision=prec)
        s.solve()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = s.solve()
        dt = (time.perf_counter() - t0) / reps
        out[name] = dict(seconds=dt, iterations=r["iterations"], refinements=r["refinements"],
                         rel_residual=r["rel_residual"])
        # dJ/dgamma of the layer-0 p-norm (p = 8): adjoint solve + sensitivity
        s.gradient("pnorm", 8.0)
        t0 = time.perf_counter()
        for _ in range(reps):
            s.gradient("pnorm", 8.0)
        out[name]["grad_seconds"] = (time.perf_counter() - t0) / reps
        out[name]["grad_iterations"] = s.last_adjoint["iterations"]
        if name == "fp64":
            out["tb_max"] = s.metrics(8.0)["tb_max"]
        s.close()
    if cpu:
        out["cpu_baseline"] = thermal_cpu_baseline(pkg, st, gamma, u, v)
    return out


def cpu_baseline(pkg, nx, ny, seconds):
    """Oracle FP32 restatement of the reference CPU path, timed for about
    `seconds` of host work (at least 3 SIMPLE iterations)."""
    import __graft_entry__ as entry

    orc = entry.load_oracle()
    cfg, gamma, inlet = pkg.configs.density_field(nx=nx, ny=ny, seed=0)  # reference knobs: PCG, 30 its
    o = orc.OracleSimple(pkg._types, cfg, gamma, inlet, prec=32)
    o.step(1)  # first step builds maps / touches memory
    n, spent = 0, 0.0
    while spent < seconds or n < 3:
        o.step(1)
        spent += o.last_seconds
        n += 1
    return dict(value=n / spent, unit="SIMPLE it/s", cores=int(os.environ.get("OMP_NUM_THREADS", "0")) or None,
                kind="port",
                sample=f"{n} SIMPLE iterations of the {nx}x{ny} density field after 1 warm-up, {spent:.1f} s: "
                       f"the oracle's FP32 restatement of the reference CPU path at the reference's inner caps "
                       f"(Jacobi-PCG 30 iterations at pTol 1e-6; Jacobi momentum to tol 1e-6, at most 400 sweeps). "
                       f"A rate per iteration: the GPU line's iterations use <= 6 multigrid-PCG iterations and "
                       f"<= 8 Jacobi sweeps (time_to_dp compares time to a converged dP)")


def _short_kernel(name):
    import re

    m = re.search(r"(k_[A-Za-z0-9_]+(<[^>]*>)?)", name)
    return m.group(1) if m else name.split("(")[0][:48]


def _csv_rows(d, suffix):
    import csv
    import glob

    for f in sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True)):
        with open(f) as fh:
            yield from csv.DictReader(fh)


def split_early(values, ratio=2.0):
    """Dispatches enqueued after a device stop flag was set (a converged
    PCG or Jacobi solve) return at once; their durations and counters form a
    cluster well below those of the dispatches that did the work.  Two-means
    clustering of log(value): returns (values of the upper cluster, number
    left out) when the cluster means differ by more than `ratio` and each
    cluster holds at least a tenth of the values, else (all values, 0)."""
    v = sorted(values)
    if len(v) < 2:
        return v, 0
    lg = [math.log(max(x, 1e-12)) for x in v]
    lo, hi = lg[0], lg[-1]
    cut = None
    for _ in range(50):
        mid = 0.5 * (lo + hi)
        c = next((i for i, x in enumerate(lg) if x > mid), len(lg))
        if c == 0 or c == len(lg):
            return v, 0
        nlo, nhi = sum(lg[:c]) / c, sum(lg[c:]) / (len(lg) - c)
        if cut == c:
            break
        cut, lo, hi = c, nlo, nhi
    low, high = v[:cut], v[cut:]
    if (sum(high) / len(high)) <= ratio * (sum(low) / len(low)) or min(len(low), len(high)) < 0.1 * len(v):
        return v, 0
    return high, len(low)


def fused_fine_down_values(rows, kernel="k_mg_down<float, 2>", markers=("k_pcg_start<float", "k_prhs<false, float>")):
    """rows: (order key, short kernel name, value) of every dispatch.  The
    values of the roofline kernel's fused launches: dispatches of `kernel`
    that do not directly follow a dispatch of one of `markers` (the kernel
    that starts a pressure solve: its own start kernel, or the right-side
    kernel that carries the start; the first V-cycle of each solve runs the
    unfused form), early exits left out (split_early).  Returns (values,
    number of early exits, number of unfused first launches)."""
    fused, first, after = [], 0, False
    for _, k, v in sorted(rows):
        if k.startswith(markers):
            after = True
        elif k == kernel:
            if after:
                first += 1
            else:
                fused.append(v)
            after = False
    work, nearly = split_early(fused)
    return work, nearly, first


def fused_fine_down(trace):
    """Durations [us] of the roofline kernel's fused launches in a kernel trace."""
    return fused_fine_down_values([(int(r["Start_Timestamp"]), _short_kernel(r["Kernel_Name"]),
                                    (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
                                   for r in _csv_rows(trace, "kernel_trace.csv")])


def fused_fine_down_counter(d, name):
    """Mean counter value [KiB] over the roofline kernel's fused launches of a --pmc run."""
    work, _, _ = fused_fine_down_values([(int(r["Dispatch_Id"]), _short_kernel(r["Kernel_Name"]),
                                          float(r["Counter_Value"]))
                                         for r in _csv_rows(d, "counter_collection.csv") if r["Counter_Name"] == name])
    return sum(work) / len(work) if work else None


def profile_table(out, trace, fetch, write, copy_fetch, copy_write, title, skip=0.0, roofline_bytes=0.0,
                  from_kernel=None):
    """Per kernel and grid: dispatches, mean duration (kernel trace, end -
    start; dispatches after the first `skip` fraction of the run), HBM
    bytes read / written per dispatch (mean FETCH_SIZE / WRITE_SIZE in
    KiB, scaled by the factor that maps the same counter of a 1 GiB device
    copy to 1 GiB) and the HBM GB/s they give.  Dispatches that exited early
    on a device stop flag (split_early) are counted apart and left out of
    the means."""
    import statistics as st

    def grid_of(r):
        return int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])

    rows = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"]), _short_kernel(r["Kernel_Name"]), grid_of(r),
             f'{r["Grid_Size_X"]}x{r["Grid_Size_Y"]}x{r["Grid_Size_Z"]}') for r in _csv_rows(trace, "kernel_trace.csv")]
    rows.sort()
    rows = rows[int(len(rows) * skip):]
    if from_kernel:  # the window from the first dispatch of this kernel on
        first = next((q for q, r in enumerate(rows) if r[2].startswith(from_kernel)), 0)
        rows = rows[first:]
    dur = {}
    for t0, t1, k, g, gs in rows:
        dur.setdefault((k, g, gs), []).append((t1 - t0) / 1e3)

    def counters(d, name):
        acc = {}
        for r in _csv_rows(d, "counter_collection.csv"):
            if r["Counter_Name"] != name:
                continue
            g = int(r["Grid_Size"]) if r.get("Grid_Size") else grid_of(r)
            acc.setdefault((_short_kernel(r["Kernel_Name"]), g), []).append(float(r["Counter_Value"]))
        out = {}
        for k, v in acc.items():
            hi, _ = split_early(v)
            out[k] = sum(hi) / len(hi)
        return out

    def scale(d, name):
        c = counters(d, name)
        big = max(c.values())  # the 1 GiB copy is the largest dispatch of the calibration run
        return (1 << 30) / (big * 1024.0)

    F, W = counters(fetch, "FETCH_SIZE"), counters(write, "WRITE_SIZE")
    fs, ws = scale(copy_fetch, "FETCH_SIZE"), scale(copy_write, "WRITE_SIZE")
    tot = sum(sum(v) for v in dur.values())
    lines = [f"## {title}", "",
             f"HBM bytes: FETCH_SIZE x {fs:.3f}, WRITE_SIZE x {ws:.3f} (KiB -> bytes of a 1 GiB copy); "
             f"GB/s = (read + written) / mean duration; fraction of the 8 TB/s peak.", "",
             "Early exits: dispatches that returned at once on a device stop flag (not in the means).", "",
             "| kernel | grid (threads) | dispatches | early exits | mean us | total ms | % | read MB | written MB | HBM GB/s | of 8 TB/s |",
             "|---|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for (k, g, gs), v in sorted(dur.items(), key=lambda kv: -sum(kv[1])):
        if not k.startswith("k_"):
            continue
        work, nearly = split_early(v)
        mean = sum(work) / len(work)
        rd = F.get((k, g))
        wr = W.get((k, g))
        if rd is None or wr is None:
            cells = "| - | - | - | - |"
        else:
            rb, wb = rd * 1024 * fs, wr * 1024 * ws
            gbs = (rb + wb) / (mean * 1e-6) / 1e9
            cells = f"| {rb / 1e6:.2f} | {wb / 1e6:.2f} | {gbs:.0f} | {gbs / 8000:.2f} |"
        lines.append(f"| {k} | {gs} | {len(v)} | {nearly} | {mean:.2f} | {sum(v) / 1e3:.3f} | {100 * sum(v) / tot:.1f} "
                     + cells)
    if roofline_bytes > 0:
        work, nearly, first = fused_fine_down(trace)
        if work:
            mean = sum(work) / len(work)
            gbs = roofline_bytes / (mean * 1e-6) / 1e9
            lines += ["", f"Roofline kernel (fused fine-level k_mg_down<float, 2>, FP32-inner form): {len(work)} dispatches "
                      f"that did the work ({nearly} early exits and {first} unfused first-V-cycle launches left out), "
                      f"mean {mean:.2f} us; {roofline_bytes / 1e6:.2f} MB algorithmic per launch -> {gbs:.0f} GB/s, "
                      f"{gbs / 8000:.3f} of 8 TB/s."]
        fr, wr = fused_fine_down_counter(fetch, "FETCH_SIZE"), fused_fine_down_counter(write, "WRITE_SIZE")
        if fr is not None and wr is not None:
            rb, wb = fr * 1024 * fs, wr * 1024 * ws
            lines += ["", f"Its measured HBM traffic (the same fused launches of the counter runs): {rb / 1e6:.2f} MB read + "
                      f"{wb / 1e6:.2f} MB written = {(rb + wb) / 1e6:.2f} MB per launch, against "
                      f"{roofline_bytes / 1e6:.2f} MB algorithmic.", "",
                      f"traffic_bytes_per_launch: {rb + wb:.0f}"]
    text = "\n".join(lines) + "\n"
    with open(out, "w") as fh:
        fh.write(text)
    print(text, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--parity-iters", type=int, default=30,
                    help="SIMPLE iterations of the FP64 GPU form against the FP64 oracle (reference knobs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison run")
    ap.add_argument("--no-grads", action="store_true", help="skip the d(dP)/d(gamma) leg")
    ap.add_argument("--no-thermal", action="store_true", help="skip the 3-D thermal leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 48-field batch leg")
    ap.add_argument("--no-time-to-dp", action="store_true", help="skip the time-to-converged-dP leg")
    ap.add_argument("--batch-plan", default=None,
                    help="batch leg plan instead of the autotuner's: class:UNROLL[:CHAINS] or per_field:STREAMS")
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--copy-calibration", action="store_true", help="only copy 1 GiB on the device (4x) and exit")
    ap.add_argument("--profile-table", metavar="OUT_MD", default=None,
                    help="digest rocprofv3 CSVs (--trace, --fetch, --write, --copy-fetch, --copy-write dirs) and exit")
    for name in ("trace", "fetch", "write", "copy-fetch", "copy-write"):
        ap.add_argument("--" + name, default=None)
    ap.add_argument("--title", default="per-kernel HBM traffic")
    ap.add_argument("--from-kernel", default=None,
                    help="profile table: durations from the first dispatch of this kernel on (counters: all)")
    ap.add_argument("--roofline-bytes", type=float, default=0.0,
                    help="profile table: algorithmic bytes per fused fine down launch (adds the roofline line)")
    ap.add_argument("--skip", type=float, default=0.0, help="profile table: fraction of early dispatches left out")
    ap.add_argument("--leg", choices=["grads", "grads_batch", "thermal", "batch", "time_to_dp", "parity"],
                    help="run only this leg and print its results (for profiling)")
    args = ap.parse_args()
    if args.profile_table:
        profile_table(args.profile_table, args.trace, args.fetch, args.write, args.copy_fetch, args.copy_write,
                      args.title, args.skip, args.roofline_bytes, args.from_kernel)
        return
    if args.copy_calibration:
        import torch

        n = (1 << 30) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 4 * n, "copy_calibration_bytes_written": 4 * n}), flush=True)
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the SIMPLE path runs one density field per device and "
              "this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    # hardware queues per process for the concurrent streams of the batch
    # leg (HIP reads it at initialisation; the caller's setting wins)
    os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")
    if "OMP_NUM_THREADS" not in os.environ:
        os.environ["OMP_NUM_THREADS"] = str(min(16, os.cpu_count() or 1))

    import __graft_entry__ as entry

    pkg = entry.load_package()
    if args.leg:
        legs = {"grads": grads_leg, "grads_batch": grads_batch_leg, "thermal": thermal_leg,
                "batch": lambda p: batch_leg(p, steps=args.batch_steps, plan=args.batch_plan),
                "time_to_dp": time_to_dp_leg,
                "parity": lambda p: oracle_parity(p, args.nx, args.ny, args.parity_iters)}
        res = legs[args.leg](pkg)
        print(json.dumps({"leg": args.leg, **res}, default=str), flush=True)
        return
    nx, ny = args.nx, args.ny
    mixed = run_form(pkg, nx, ny, 32, args.steps, args.warmup, kernel="mg_smooth0")
    fp64 = None if args.no_fp64 else run_form(pkg, nx, ny, 64, args.steps, args.warmup)

    kt = mixed["kernel"]
    roof = None
    if kt and kt["launches"] > 0 and kt["avg_us"] > 0:
        gbs = kt["bytes_per_launch"] / (kt["avg_us"] * 1e-6) / 1e9
        roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": measured_traffic(),
                "kernel": "k_mg_down<float> on the fine grid with the single-reduction PCG update fused "
                          "(p = z + beta p, s = w + beta s, x += alpha p, r -= alpha s, 2 pressure smoother "
                          "sweeps, residual, restriction)",
                "avg_us": round(kt["avg_us"], 3), "launches": kt["launches"],
                "bytes_per_launch": kt["bytes_per_launch"]}

    grads = None if args.no_grads else grads_leg(pkg)
    thermal = None if args.no_thermal else thermal_leg(pkg, cpu=not args.no_cpu)
    batch = None if args.no_batch else batch_leg(pkg, steps=args.batch_steps, plan=args.batch_plan)
    if batch is not None:
        # the same fields in plain FP64 under the same plan: the precision
        # ratio where the kernels carry whole size classes
        same = (f"class:{batch['unroll']}:{batch['plan_param']}" if batch["plan"] == "class"
                else f"per_field:{batch['streams']}")
        b64 = batch_leg(pkg, steps=args.batch_steps, plan=same, precision=pkg._types.CPS_PREC_FP64,
                        start_at=batch["start_iteration"])
        batch["fp64_start_iteration"] = b64["start_iteration"]
        batch["fp64_cell_updates_per_s"] = b64["cell_updates_per_s"]
        batch["speedup_vs_fp64"] = batch["cell_updates_per_s"] / b64["cell_updates_per_s"]
    grads_batch = None if args.no_grads else grads_batch_leg(pkg)
    tdp = None if args.no_fp64 or args.no_time_to_dp else time_to_dp_leg(pkg, nx, ny)
    cpu = None if args.no_cpu else cpu_baseline(pkg, nx, ny, args.cpu_seconds)
    parity = None if args.no_cpu else oracle_parity(pkg, nx, ny, args.parity_iters)

    dp_err = None
    if fp64 is not None and fp64["dp"] != 0:
        dp_err = abs(mixed["dp"] - fp64["dp"]) / abs(fp64["dp"])
    line = {
        "metric": METRIC,
        "value": round(mixed["it_s"], 2),
        "unit": "SIMPLE it/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 / mixed["it_s"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-inner/f64-refined",
        "data": "synthetic",
        "config": {
            "workload": f"2-D density field {nx}x{ny}: seeded projected random gamma, Brinkman q=0.1 "
                        f"alpha_max=1e8, 2.5-D Ht=1e-4 m, SOU, parabolic inlet 0.5 m/s; GMG-PCG pressure "
                        f"(p_rtol 1e-3, <=6 its), <=8 Jacobi sweeps",
            "nx": nx, "ny": ny, "cells": nx * ny,
        },
        "fp64_it_s": None if fp64 is None else round(fp64["it_s"], 2),
        "speedup_vs_fp64": None if fp64 is None else round(mixed["it_s"] / fp64["it_s"], 3),
        "dp_pa": mixed["dp"],
        "dp_fp64_pa": None if fp64 is None else fp64["dp"],
        "dp_rel_err": dp_err,
        "fields_rel_l2_vs_fp64": None if fp64 is None else {
            k: float(f"{rel_l2(mixed['fields'][k], fp64['fields'][k]):.3g}") for k in ("u", "v", "p")},
        "fp64_vs_oracle": parity,
        "time_to_dp": tdp,
        "cell_updates_per_s": round(mixed["it_s"] * nx * ny, 1),
        "grads_per_s": None if grads is None else round(grads["grads_per_s"], 3),
        "grads": None if grads is None else {
            "workload": "d(dP)/d(gamma), 512x256 density field (seed 0, SOU) at its SIMPLE fixed point; "
                        "Jacobian build + adjoint FGMRES (rtol 1e-8; LSC block preconditioner, x/y line "
                        "momentum solves) + gradient",
            "seconds_per_grad": round(grads["seconds_per_grad"], 4),
            "fgmres_iterations": grads["adjoint"]["iterations"],
            "adjoint_rel_residual": grads["adjoint"]["rel_residual"],
            "forward_steps": grads["forward_steps"], "forward_seconds": round(grads["forward_seconds"], 3),
            "state_residual_max": max(grads["state_residual"]),
        },
        "grads_batch": None if grads_batch is None else {
            "workload": "32 seeded 512x256 density fields (SOU) after 1500 batched SIMPLE iterations; all 32 "
                        "d(dP)/d(gamma) at once (host threads, per-field streams), rtol 1e-8, LSC x/y line "
                        "preconditioner",
            **{k: (float(f"{v:.5g}") if isinstance(v, float) else v) for k, v in grads_batch.items()},
        },
        "batch": None if batch is None else {
            "workload": "48 density fields, sizes from {256x128, 512x256, 768x384, 1024x512}, sigma U[4,12], "
                        "headline knobs; autotuned plan (class plan: one launch per kernel over all fields "
                        "of a grid size, one hipGraph per size class); fp64_*: the same fields and plan in plain FP64, "
                        "timed from the same field iteration (start_iteration)",
            **{k: (float(f"{v:.5g}") if isinstance(v, float) else v) for k, v in batch.items()},
        },
        "thermal": None if thermal is None else {
            "workload": "512x256x12 CHT stack (4 Cu base, 4 channel, 4 Cu lid), rtol 1e-6; channel velocity from "
                        "the 512x256 density-field SIMPLE solve on the same gamma; grad = d(layer-0 8-norm)/d(gamma)",
            **{k: ({kk: (float(f"{vv:.4g}") if isinstance(vv, float) else vv) for kk, vv in d.items()}
                   if isinstance(d, dict) else (float(f"{d:.6g}") if isinstance(d, float) else d))
               for k, d in thermal.items()},
        },
        "p_iters_last": mixed["stats"]["p_iters"],
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
