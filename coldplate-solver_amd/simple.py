"""Host-side mirror of the reference SIMPLE solver interface on the HIP path.

`SIMPLE` follows the reference class (SIMPLE.h:55-614): construction loads
the grid, density and Brinkman field (SIMPLE::SIMPLE, SIMPLE.cpp:18-29),
`calculateStep` performs one SIMPLE iteration (iterations.cpp:82-659) and
returns the max RMS residual, `runIterations` iterates to `epsilon` or
`maxIterations` (SIMPLE.cpp:378), and `sampleBoundaryPatch` /
`full_static_drop` give the full-system pressure drop
(postprocessing.cpp:245-327, SIMPLE.cpp:438).  Every call goes through the C
ABI of libcoldplate_hip.so; the fields stay on the device.

`pressure_solve_csr` / `momentum_jacobi_csr` are the drop-ins for the
reference GPU backend calls amgcl_cuda::solve and
amgcl_cuda::solveMomentumU/V (cuda_solver.hpp:24-57).
"""
import ctypes as C

import numpy as np

from . import _capi
from . import _types as T


class SIMPLE:
    def __init__(self, cfg, gamma, inlet=None, device=0):
        self.ctx = _capi.context(device)
        self.L = self.ctx.L
        self.cfg = cfg.copy()
        g = np.ascontiguousarray(gamma, dtype=np.float64)
        if g.shape != (cfg.ny, cfg.nx):
            raise ValueError(f"gamma must be ny x nx = {(cfg.ny, cfg.nx)}, got {g.shape}")
        inl = None
        if inlet is not None:
            inl = np.ascontiguousarray(inlet, dtype=np.float64)
            if inl.shape != (cfg.ny + 2,):
                raise ValueError("inlet profile must have ny + 2 (= M + 1) entries")
        h = C.c_void_p()
        _capi.check(self.L.cps_simple_create(self.ctx.h, C.byref(self.cfg), _capi.dptr(g),
                                             None if inl is None else _capi.dptr(inl), C.byref(h)))
        self.h = h
        self.M, self.N = cfg.ny + 1, cfg.nx + 1
        self.last = None

    def close(self):
        if getattr(self, "h", None):
            self.L.cps_simple_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- iteration -------------------------------------------------------
    def step(self, n=1):
        st = T.StepStats()
        _capi.check(self.L.cps_simple_step(self.h, n, C.byref(st)))
        self.last = st.as_dict()
        return self.last

    def calculateStep(self):
        return self.step(1)["error"]

    def runIterations(self):
        """SIMPLE::runIterations (SIMPLE.cpp:310-491): stop on error <=
        epsilon, max_iterations or the settled core-dP window; the returned
        stats carry the four pressure drops and `dp_converged`."""
        st = T.StepStats()
        _capi.check(self.L.cps_simple_run(self.h, C.byref(st)))
        self.last = st.as_dict()
        return self.last

    def runIterationsLogged(self, poll_every=None):
        """runIterations plus the per-iteration log rows (residuals and the
        four pressure drops, writeIterationLogs output.cpp:330-360) as a
        dict of numpy arrays.  poll_every: iterations the host enqueues
        between reads of the device's stop flag (default 32); the host
        synchronisations of the run are in self.last_host_syncs."""
        st = T.StepStats()
        cap = max(int(self.cfg.max_iterations), 0)
        log = (T.IterLog * max(cap, 1))()
        n, syncs = C.c_int(), C.c_int()
        if poll_every is None:
            _capi.check(self.L.cps_simple_run_logged(self.h, C.byref(st), log, cap, C.byref(n)))
            syncs.value = -1
        else:
            _capi.check(self.L.cps_simple_run_polled(self.h, int(poll_every), C.byref(st), log, cap, C.byref(n),
                                                     C.byref(syncs)))
        self.last = st.as_dict()
        self.last_host_syncs = syncs.value
        rows = log[: n.value]
        hist = {k: np.array([getattr(r, k) for r in rows]) for k, _ in T.IterLog._fields_ if k != "pad_"}
        return self.last, hist

    # ---- fields ----------------------------------------------------------
    def get(self, name):
        out = np.empty(T.field_shape(name, self.cfg.nx, self.cfg.ny))
        _capi.check(self.L.cps_simple_get_field(self.h, T.FIELDS[name], _capi.dptr(out)))
        return out

    def set(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != T.field_shape(name, self.cfg.nx, self.cfg.ny):
            raise ValueError(f"{name} must have shape {T.field_shape(name, self.cfg.nx, self.cfg.ny)}")
        _capi.check(self.L.cps_simple_set_field(self.h, T.FIELDS[name], _capi.dptr(a)))

    u = property(lambda self: self.get("u"))
    v = property(lambda self: self.get("v"))
    p = property(lambda self: self.get("p"))

    # ---- post-processing -------------------------------------------------
    def sampleBoundaryPatch(self, inlet):
        m = T.Patch()
        _capi.check(self.L.cps_simple_boundary_patch(self.h, 1 if inlet else 0, C.byref(m)))
        return m.as_dict()

    def samplePlaneAtX(self, x):
        """Plane metrics at physical x [m] (postprocessing.cpp:102-243)."""
        m = T.Patch()
        _capi.check(self.L.cps_simple_plane(self.h, float(x), C.byref(m)))
        return m.as_dict()

    # ---- adjoint ---------------------------------------------------------
    def gradient_dp(self, rtol=1e-10, max_iter=400):
        """(dP, d(dP)/dgamma (ny, nx)) at the current state by the discrete
        adjoint of the converged SIMPLE equations."""
        grad = np.empty((self.cfg.ny, self.cfg.nx))
        dp = C.c_double()
        st = T.SolveStats()
        _capi.check(self.L.cps_simple_gradient_dp(self.h, rtol, max_iter, C.byref(dp), _capi.dptr(grad),
                                                  C.byref(st)))
        self.last_adjoint = st.as_dict()
        return dp.value, grad

    def gradient_vjp(self, dJ_du=None, dJ_dv=None, dJ_dp=None, rtol=1e-10, max_iter=400):
        """d J / d gamma (ny, nx) of a caller's objective J(u, v, p) of the
        converged state, given dJ/du, dJ/dv, dJ/dp in the field shapes (None
        = 0), by the same discrete adjoint as gradient_dp."""
        arrs = []
        for name, a in (("u", dJ_du), ("v", dJ_dv), ("p", dJ_dp)):
            if a is None:
                arrs.append(None)
                continue
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.shape != T.field_shape(name, self.cfg.nx, self.cfg.ny):
                raise ValueError(f"dJ_d{name} must have shape {T.field_shape(name, self.cfg.nx, self.cfg.ny)}")
            arrs.append(a)
        grad = np.empty((self.cfg.ny, self.cfg.nx))
        st = T.SolveStats()
        ptr = [None if a is None else _capi.dptr(a) for a in arrs]
        _capi.check(self.L.cps_simple_gradient_vjp(self.h, *ptr, rtol, max_iter, _capi.dptr(grad), C.byref(st)))
        self.last_adjoint = st.as_dict()
        return grad

    def adjoint_options(self, preconditioner="lsc_adi", line_sweeps=2, omega=0.8, schur_iters=2):
        """Preconditioner of the adjoint solve: "lsc_adi" (default: x- and
        y-line momentum solves with damped alternating sweeps, least-squares-
        commutator Schur complement), "lsc_line" (x-lines only) or "simple"
        (1/aP0 momentum, pressure-correction Schur complement); schur_iters
        caps the multigrid-PCG iterations of each Schur-complement solve."""
        kinds = {"simple": T.CPS_ADJ_PRECOND_SIMPLE, "lsc_line": T.CPS_ADJ_PRECOND_LSC_LINE,
                 "lsc_adi": T.CPS_ADJ_PRECOND_LSC_ADI}
        _capi.check(self.L.cps_simple_adjoint_options(self.h, kinds[preconditioner], int(line_sweeps), float(omega),
                                                      int(schur_iters)))

    def state_residual(self):
        """max |R| of the U-momentum, V-momentum and continuity equations."""
        out = np.empty(3)
        _capi.check(self.L.cps_simple_state_residual(self.h, _capi.dptr(out)))
        return out

    def jacobian_apply(self, x, transpose=False):
        """J x or J^T x; x, y are dicts (or tuples) of u, v, p arrays."""
        shapes = [T.field_shape(k, self.cfg.nx, self.cfg.ny) for k in ("u", "v", "p")]
        parts = [x[k] for k in ("u", "v", "p")] if isinstance(x, dict) else list(x)
        flat = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float64).ravel() for a in parts]))
        y = np.empty_like(flat)
        _capi.check(self.L.cps_simple_jacobian_apply(self.h, 1 if transpose else 0, _capi.dptr(flat),
                                                     _capi.dptr(y)))
        out, off = {}, 0
        for k, sh in zip(("u", "v", "p"), shapes):
            m = sh[0] * sh[1]
            out[k] = y[off:off + m].reshape(sh)
            off += m
        return out

    # ---- measurement -----------------------------------------------------
    KERNELS = {"mg_smooth0": 0, "mom_jacobi": 1, "pcg_stencil": 2}

    def time_kernel(self, name):
        """Bracket every launch of one kernel with HIP events on the solver
        stream (None stops timing); clears the counters."""
        _capi.check(self.L.cps_simple_time_kernel(self.h, -1 if name is None else self.KERNELS[name]))

    def kernel_timing(self):
        n, us, by = C.c_int(), C.c_double(), C.c_double()
        _capi.check(self.L.cps_simple_kernel_timing(self.h, C.byref(n), C.byref(us), C.byref(by)))
        return {"launches": n.value, "avg_us": us.value, "bytes_per_launch": by.value}

    def synchronize(self):
        self.ctx.synchronize()

    def full_static_drop(self):
        return self.sampleBoundaryPatch(True)["avg_static"] - self.sampleBoundaryPatch(False)["avg_static"]

    def full_total_drop(self):
        return self.sampleBoundaryPatch(True)["avg_total"] - self.sampleBoundaryPatch(False)["avg_total"]


class Batch:
    """Many density fields stepped together.  Fields with the same grid and
    step settings form a size class; under the class plan (the default pick
    of `tune`) one launch per kernel of the SIMPLE chain steps every field of
    a class, one class iteration (or `unroll` of them) is one hipGraph and
    the classes' graphs run on streams of their own.  The per-field plan
    spreads the fields' own chains over several streams.  `fields` is a list
    of (cfg, gamma, inlet) tuples; iteration caps must allow graph capture
    (momentum and pressure iterations <= 32)."""

    PLANS = {"class": T.CPS_BATCH_CLASS, "per_field": T.CPS_BATCH_PER_FIELD}

    def __init__(self, fields, device=0):
        self.ctx = _capi.context(device)
        self.L = self.ctx.L
        n = len(fields)
        self.cfgs = (T.SimpleConfig * n)(*[f[0] for f in fields])
        self._keep = []
        gptr = (C.POINTER(C.c_double) * n)()
        iptr = (C.POINTER(C.c_double) * n)()
        for k, (cfg, gamma, inlet) in enumerate(fields):
            g = np.ascontiguousarray(gamma, dtype=np.float64)
            if g.shape != (cfg.ny, cfg.nx):
                raise ValueError(f"field {k}: gamma must be ny x nx")
            self._keep.append(g)
            gptr[k] = _capi.dptr(g)
            if inlet is not None:
                a = np.ascontiguousarray(inlet, dtype=np.float64)
                self._keep.append(a)
                iptr[k] = _capi.dptr(a)
        h = C.c_void_p()
        _capi.check(self.L.cps_batch_create(self.ctx.h, n, self.cfgs, gptr, iptr, C.byref(h)))
        self.h = h
        self.n = n
        self.streams = None
        self.us_per_step = None

    def close(self):
        if getattr(self, "h", None):
            self.L.cps_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def signature(self):
        """Key of the tuner cache: the size classes (grid, count) and the
        step settings that set the kernels' work."""
        cls = {}
        for c in self.cfgs:
            k = (c.nx, c.ny, c.precision, c.pressure_solver, c.max_momentum_iter, c.max_pressure_iter, c.max_refine,
                 c.convection_scheme)
            cls[k] = cls.get(k, 0) + 1
        return ";".join(f"{nx}x{ny}:{m} prec{pr} p{ps} mom{mi} pit{pi} ref{rf} conv{cv}"
                        for (nx, ny, pr, ps, mi, pi, rf, cv), m in sorted(cls.items()))

    def tune(self, streams=0, reps=3, cache=None):
        """Captures the batch iteration.  streams > 0: the per-field plan on
        that many streams.  streams = 0: the autotuner times the class plan
        (1 .. 4 chains per class, unroll 1, 4, 8) and the per-field plan
        (1 .. 32 streams), max(reps x unroll, 16) iterations each in two
        passes, the second in reverse order, after fields younger than 128
        iterations have reached it (all of which advance the fields), and
        keeps the fastest.
        cache: path of a JSON file of earlier picks keyed by signature(): a
        hit is captured without timing, a miss is tuned and stored.
        Returns (streams used, microseconds per batch iteration)."""
        import json
        import os

        key = self.signature()
        if streams == 0 and cache is not None and os.path.exists(cache):
            with open(cache) as fh:
                entry = json.load(fh).get("entries", {}).get(key)
            if entry is not None:
                self.plan(entry["plan"], entry.get("param", 0), entry.get("unroll", 1))
                self.us_per_step = entry.get("us_per_step")
                return self.streams, self.us_per_step
        ch, us = C.c_int(), C.c_double()
        _capi.check(self.L.cps_batch_tune(self.h, streams, reps, C.byref(ch), C.byref(us)))
        self.streams, self.us_per_step = ch.value, us.value
        if streams == 0 and cache is not None:
            info = self.plan_info()
            data = {"version": 1, "entries": {}}
            if os.path.exists(cache):
                with open(cache) as fh:
                    data = json.load(fh)
            data.setdefault("entries", {})[key] = dict(plan=info["plan"], param=info["param"], streams=info["streams"],
                                                       unroll=info["unroll"], us_per_step=info["us_per_step"])
            tmp = cache + ".tmp"
            with open(tmp, "w") as fh:
                json.dump(data, fh, indent=1, sort_keys=True)
            os.replace(tmp, cache)
        return self.streams, self.us_per_step

    def plan(self, plan="class", streams=0, unroll=1):
        """Captures a given plan without timing: "class" (streams = chains
        per size class, 0 = one) or "per_field" (streams = streams)."""
        p = self.PLANS[plan] if isinstance(plan, str) else int(plan)
        _capi.check(self.L.cps_batch_plan(self.h, p, streams, unroll))
        self.streams = self.plan_info()["streams"]

    def plan_info(self):
        """plan, param (chains per class or streams), streams (used by the
        graphs), unroll, classes (size classes), us_per_step (tuner's time)."""
        p, pa, ns, un, ncls, us = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_double()
        _capi.check(self.L.cps_batch_plan_info(self.h, C.byref(p), C.byref(pa), C.byref(ns), C.byref(un),
                                               C.byref(ncls), C.byref(us)))
        name = {v: k for k, v in self.PLANS.items()}[p.value]
        return dict(plan=name, param=pa.value, streams=ns.value, unroll=un.value, classes=ncls.value,
                    us_per_step=us.value)

    def step(self, n=1):
        _capi.check(self.L.cps_batch_step(self.h, n))

    def synchronize(self):
        _capi.check(self.L.cps_batch_synchronize(self.h))

    def stats(self, k):
        st = T.StepStats()
        _capi.check(self.L.cps_batch_field_stats(self.h, k, C.byref(st)))
        return st.as_dict()

    def get(self, k, name):
        cfg = self.cfgs[k]
        out = np.empty(T.field_shape(name, cfg.nx, cfg.ny))
        _capi.check(self.L.cps_batch_get_field(self.h, k, T.FIELDS[name], _capi.dptr(out)))
        return out

    def gradient_dp(self, k, rtol=1e-10, max_iter=4000):
        """(dP, d(dP)/dgamma) of field k at its current state (discrete adjoint)."""
        cfg = self.cfgs[k]
        grad = np.empty((cfg.ny, cfg.nx))
        dp = C.c_double()
        st = T.SolveStats()
        _capi.check(self.L.cps_batch_gradient_dp(self.h, k, rtol, max_iter, C.byref(dp), _capi.dptr(grad),
                                                 C.byref(st)))
        return dp.value, grad, st.as_dict()

    def gradients_dp(self, fields=None, rtol=1e-10, max_iter=4000, workers=8):
        """Gradients of several fields at once: one host thread per field in
        flight.  Each field's adjoint runs on the stream its field was given
        (field k on stream k mod 32 under the class plan, its owner stream
        under the per-field plan), so fields on different streams overlap and
        fields sharing a stream run one after another."""
        from concurrent.futures import ThreadPoolExecutor

        fields = list(range(self.n)) if fields is None else list(fields)
        self.synchronize()
        with ThreadPoolExecutor(max_workers=workers) as ex:
            return list(ex.map(lambda k: self.gradient_dp(k, rtol, max_iter), fields))

    def pressure_drop(self, k):
        a, b = T.Patch(), T.Patch()
        _capi.check(self.L.cps_batch_boundary_patch(self.h, k, 1, C.byref(a)))
        _capi.check(self.L.cps_batch_boundary_patch(self.h, k, 0, C.byref(b)))
        return a.avg_static - b.avg_static


def _csr_args(row_ptr, col_idx, values, rhs, solution):
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    ci = np.ascontiguousarray(col_idx, dtype=np.int32)
    val = np.ascontiguousarray(values, dtype=np.float32)
    b = np.ascontiguousarray(rhs, dtype=np.float32)
    if not (isinstance(solution, np.ndarray) and solution.dtype == np.float32 and solution.flags.c_contiguous):
        raise TypeError("solution must be a contiguous float32 array (updated in place)")
    n = len(b)
    if rp.shape != (n + 1,) or solution.shape != (n,) or len(ci) != rp[-1] or len(val) != rp[-1]:
        raise ValueError("inconsistent CSR sizes")
    return n, rp, ci, val, b


def pressure_solve_csr(row_ptr, col_idx, values, rhs, solution, tolerance, max_iterations, device=0):
    """Drop-in for amgcl_cuda::solve (cuda_solver.hpp:24-33) with the
    Jacobi-PCG semantics of solvePCG_CSR (pressure_solver.cpp:121-184);
    `solution` (float32) is the initial guess in and the result out.
    Returns the iteration count."""
    ctx = _capi.context(device)
    n, rp, ci, val, b = _csr_args(row_ptr, col_idx, values, rhs, solution)
    it = C.c_int()
    _capi.check(ctx.L.cps_pressure_solve_csr(ctx.h, n, _capi.iptr(rp), _capi.iptr(ci), _capi.fptr(val),
                                             _capi.fptr(b), _capi.fptr(solution), tolerance, max_iterations,
                                             C.byref(it)))
    return it.value


def pressure_solve_csr_noalloc(row_ptr, col_idx, values, rhs, solution, tolerance, max_iterations, device=0):
    """pressure_solve_csr without device allocation (cps_pressure_solve_csr_noalloc:
    static per-GPU buffers, n <= 524288, nnz <= 2621440; calls overlapping
    on one GPU race).  Returns the iteration count."""
    ctx = _capi.context(device)
    n, rp, ci, val, b = _csr_args(row_ptr, col_idx, values, rhs, solution)
    it = C.c_int()
    _capi.check(ctx.L.cps_pressure_solve_csr_noalloc(ctx.h, n, _capi.iptr(rp), _capi.iptr(ci), _capi.fptr(val),
                                                     _capi.fptr(b), _capi.fptr(solution), tolerance,
                                                     max_iterations, C.byref(it)))
    return it.value


def momentum_jacobi_csr(row_ptr, col_idx, values, rhs, solution, tolerance, max_iterations, device=0):
    """Drop-in for amgcl_cuda::solveMomentumU/V (cuda_solver.hpp:36-57) with
    the stop rule of solvers::cpu::solveMomentumJacobi (jacobi_solver.cpp:
    20-75).  Returns the sweep count."""
    ctx = _capi.context(device)
    n, rp, ci, val, b = _csr_args(row_ptr, col_idx, values, rhs, solution)
    it = C.c_int()
    _capi.check(ctx.L.cps_momentum_jacobi_csr(ctx.h, n, _capi.iptr(rp), _capi.iptr(ci), _capi.fptr(val),
                                              _capi.fptr(b), _capi.fptr(solution), tolerance, max_iterations,
                                              C.byref(it)))
    return it.value


def _pressure_drop_function():
    import torch

    class PressureDrop(torch.autograd.Function):
        """dP(gamma): forward iterates SIMPLE to a fixed point, backward is
        one discrete-adjoint solve at that state."""

        @staticmethod
        def forward(ctx, gamma, cfg, inlet, steps, rtol, device):
            g = gamma.detach().to("cpu", torch.float64).numpy()
            s = SIMPLE(cfg, g, inlet, device=device)
            s.step(steps)
            ctx.solver, ctx.rtol = s, rtol
            ctx.gdev, ctx.gdtype = gamma.device, gamma.dtype
            return gamma.new_tensor(s.full_static_drop())

        @staticmethod
        def backward(ctx, gout):
            import torch as _t

            _, grad = ctx.solver.gradient_dp(rtol=ctx.rtol, max_iter=4000)
            ctx.solver.close()
            g = _t.from_numpy(grad).to(ctx.gdev, ctx.gdtype)
            return gout * g, None, None, None, None, None

    return PressureDrop


_PD = None


def pressure_drop(gamma, cfg, inlet=None, steps=4000, rtol=1e-8, device=0):
    """Differentiable full static pressure drop (0-d tensor) of the density
    field `gamma` (ny x nx tensor): `steps` SIMPLE iterations from rest, then
    d(dP)/d(gamma) by the discrete adjoint on backward."""
    global _PD
    if _PD is None:
        _PD = _pressure_drop_function()
    return _PD.apply(gamma, cfg, inlet, steps, rtol, device)
