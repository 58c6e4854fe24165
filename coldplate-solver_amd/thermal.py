"""Host-side mirror of the reference 3-D conjugate-heat-transfer solver on the
HIP path.

`ThermalSolver3D` follows ThermalSolver3D (ThermalSolver/thermal_solver.cpp:
223-620): construction takes the layer stack, density field, cell-centred
velocity and power map; `solve` assembles and solves A T = b (the reference's
assembly + AMGCL IDR(8) solve, :422-553) to the same relative tolerance;
`compute_pnorm_objective` / `compute_optimization_results` (:345-381),
`metrics` (compute_thermal_metrics, thermal_output.cpp:135-270) and
`energy_balance` (:592-620) post-process the device-resident field.
`gradient` adds what the reference leaves as a stub (thermal_adjoint.cpp):
dJ/dgamma of T_avg_base or the layer-0 p-norm by one adjoint solve, and
`ThermalObjective` exposes it as a torch.autograd.Function.

Every call goes through the C ABI of libcoldplate_hip.so.
"""
import ctypes as C

import numpy as np

from . import _capi
from . import _types
from . import _types as T


def _f64(a, shape, name):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.shape != shape:
        raise ValueError(f"{name} must have shape {shape}, got {a.shape}")
    return a


class ThermalSolver3D:
    def __init__(self, stack, gamma, u, v, q0=None, precision=T.CPS_PREC_FP64, rtol=1e-6, max_iter=1000,
                 max_refine=8, device=0):
        self.ctx = _capi.context(device)
        self.L = self.ctx.L
        self.cfg = T.thermal_config(stack, precision=precision, rtol=rtol, max_iter=max_iter, max_refine=max_refine)
        self.nx, self.ny, self.nz = self.cfg.nx, self.cfg.ny, self.cfg.nz
        s2 = (self.ny, self.nx)
        dz = np.ascontiguousarray(stack["dz"], dtype=np.float64)
        ch = np.ascontiguousarray(stack["channel"], dtype=np.int32)
        if dz.shape != (self.nz,) or ch.shape != (self.nz,):
            raise ValueError("dz and channel need one entry per layer")
        g = _f64(gamma, s2, "gamma")
        uu, vv = _f64(u, s2, "u"), _f64(v, s2, "v")
        q = _f64(stack["q0"] if q0 is None else q0, s2, "q0")
        h = C.c_void_p()
        _capi.check(self.L.cps_thermal_create(self.ctx.h, C.byref(self.cfg), _capi.dptr(dz), _capi.iptr(ch),
                                              _capi.dptr(g), _capi.dptr(uu), _capi.dptr(vv), _capi.dptr(q),
                                              C.byref(h)))
        self.h = h
        self.dz, self.channel = dz, ch
        self.last = None

    def close(self):
        if getattr(self, "h", None):
            self.L.cps_thermal_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- inputs ----------------------------------------------------------
    def update(self, gamma=None, u=None, v=None, q0=None):
        s2 = (self.ny, self.nx)
        arrs = [None if a is None else _f64(a, s2, n) for a, n in ((gamma, "gamma"), (u, "u"), (v, "v"), (q0, "q0"))]
        ptrs = [None if a is None else _capi.dptr(a) for a in arrs]
        _capi.check(self.L.cps_thermal_update(self.h, *ptrs))

    # ---- solve -----------------------------------------------------------
    def solve(self, T0=None):
        st = T.SolveStats()
        t0 = None
        if T0 is not None:
            t0 = _f64(T0, (self.nz, self.ny, self.nx), "T0")
        _capi.check(self.L.cps_thermal_solve(self.h, None if t0 is None else _capi.dptr(t0), C.byref(st)))
        self.last = st.as_dict()
        return self.last

    @property
    def T(self):
        out = np.empty((self.nz, self.ny, self.nx))
        _capi.check(self.L.cps_thermal_get_temperature(self.h, _capi.dptr(out)))
        return out

    def operator(self):
        """Assembled stencil [d, w, e, s, n, b, t] (7, nz, ny, nx) and b."""
        coef = np.empty((7, self.nz, self.ny, self.nx))
        b = np.empty((self.nz, self.ny, self.nx))
        _capi.check(self.L.cps_thermal_get_operator(self.h, _capi.dptr(coef), _capi.dptr(b)))
        return coef, b

    # ---- post-processing -------------------------------------------------
    def metrics(self, pnorm_exp=10.0):
        m = T.ThermalMetrics()
        _capi.check(self.L.cps_thermal_get_metrics(self.h, float(pnorm_exp), C.byref(m)))
        return m.as_dict()

    def compute_pnorm_objective(self, pnorm_exp=10):
        return self.metrics(pnorm_exp)["tb_pnorm"]

    def compute_optimization_results(self, pnorm_exp=10):
        m = self.metrics(pnorm_exp)
        return {"Tb_max": m["tb_max"], "Tb_pnorm": m["tb_pnorm"], "energy_imbalance": m["imbalance_pct"]}

    def energy_balance(self):
        m = self.metrics()
        return {k: m[k] for k in ("q_in", "q_out", "t_out", "imbalance_pct")}

    # ---- adjoint ---------------------------------------------------------
    def solve_adjoint(self, rhs):
        rhs = _f64(rhs, (self.nz, self.ny, self.nx), "rhs")
        lam = np.empty_like(rhs)
        st = T.SolveStats()
        _capi.check(self.L.cps_thermal_solve_adjoint(self.h, _capi.dptr(rhs), _capi.dptr(lam), C.byref(st)))
        self.last_adjoint = st.as_dict()
        return lam

    def sensitivities(self, lam):
        """-lambda^T d(A T - b)/d(gamma, u, v) at the last solve's T for an
        adjoint lambda (nz, ny, nx): (dJ/dgamma, dJ/du, dJ/dv), each (ny, nx),
        u and v being the cell-centred channel velocity."""
        lam = _f64(lam, (self.nz, self.ny, self.nx), "lambda")
        out = [np.empty((self.ny, self.nx)) for _ in range(3)]
        _capi.check(self.L.cps_thermal_sensitivities(self.h, _capi.dptr(lam), *[_capi.dptr(a) for a in out]))
        return tuple(out)

    def gradient(self, objective="t_avg_base", pnorm_exp=8.0):
        """(J, dJ/dgamma (ny, nx)) of the last solve."""
        grad = np.empty((self.ny, self.nx))
        J = C.c_double()
        st = T.SolveStats()
        _capi.check(self.L.cps_thermal_gradient(self.h, T.THERMAL_OBJECTIVES[objective], float(pnorm_exp),
                                                _capi.dptr(grad), C.byref(J), C.byref(st)))
        self.last_adjoint = st.as_dict()
        return J.value, grad


def solve_csr(nx, ny, nz, row_ptr, col_idx, values, rhs, T0, rtol=1e-6, max_iter=1000, precision=T.CPS_PREC_FP64,
              device=0):
    """Drop-in for the reference's AMGCL call (thermal_solver.cpp:534-552):
    solves the caller's assembled 7-point matrix (CSR, rows numbered
    (k ny + j) nx + i) from T0 to ||b - A T|| <= rtol ||b||.  Returns
    (T (n,), (iterations, relative residual))."""
    ctx = _capi.context(device)
    n = nx * ny * nz
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    ci = np.ascontiguousarray(col_idx, dtype=np.int32)
    val = np.ascontiguousarray(values, dtype=np.float64)
    b = np.ascontiguousarray(rhs, dtype=np.float64)
    x = np.array(T0, dtype=np.float64).reshape(-1).copy()
    if rp.shape != (n + 1,) or b.shape != (n,) or x.shape != (n,) or ci.shape != val.shape:
        raise ValueError("CSR arrays do not match the grid")
    st = T.SolveStats()
    _capi.check(ctx.L.cps_thermal_solve_csr(ctx.h, nx, ny, nz, _capi.iptr(rp), _capi.iptr(ci), _capi.dptr(val),
                                            _capi.dptr(b), _capi.dptr(x), rtol, max_iter, precision, C.byref(st)))
    d = st.as_dict()
    return x, (d["iterations"], d["rel_residual"])


def surrogate_adjoint(k, P=None, T=None, gamma=None, k_s=400.0, k_f=0.6, q_k=1.0, rtol=1e-12, device=0):
    """The 2-D conduction surrogate adjoint of AdjointMethod/thermal_adjoint.py
    (solve_thermal_adjoint) on the device: T (solved from P unless given),
    lambda, sens = dQ/dk, and with gamma dJ/dgamma = sens dk/dgamma (RAMP),
    as run_coldplate_adjoint.py applies it.  Arrays are ny x nx."""
    ctx = _capi.context(device)
    k = np.ascontiguousarray(k, dtype=np.float64)
    ny, nx = k.shape
    arr = {}
    for name, a in (("P", P), ("T", T), ("gamma", gamma)):
        arr[name] = None if a is None else _f64(a, (ny, nx), name)
    if arr["P"] is None and arr["T"] is None:
        raise ValueError("need P or T")
    outT, lam, sens = np.empty((ny, nx)), np.empty((ny, nx)), np.empty((ny, nx))
    dJ = np.empty((ny, nx)) if gamma is not None else None
    st = _types.SolveStats()  # (the argument T shadows the module alias)
    p = lambda a: None if a is None else _capi.dptr(a)  # noqa: E731
    _capi.check(ctx.L.cps_surrogate_adjoint(ctx.h, nx, ny, _capi.dptr(k), p(arr["P"]), p(arr["T"]), rtol,
                                            p(arr["gamma"]), k_s, k_f, q_k, _capi.dptr(outT), _capi.dptr(lam),
                                            _capi.dptr(sens), p(dJ), C.byref(st)))
    out = {"T": outT, "lambda": lam, "sens": sens, "stats": st.as_dict()}
    if dJ is not None:
        out["dJ_dgamma"] = dJ
    return out


def _torch():
    import torch
    return torch


def _objective_function():
    torch = _torch()

    class ThermalObjective(torch.autograd.Function):
        """J(gamma) of a ThermalSolver3D: forward = update + solve + metric,
        backward = one adjoint solve (dJ/dgamma)."""

        @staticmethod
        def forward(ctx, gamma, solver, objective, pnorm_exp):
            g = gamma.detach().to("cpu", torch.float64).numpy()
            solver.update(gamma=g)
            solver.solve()
            J, grad = solver.gradient(objective, pnorm_exp)
            ctx.save_for_backward(torch.from_numpy(grad).to(gamma.device, gamma.dtype))
            return gamma.new_tensor(J)

        @staticmethod
        def backward(ctx, gout):
            (grad,) = ctx.saved_tensors
            return gout * grad, None, None, None

    return ThermalObjective


_TO = None


def thermal_objective(gamma, solver, objective="t_avg_base", pnorm_exp=8.0):
    """Differentiable J(gamma) (a 0-d tensor) through the HIP thermal solve."""
    global _TO
    if _TO is None:
        _TO = _objective_function()
    return _TO.apply(gamma, solver, objective, pnorm_exp)
