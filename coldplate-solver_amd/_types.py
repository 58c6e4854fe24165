"""ctypes mirrors of the plain C structs in include/coldplate_hip.h.

Kept free of any library loading so that both the HIP binding and the test
oracle binding can share them.
"""
import ctypes as C

CPS_OK = 0
CPS_ERR_ARG = 1
CPS_ERR_HIP = 2
CPS_ERR_STATE = 3
CPS_ERR_DIVERGED = 4

CPS_PRESSURE_SOR = 0
CPS_PRESSURE_PCG = 1
CPS_PRESSURE_AMGCL = 2        # falls back to GMG-PCG, stop ||r|| <= p_tol ||b||
CPS_PRESSURE_LDLT = 3         # falls back to GMG-PCG solved to 1e-12 ||b||
CPS_PRESSURE_AMGCL_CUDA = 4   # the reference default; as type 2
CPS_PRESSURE_GMG = 5
CPS_BATCH_CLASS = 0
CPS_BATCH_PER_FIELD = 1

CPS_ADJ_PRECOND_SIMPLE = 0
CPS_ADJ_PRECOND_LSC_LINE = 1
CPS_ADJ_PRECOND_LSC_ADI = 2

CPS_PREC_FP64 = 64
CPS_PREC_MIXED = 32

FIELDS = {
    "u": 0, "v": 1, "p": 2, "uStar": 3, "vStar": 4, "pStar": 5,
    "dE": 6, "dN": 7, "gamma": 8, "alpha": 9,
}


class SimpleConfig(C.Structure):
    """cps_simple_config; knob names follow SIMPLE.h (see the header)."""

    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int),
        ("hx", C.c_double), ("hy", C.c_double),
        ("rho", C.c_double), ("eta", C.c_double),
        ("u_in", C.c_double),
        ("convection_scheme", C.c_int),
        ("enable_25d", C.c_int),
        ("ht", C.c_double),
        ("conv_factor_25d", C.c_double),
        ("sink_mult_25d", C.c_double),
        ("brinkman_kmin", C.c_double),
        ("brinkman_q", C.c_double),
        ("uv_alpha", C.c_double), ("p_alpha", C.c_double),
        ("momentum_tol", C.c_double),
        ("max_momentum_iter", C.c_int),
        ("p_tol", C.c_double),
        ("max_pressure_iter", C.c_int),
        ("pressure_solver", C.c_int),
        ("precision", C.c_int),
        ("p_rtol", C.c_double),
        ("max_refine", C.c_int),
        ("epsilon", C.c_double),
        ("max_iterations", C.c_int),
        ("dp_fluid_tol", C.c_double),
        ("x_core_in", C.c_double), ("x_core_out", C.c_double),
        ("use_dp_convergence", C.c_int),
        ("dp_window", C.c_int),
        ("dp_percent", C.c_double),
        ("use_dp_slope_gate", C.c_int),
        ("dp_slope_window", C.c_int),
        ("dp_slope_max_delta", C.c_double),
        ("inlet_ramp", C.c_int),
        ("ramp_steps", C.c_int),
    ]

    def copy(self):
        c = SimpleConfig()
        C.pointer(c)[0] = self
        return c

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def default_config(**overrides):
    """Reference defaults of SIMPLE.h (SIMPLE.h:68-69,152-153,176-177,193,
    205-208,242-244,374-375,397-398).  The reference's default pressure
    solver (4, AMGCL-CUDA) is replaced by Jacobi-PCG (1), its CPU fallback."""
    c = SimpleConfig()
    c.nx, c.ny = 0, 0
    c.hx = c.hy = 0.0
    c.rho, c.eta = 997.0, 0.00089
    c.u_in = 0.0
    c.convection_scheme = 0
    c.enable_25d = 0
    c.ht = 0.0
    c.conv_factor_25d = 1.2
    c.sink_mult_25d = 1.0
    c.brinkman_kmin = 1e-14
    c.brinkman_q = 1.0
    c.uv_alpha, c.p_alpha = 0.7, 0.3
    c.momentum_tol = 1e-6
    c.max_momentum_iter = 400
    c.p_tol = 1e-6
    c.max_pressure_iter = 30
    c.pressure_solver = CPS_PRESSURE_PCG
    c.precision = CPS_PREC_FP64
    c.p_rtol = 0.0
    c.max_refine = 4
    c.epsilon = 1e-7
    c.max_iterations = 1000
    c.dp_fluid_tol = 1e-6
    c.x_core_in, c.x_core_out = 0.01, 0.05
    c.use_dp_convergence = 1
    c.dp_window = 500
    c.dp_percent = 1.0
    c.use_dp_slope_gate = 0
    c.dp_slope_window = 1000
    c.dp_slope_max_delta = 10.0
    c.inlet_ramp = 0
    c.ramp_steps = 1000
    for k, v in overrides.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


class StepStats(C.Structure):
    _fields_ = [
        ("error", C.c_double),
        ("resid_u", C.c_double), ("resid_v", C.c_double), ("resid_mass", C.c_double),
        ("resid_u_max", C.c_double), ("resid_v_max", C.c_double), ("resid_mass_max", C.c_double),
        ("u_iters", C.c_int), ("v_iters", C.c_int), ("p_iters", C.c_int),
        ("iteration", C.c_int),
        ("dp_core_total", C.c_double), ("dp_core_static", C.c_double),
        ("dp_full_total", C.c_double), ("dp_full_static", C.c_double),
        ("dp_converged", C.c_int),
        ("pad_", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class IterLog(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("resid_mass", "resid_u", "resid_v", "dp_core_total", "dp_full_total",
                                          "dp_core_static", "dp_full_static")] + [
        ("p_iters", C.c_int), ("pad_", C.c_int)]


class Patch(C.Structure):
    _fields_ = [
        ("area", C.c_double), ("mass_flux", C.c_double), ("avg_static", C.c_double),
        ("avg_dynamic", C.c_double), ("avg_total", C.c_double), ("valid", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def field_shape(name, nx, ny):
    """Padded staggered shapes, M = ny+1, N = nx+1 (SIMPLE.cpp:53-54, 252-269)."""
    M, N = ny + 1, nx + 1
    return {
        "u": (M + 1, N), "uStar": (M + 1, N), "dE": (M + 1, N),
        "v": (M, N + 1), "vStar": (M, N + 1), "dN": (M, N + 1),
        "p": (M + 1, N + 1), "pStar": (M + 1, N + 1),
        "gamma": (M, N), "alpha": (M, N),
    }[name]


# ---------------------------------------------------------------------------
# 3-D thermal solve

CPS_THERMAL_OBJ_T_AVG_BASE = 0
CPS_THERMAL_OBJ_PNORM_L0 = 1
THERMAL_OBJECTIVES = {"t_avg_base": CPS_THERMAL_OBJ_T_AVG_BASE, "pnorm": CPS_THERMAL_OBJ_PNORM_L0}


class ThermalConfig(C.Structure):
    """cps_thermal_config; property names follow Params (thermal_solver.cpp:148-192)."""

    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double),
        ("ht", C.c_double),
        ("k_s", C.c_double), ("k_f", C.c_double), ("rho", C.c_double), ("cp", C.c_double),
        ("t_in", C.c_double),
        ("k_mode", C.c_int),
        ("qk", C.c_double),
        ("precision", C.c_int),
        ("rtol", C.c_double),
        ("max_iter", C.c_int),
        ("max_refine", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("refinements", C.c_int), ("rel_residual", C.c_double),
                ("converged", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ThermalMetrics(C.Structure):
    """cps_thermal_metrics (ThermalMetrics of thermal_output.cpp:16-44, plus
    Tb_max / Tb_pnorm and the energy balance)."""

    _fields_ = [(k, C.c_double) for k in (
        "t_avg_base", "t_avg_fluid", "t_avg_solid", "t_avg_global",
        "t_std_base", "t_std_fluid", "t_std_global",
        "t_range_base", "t_range_fluid", "t_range_global",
        "t_uniformity_base", "t_uniformity_global",
        "tb_max", "tb_pnorm", "q_in", "q_out", "t_out", "imbalance_pct",
        "grad_max", "grad_avg", "grad_max_base", "grad_avg_base")] + [
        (k, C.c_int) for k in ("grad_max_i", "grad_max_j", "grad_max_k", "pad_")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def thermal_config(stack, precision=CPS_PREC_FP64, rtol=1e-6, max_iter=1000, max_refine=8, k_mode=1):
    """cps_thermal_config from a layer-stack description (configs.thermal_stack);
    solver knobs default to Params (thermal_solver.cpp:167-168)."""
    c = ThermalConfig()
    c.nx, c.ny, c.nz = int(stack["nx"]), int(stack["ny"]), len(stack["dz"])
    c.dx, c.dy = float(stack["dx"]), float(stack["dy"])
    c.ht = float(stack["ht"])
    c.k_s, c.k_f = float(stack["k_s"]), float(stack["k_f"])
    c.rho, c.cp, c.t_in = float(stack["rho"]), float(stack["cp"]), float(stack["t_in"])
    c.k_mode = int(stack.get("k_mode", k_mode))
    c.qk = float(stack["qk"])
    c.precision = precision
    c.rtol = rtol
    c.max_iter = max_iter
    c.max_refine = max_refine
    return c
