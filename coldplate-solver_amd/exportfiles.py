"""The reference's ExportFiles text formats, so that a user of
GeometryGenerator_V6.py / run_solvers.py drives the device solvers directly
instead of through simple.exe and thermal_solver.exe subprocesses.

Inputs (written by GeometryGenerator_V6.py:507-518):
  fluid_params.txt      "Ny Nx hy hx U_in N_in_buffer N_out_buffer [Ht]"
                        (SIMPLE.cpp:31-60 reads it; extra columns ignored)
  geometry_fluid.txt    Ny x Nx cellType, 0 = fluid, 1 = solid, continuous
                        in between (loadTopology SIMPLE.cpp:636-690,
                        loadDensityField :694-746)
  thermal_params.txt    "rows cols dx dy q_flux[W/m^2] Ht" as
                        thermal_solver.cpp:700-706 reads it
  geometry_thermal.txt  rows x cols cellType of the heat-sink region
  power_map.txt         optional, W/cm^2 (thermal_solver.cpp:750-770)

Outputs (same names, layouts and number formats as the reference):
  u.txt, v.txt, p.txt, u_full.txt, v_full.txt, pressure_full.txt,
  u_thermal.txt, v_thermal.txt, pressure_thermal.txt, p_gradient.txt
                        (SIMPLE::saveAll, output.cpp:48-200)
  residuals.txt, pressure_drop_history.txt
                        (initLogFiles / writeIterationLogs, output.cpp:302-360)
  thermal_metrics.txt, k_thermal.txt, T_bottom.txt
                        (thermal_output.cpp:323-370, 524-560, 619-650)
  objective.txt, dJ_dgamma.txt (thermal_solver.cpp:383-421, on request)
  sensitivity_field.txt, dJ_dgamma.txt
                        (run_coldplate_adjoint.py:110-150, by run_adjoint)

The solvers are the device ones (SIMPLE, ThermalSolver3D); this module only
reads and writes text.
"""
import os
import re

import numpy as np

from . import _types as T

# thermal_solver.cpp:148-192 (Params): 2 base-plate layers over H_b, 7
# channel layers over Ht, no lid
THERMAL_NZ_SOLID = 2
THERMAL_NZ_FLUID = 7
THERMAL_H_BASE = 0.0005


# ---------------------------------------------------------------------------
# readers

def read_matrix(path):
    """read_2d (thermal_solver.cpp:210-220): whitespace-separated rows; blank
    lines and lines starting with '#' skipped.  Returns a list of rows (they
    may be ragged, as in the reference)."""
    rows = []
    with open(path) as f:
        for line in f:
            if not line.strip() or line.startswith("#"):
                continue
            vals = [float(t) for t in line.split()]
            if vals:
                rows.append(vals)
    return rows


def _stream(path):
    """All numbers of a file in order (operator>> over whitespace)."""
    with open(path) as f:
        return [float(t) for t in f.read().split()]


def read_fluid_params(path):
    """loadParameters (SIMPLE.cpp:31-60): Ny Nx hy hx U_in N_in N_out, then an
    optional Ht; further columns are ignored."""
    v = _stream(path)
    if len(v) < 7:
        raise ValueError(f"{path}: expected 'Ny Nx hy hx U_in N_in N_out [Ht]'")
    ny, nx = int(v[0]), int(v[1])
    if ny <= 0 or nx <= 0 or v[2] <= 0 or v[3] <= 0:
        raise ValueError(f"{path}: non-positive grid size or spacing")
    return dict(ny=ny, nx=nx, hy=v[2], hx=v[3], u_in=v[4], n_in_buffer=int(v[5]), n_out_buffer=int(v[6]),
                ht=v[7] if len(v) > 7 else 0.0)


def read_celltype(path, ny, nx):
    """loadTopology (SIMPLE.cpp:636-667): ny x nx values read in row order;
    entries missing at the end of the file stay 0 (fluid), extra ones are
    ignored, a missing file means all fluid."""
    ct = np.zeros(ny * nx)
    if os.path.exists(path):
        v = _stream(path)[: ny * nx]
        ct[: len(v)] = v
    return ct.reshape(ny, nx)


def gamma_from_celltype(ct):
    """loadDensityField (SIMPLE.cpp:694-717): gamma = 1 - clamp(cellType, 0, 1)."""
    return 1.0 - np.clip(np.asarray(ct, dtype=np.float64), 0.0, 1.0)


def fluid_case(export_dir, **overrides):
    """SIMPLE::SIMPLE (SIMPLE.cpp:18-29) from an ExportFiles directory:
    (config, gamma, meta) with the SIMPLE.h defaults for every knob the
    files do not set; overrides set config fields (e.g. pressure_solver,
    precision).  2.5-D needs Ht > 0 (SIMPLE.cpp:66-72)."""
    prm = read_fluid_params(os.path.join(export_dir, "fluid_params.txt"))
    cfg = T.default_config(nx=prm["nx"], ny=prm["ny"], hx=prm["hx"], hy=prm["hy"], u_in=prm["u_in"], ht=prm["ht"])
    for k, v in overrides.items():
        if not hasattr(cfg, k):
            raise KeyError(k)
        setattr(cfg, k, v)
    if cfg.enable_25d and cfg.ht <= 0.0:
        cfg.enable_25d = 0
    ct = read_celltype(os.path.join(export_dir, "geometry_fluid.txt"), prm["ny"], prm["nx"])
    return cfg, gamma_from_celltype(ct), prm


def read_thermal_params(path):
    """thermal_solver.cpp:700-706: 'pf >> M >> N >> dx >> dy >> qf >> Ht'.
    GeometryGenerator_V6.py:515 writes 'rows cols dy dx q Ht', so the third
    number is taken as dx and the fourth as dy, as the reference does."""
    v = _stream(path)
    if len(v) < 6:
        raise ValueError(f"{path}: expected 'rows cols dx dy q_flux Ht'")
    return dict(rows=int(v[0]), cols=int(v[1]), dx=v[2], dy=v[3], q_flux=v[4], ht=v[5])


def thermal_case(export_dir, qk=1.0, k_mode=1, k_s=400.0, k_f=0.6, rho=998.0, cp=4180.0, t_in=30.0):
    """thermal_solver.cpp main (:675-772) up to the solve: the stack of
    Params (2 base layers over H_b = 0.5 mm, 7 channel layers over Ht), the
    grid from geometry_thermal.txt (gamma = 1 - value, unclamped), u / v
    from u_thermal.txt / v_thermal.txt (zeros where missing), the bottom
    flux from power_map.txt in W/cm^2 (x 1e4) when its shape matches, else
    the uniform q_flux.  Returns (stack, gamma, u, v)."""
    prm = read_thermal_params(os.path.join(export_dir, "thermal_params.txt"))
    gam = read_matrix(os.path.join(export_dir, "geometry_thermal.txt"))
    if not gam:
        raise ValueError("could not load geometry_thermal.txt")
    ny, nx = len(gam), len(gam[0])
    gamma = 1.0 - np.array([r[:nx] + [0.0] * (nx - len(r[:nx])) for r in gam])

    def field(name):
        out = np.zeros((ny, nx))
        path = os.path.join(export_dir, name)
        if os.path.exists(path):
            for j, r in enumerate(read_matrix(path)[:ny]):
                out[j, : min(nx, len(r))] = r[:nx]
        return out

    u, v = field("u_thermal.txt"), field("v_thermal.txt")
    q0 = np.full((ny, nx), prm["q_flux"])
    pm_path = os.path.join(export_dir, "power_map.txt")
    if os.path.exists(pm_path):
        pm = read_matrix(pm_path)
        if len(pm) == ny and len(pm[0]) == nx:
            q0 = np.array(pm) * 10000.0
    dz = [THERMAL_H_BASE / THERMAL_NZ_SOLID] * THERMAL_NZ_SOLID + [prm["ht"] / THERMAL_NZ_FLUID] * THERMAL_NZ_FLUID
    channel = [0] * THERMAL_NZ_SOLID + [1] * THERMAL_NZ_FLUID
    stack = dict(nx=nx, ny=ny, dx=prm["dx"], dy=prm["dy"], dz=np.array(dz), channel=np.array(channel, np.int32),
                 ht=prm["ht"], k_s=k_s, k_f=k_f, rho=rho, cp=cp, t_in=t_in, qk=qk, k_mode=k_mode, q0=q0)
    return stack, gamma, u, v


# ---------------------------------------------------------------------------
# writers

def _g(x):
    """operator<< of a float with the default stream format (6 significant
    digits, %g)."""
    return "%g" % x


def save_matrix(path, a, fmt=_g):
    """saveMatrix (output.cpp:48-69): one row per line, tab-separated."""
    a = np.atleast_2d(np.asarray(a, dtype=np.float64))
    with open(path, "w") as f:
        for row in a:
            f.write("\t".join(fmt(x) for x in row) + "\n")


def _fixed(digits):
    return lambda x: "%.*f" % (digits, x)


def cell_fields(u, v, p):
    """Cell-centred u, v, p of the physical cells (output.cpp:96-120)."""
    ny, nx = p.shape[0] - 2, p.shape[1] - 2
    uc = 0.5 * (u[1:ny + 1, 0:nx] + u[1:ny + 1, 1:nx + 1])
    vc = 0.5 * (v[0:ny, 1:nx + 1] + v[1:ny + 1, 1:nx + 1])
    pc = p[1:ny + 1, 1:nx + 1].copy()
    return uc, vc, pc


def pressure_gradient(pc, hx, hy):
    """|grad p| per cell (output.cpp:137-165): central differences inside,
    first differences at the edges."""
    ny, nx = pc.shape
    gx = np.zeros_like(pc)
    gy = np.zeros_like(pc)
    if nx > 1:
        gx[:, 1:-1] = (pc[:, 2:] - pc[:, :-2]) / (2.0 * hx)
        gx[:, 0] = (pc[:, 1] - pc[:, 0]) / hx
        gx[:, -1] = (pc[:, -1] - pc[:, -2]) / hx
    if ny > 1:
        gy[1:-1] = (pc[2:] - pc[:-2]) / (2.0 * hy)
        gy[0] = (pc[1] - pc[0]) / hy
        gy[-1] = (pc[-1] - pc[-2]) / hy
    return np.sqrt(gx * gx + gy * gy)


def save_fluid_outputs(export_dir, solver, meta):
    """The text part of SIMPLE::saveAll (output.cpp:74-200): staggered u, v,
    p; cell-centred full-domain fields; the heat-sink crop (columns
    N_in_buffer .. Nx - N_out_buffer - 1) of u, v, p and |grad p|.  The
    momentum-residual fields and the VTK file are not written."""
    u, v, p = solver.get("u"), solver.get("v"), solver.get("p")
    d = export_dir
    save_matrix(os.path.join(d, "u.txt"), u)
    save_matrix(os.path.join(d, "v.txt"), v)
    save_matrix(os.path.join(d, "p.txt"), p)
    uc, vc, pc = cell_fields(u, v, p)
    save_matrix(os.path.join(d, "u_full.txt"), uc)
    save_matrix(os.path.join(d, "v_full.txt"), vc)
    save_matrix(os.path.join(d, "pressure_full.txt"), pc)
    gp = pressure_gradient(pc, solver.cfg.hx, solver.cfg.hy)
    nx = pc.shape[1]
    n_th = nx - meta["n_in_buffer"] - meta["n_out_buffer"]
    if n_th > 0:
        sl = slice(meta["n_in_buffer"], meta["n_in_buffer"] + n_th)
        save_matrix(os.path.join(d, "u_thermal.txt"), uc[:, sl])
        save_matrix(os.path.join(d, "v_thermal.txt"), vc[:, sl])
        save_matrix(os.path.join(d, "pressure_thermal.txt"), pc[:, sl])
        save_matrix(os.path.join(d, "p_gradient.txt"), gp[:, sl])
    return n_th


def save_iteration_logs(export_dir, hist):
    """residuals.txt and pressure_drop_history.txt (output.cpp:302-360) from
    SIMPLE.runIterationsLogged's history.  Normalised residual columns are
    0 and NormEnabled 0 (enableNormalizedResiduals is off by default); the
    CFL columns are 0 (the pseudo-time controller is not built)."""
    n = len(hist["resid_mass"])
    with open(os.path.join(export_dir, "residuals.txt"), "w") as f:
        f.write("Iter MassRMS U_RMS V_RMS MassRMSn U_RMSn V_RMSn NormEnabled "
                "Core_dP_AfterInletBuffer(Pa) Full_dP_FullSystem(Pa) CFL CFL_ratio\n")
        for k in range(n):
            f.write(" ".join([str(k), _g(hist["resid_mass"][k]), _g(hist["resid_u"][k]), _g(hist["resid_v"][k]),
                              "0", "0", "0", "0", _g(hist["dp_core_total"][k]), _g(hist["dp_full_total"][k]),
                              "0", "0"]) + "\n")
    with open(os.path.join(export_dir, "pressure_drop_history.txt"), "w") as f:
        f.write("Iter Core_Total(Pa) Full_Total(Pa) Core_Static(Pa) Full_Static(Pa)\n")
        for k in range(n):
            f.write(" ".join([str(k), _g(hist["dp_core_total"][k]), _g(hist["dp_full_total"][k]),
                              _g(hist["dp_core_static"][k]), _g(hist["dp_full_static"][k])]) + "\n")


_METRIC_LINES = [
    "# Thermal Metrics - Volume-Weighted Statistics", "# All temperatures in [C]", "#", "# Average Temperatures",
    ("T_avg_base    ", "t_avg_base"), ("T_avg_fluid   ", "t_avg_fluid"), ("T_avg_solid   ", "t_avg_solid"),
    ("T_avg_global  ", "t_avg_global"), "#", "# Standard Deviations", ("T_std_base    ", "t_std_base"),
    ("T_std_fluid   ", "t_std_fluid"), ("T_std_global  ", "t_std_global"), "#", "# Temperature Ranges (max - min)",
    ("T_range_base  ", "t_range_base"), ("T_range_fluid ", "t_range_fluid"), ("T_range_global", "t_range_global"),
    "#", "# Uniformity Indices (1 = uniform, 0 = non-uniform)", ("T_uniformity_base   ", "t_uniformity_base"),
    ("T_uniformity_global ", "t_uniformity_global"), "#", "# Thermal Gradients [C/m]",
    ("grad_max          ", "grad_max"), ("grad_avg          ", "grad_avg"), ("grad_max_base     ", "grad_max_base"),
    ("grad_avg_base     ", "grad_avg_base"),
]


def save_thermal_metrics(export_dir, m):
    """thermal_metrics.txt (save_thermal_metrics, thermal_output.cpp:323-367):
    fixed notation, 4 decimals."""
    with open(os.path.join(export_dir, "thermal_metrics.txt"), "w") as f:
        for line in _METRIC_LINES:
            if isinstance(line, str):
                f.write(line + "\n")
            else:
                f.write("%s%.4f\n" % (line[0], m[line[1]]))
        f.write("grad_max_loc      %d %d %d\n" % (m["grad_max_i"], m["grad_max_j"], m["grad_max_k"]))


_LEADING_NUMBER = re.compile(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?")


def parse_thermal_metrics(path):
    """name -> value as run_solvers.parse_thermal_metrics (run_solvers.py:
    158-187) reads thermal_metrics.txt: a name, whitespace, then the number
    the next token starts with.  So 'T_range_global', written without a
    space before its value (thermal_output.cpp:349), is absent, and
    grad_max_loc maps to its i index, as in the reference pipeline."""
    out = {}
    with open(path) as f:
        for line in f:
            t = line.strip().split(None, 1)
            if len(t) < 2 or t[0].startswith("#"):
                continue
            m = _LEADING_NUMBER.match(t[1])
            if m:
                out[t[0]] = float(m.group(0))
    return out


def k_field(gamma, k_s, k_f, qk):
    """RAMP k(gamma) at the clamped gamma (save_k_field, thermal_output.cpp:530-555)."""
    g = np.clip(np.asarray(gamma, dtype=np.float64), 0.0, 1.0)
    return k_f + (k_s - k_f) * (1.0 - g) / (1.0 + qk * g)


def save_thermal_outputs(export_dir, solver, stack, gamma):
    """What thermal_solver.cpp main writes after the solve (:774-790):
    thermal_metrics.txt, k_thermal.txt (6 decimals), T_bottom.txt (layer
    k = 0, 4 decimals).  Returns the metrics (p-norm with p = 10)."""
    m = solver.metrics(10.0)
    d = export_dir
    save_thermal_metrics(d, m)
    save_matrix(os.path.join(d, "k_thermal.txt"), k_field(gamma, stack["k_s"], stack["k_f"], stack["qk"]),
                fmt=_fixed(6))
    save_matrix(os.path.join(d, "T_bottom.txt"), solver.T[0], fmt=_fixed(4))
    return m


def save_optimization_outputs(export_dir, metrics, dJ_dgamma):
    """objective.txt (Tb_max, Tb_pnorm, '10', energy imbalance) and
    dJ_dgamma.txt, 10 decimals (save_optimization_outputs,
    thermal_solver.cpp:383-421; the reference's main does not call it).
    metrics: ThermalSolver3D.metrics(10.0); dJ_dgamma: ny x nx, e.g. from
    ThermalSolver3D.gradient("pnorm", 10.0)."""
    with open(os.path.join(export_dir, "objective.txt"), "w") as f:
        f.write("%.10f\n%.10f\n10\n%.10f\n" % (metrics["tb_max"], metrics["tb_pnorm"], metrics["imbalance_pct"]))
    save_matrix(os.path.join(export_dir, "dJ_dgamma.txt"), dJ_dgamma, fmt=_fixed(10))


# ---------------------------------------------------------------------------
# the optimiser's adjoint step and objective

def run_adjoint(export_dir, q_k=1.0, k_s=400.0, k_f=0.6, device=0):
    """run_coldplate_adjoint.main (run_coldplate_adjoint.py:73-161) on the
    device: the 2-D surrogate adjoint with T = T_bottom.txt, P =
    pressure_thermal.txt and k = k_thermal.txt (external-T mode), dQ/dk to
    sensitivity_field.txt and, when geometry_thermal.txt has the same shape,
    dJ/dgamma = dQ/dk * dk/dgamma (RAMP at gamma = 1 - clip(value)) to
    dJ_dgamma.txt; both tab-separated in %.10e.  Returns (sens, dJ or None)."""
    from .thermal import surrogate_adjoint

    load = lambda n: np.atleast_2d(np.loadtxt(os.path.join(export_dir, n), delimiter="\t"))  # noqa: E731
    Tb, P, k = load("T_bottom.txt"), load("pressure_thermal.txt"), load("k_thermal.txt")
    if P.shape != Tb.shape or k.shape != Tb.shape:
        raise ValueError(f"shape mismatch T {Tb.shape}, P {P.shape}, k {k.shape}")
    gamma = None
    geo_path = os.path.join(export_dir, "geometry_thermal.txt")
    if os.path.exists(geo_path):
        geo = np.atleast_2d(np.loadtxt(geo_path, delimiter="\t"))
        if geo.shape == Tb.shape:
            gamma = 1.0 - np.clip(geo, 0.0, 1.0)
    out = surrogate_adjoint(k, P=P, T=Tb, gamma=gamma, k_s=k_s, k_f=k_f, q_k=q_k, device=device)
    np.savetxt(os.path.join(export_dir, "sensitivity_field.txt"), out["sens"], delimiter="\t", fmt="%.10e")
    dJ = out.get("dJ_dgamma")
    if dJ is not None:
        np.savetxt(os.path.join(export_dir, "dJ_dgamma.txt"), dJ, delimiter="\t", fmt="%.10e")
    return out["sens"], dJ


def objective(metrics, w_peak=1.0, w_uniformity=0.0):
    """evaluate_objective's weighted objective (mma_optimizer.py:475-489)
    from the parsed thermal_metrics.txt dict: w_peak (T_avg_base +
    T_range_base / 2) + w_uniformity (1 - T_uniformity_base), with the same
    fallbacks (100, 0, 1) for missing entries."""
    t_peak = metrics.get("T_avg_base", 100.0) + 0.5 * metrics.get("T_range_base", 0.0)
    return w_peak * t_peak + w_uniformity * (1.0 - metrics.get("T_uniformity_base", 1.0))


# ---------------------------------------------------------------------------
# run_solvers.py without subprocesses

def run_solvers(export_dir, skip_fluid=False, qk=1.0, mode=1, fluid_overrides=None, thermal_precision=None,
                device=0):
    """The fluid solve (SIMPLE::SIMPLE, runIterations, saveAll) and then the
    thermal solve (thermal_solver.cpp main) on one directory, as
    run_solvers.py runs simple.exe and thermal_solver.exe.  Returns
    (fluid stats or None, thermal metrics)."""
    from .simple import SIMPLE
    from .thermal import ThermalSolver3D

    fstats = None
    if not skip_fluid:
        cfg, gamma, meta = fluid_case(export_dir, **(fluid_overrides or {}))
        s = SIMPLE(cfg, gamma, device=device)
        try:
            fstats, hist = s.runIterationsLogged()
            save_iteration_logs(export_dir, hist)
            save_fluid_outputs(export_dir, s, meta)
        finally:
            s.close()
    stack, gamma, u, v = thermal_case(export_dir, qk=qk, k_mode=mode)
    kw = {} if thermal_precision is None else dict(precision=thermal_precision)
    th = ThermalSolver3D(stack, gamma, u, v, device=device, **kw)
    try:
        th.solve()
        m = save_thermal_outputs(export_dir, th, stack, gamma)
    finally:
        th.close()
    return fstats, m
