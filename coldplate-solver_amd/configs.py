"""Seeded synthetic inputs for the benchmark workloads.

Written from the workload descriptions (not from the reference's interactive
generators, GeometryGenerator_V6.py / PowerMapGenerator_V2.py, which prompt
for input and use unseeded RNG):

  channel_200x50      straight all-fluid channel, uniform inlet, FOU
  density_field(...)  smoothed Gaussian random field -> tanh projection,
                      Brinkman RAMP sink, 2.5-D sink, SOU, parabolic inlet
  thermal_stack(...)  copper base / gamma-interpolated channel / copper lid
                      with a seeded hotspot power map
"""
import hashlib

import numpy as np

from ._types import default_config, CPS_PRESSURE_PCG


def sha256(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def channel_200x50(**overrides):
    """Plain 2-D SIMPLE, straight channel 200x50 cells, dx = dy = 1e-4 m,
    all fluid, uniform inlet 0.1 m/s, FOU, Jacobi momentum (tol 1e-6, 400
    sweeps), Jacobi-PCG pressure (pTol 1e-8, cap raised to 1000), uvAlpha
    0.7, pAlpha 0.3, 2.5-D off, 2000 iterations or residual 1e-6 (the
    dP-window stop is off)."""
    cfg = default_config(
        nx=200, ny=50, hx=1e-4, hy=1e-4, u_in=0.1, convection_scheme=0,
        p_tol=1e-8, max_pressure_iter=1000, pressure_solver=CPS_PRESSURE_PCG,
        epsilon=1e-6, max_iterations=2000, use_dp_convergence=0,
    )
    for k, v in overrides.items():
        setattr(cfg, k, v)
    gamma = np.ones((cfg.ny, cfg.nx))
    return cfg, gamma, None


def projected_random_field(ny, nx, seed, sigma=6.0, beta=8.0, eta=0.5):
    """gamma in [0,1]: N(0,1) noise -> gaussian_filter(sigma, reflect) ->
    min-max to [0,1] -> tanh projection
    (tanh(b e) + tanh(b (x - e))) / (tanh(b e) + tanh(b (1 - e)))."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    w = rng.standard_normal((ny, nx))
    w = gaussian_filter(w, sigma, mode="reflect")
    x = (w - w.min()) / (w.max() - w.min())
    num = np.tanh(beta * eta) + np.tanh(beta * (x - eta))
    den = np.tanh(beta * eta) + np.tanh(beta * (1.0 - eta))
    return num / den


def parabolic_inlet(ny, u_max, r0=None, r1=None):
    """Per-u-row inlet speeds (M+1 = ny+2 entries).  Physical rows r0..r1
    carry u_max 4 xi (1 - xi), xi = (r + 0.5 - r0) / (r1 - r0 + 1); u-row
    i = r + 1.  Defaults are rows ny/8 .. 7ny/8 (64..448 at ny = 512)."""
    r0 = ny // 8 if r0 is None else r0
    r1 = 7 * ny // 8 if r1 is None else r1
    prof = np.zeros(ny + 2)
    r = np.arange(r0, r1 + 1)
    xi = (r + 0.5 - r0) / (r1 - r0 + 1)
    prof[r + 1] = u_max * 4.0 * xi * (1.0 - xi)
    return prof


def density_field(nx=1024, ny=512, seed=0, sigma=6.0, h=2e-5, **overrides):
    """2-D density-field solve: dx = dy = 2e-5 m, projected random gamma,
    Brinkman alpha_max q (1-g)/(q+g) with q = 0.1, alpha_max = 1e8 Pa s/m^2
    (the reference RAMP alpha_max (1-g)/(1+q'g) with q' = 1/q = 10 and
    K_min = mu/alpha_max), 2.5-D on with Ht = 1e-4 m (sink 12 mu/Ht^2,
    reference convection factor 1.2), SOU, parabolic inlet 0.5 m/s on rows
    ny/8..7ny/8, pressure outlet, a fixed iteration count (dP-window stop
    off).  The boundary-patch dP averages rows whose adjacent cell has
    gamma >= 0.5 (the reference's 1 - 1e-6 test selects almost no rows of a
    projected random field)."""
    mu = 0.00089
    cfg = default_config(
        nx=nx, ny=ny, hx=h, hy=h, u_in=0.5, convection_scheme=1,
        enable_25d=1, ht=1e-4, conv_factor_25d=1.2, sink_mult_25d=1.0,
        brinkman_q=10.0, brinkman_kmin=mu / 1e8,
        max_iterations=1000, dp_fluid_tol=0.5, use_dp_convergence=0,
    )
    for k, v in overrides.items():
        setattr(cfg, k, v)
    gamma = projected_random_field(ny, nx, seed, sigma=sigma)
    inlet = parabolic_inlet(ny, cfg.u_in)
    return cfg, gamma, inlet


BATCH_SIZES = [(256, 128), (512, 256), (768, 384), (1024, 512)]


def batch_sweep(n_fields=48, seed=0):
    """48 density fields, sizes uniform over BATCH_SIZES, sigma ~ U[4, 12],
    one seed per field."""
    rng = np.random.default_rng(seed)
    out = []
    for f in range(n_fields):
        nx, ny = BATCH_SIZES[rng.integers(len(BATCH_SIZES))]
        sigma = rng.uniform(4.0, 12.0)
        out.append(density_field(nx=nx, ny=ny, seed=1000 + f, sigma=sigma, max_iterations=500))
    return out


def power_map(nx, ny, seed=0, background=20.0, n_spots=8, sigma=20.0, amp=(50.0, 300.0)):
    """Bottom heat flux in W/m^2: 20 W/cm^2 + 8 Gaussian hotspots (sigma 20
    cells, amplitudes U[50, 300] W/cm^2, seeded centres), x 1e4."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:ny, 0:nx]
    q = np.full((ny, nx), background)
    for _ in range(n_spots):
        cx, cy = rng.uniform(0, nx), rng.uniform(0, ny)
        a = rng.uniform(*amp)
        q += a * np.exp(-((x + 0.5 - cx) ** 2 + (y + 0.5 - cy) ** 2) / (2 * sigma ** 2))
    return q * 1e4


def thermal_stack(nx=512, ny=256, n_base=4, n_channel=4, n_lid=4, dz_solid=50e-6, ht=1e-4, h=2e-5,
                  k_s=390.0, k_f=0.6, rho=997.0, cp=4180.0, t_in=30.0, qk=1.0, seed=0):
    """Layer description of the 3-D conjugate-heat-transfer stack: copper
    base, channel (dz = Ht / n_channel, k RAMP-interpolated by gamma),
    copper lid."""
    dz = [dz_solid] * n_base + [ht / n_channel] * n_channel + [dz_solid] * n_lid
    channel = [0] * n_base + [1] * n_channel + [0] * n_lid
    return dict(nx=nx, ny=ny, dx=h, dy=h, dz=np.array(dz), channel=np.array(channel, dtype=np.int32), ht=ht,
                k_s=k_s, k_f=k_f, rho=rho, cp=cp, t_in=t_in, qk=qk, q0=power_map(nx, ny, seed=seed))


def cell_centre_velocity(u, v):
    """Cell-centred velocities of the physical cells from staggered u, v
    (output.cpp:98-107): uc(i,j) = (u(i+1,j) + u(i+1,j+1))/2,
    vc(i,j) = (v(i,j+1) + v(i+1,j+1))/2."""
    ny, nx = u.shape[0] - 2, u.shape[1] - 1
    uc = 0.5 * (u[1:ny + 1, 0:nx] + u[1:ny + 1, 1:nx + 1])
    vc = 0.5 * (v[0:ny, 1:nx + 1] + v[1:ny + 1, 1:nx + 1])
    return uc, vc
