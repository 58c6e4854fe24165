"""The optimiser's objective as one differentiable function of the density
field: gamma -> SIMPLE fixed point -> cell-centred velocity -> 3-D thermal
solve -> f0.

The reference evaluates f0 = w_peak (T_avg_base + 0.5 T_range_base)
+ w_uniformity (1 - T_uniformity_base) (mma_optimizer.py:477-487) and takes
its sensitivities from the 2-D conduction surrogate, which holds the flow
fixed (compute_sensitivities, mma_optimizer.py:498-567).  Here the range,
a max minus a min, is replaced by a smooth stand-in, the difference of the
volume-weighted p-norms of T and of 1/T over the base layers,

    R_p = (mean_w T^p)^(1/p) - (mean_w T^-p)^(-1/p)   (-> max - min as p grows),

and the uniformity by 1 - std/avg as thermal_output.cpp:259-267 defines it
(without its clamp to [0, 1]):

    f0_smooth = w_peak (T_avg_base + 0.5 R_p) + w_uniformity std_base / T_avg_base.

Its gradient includes the flow: with A(u) T = b(u) the thermal system and
R(w; gamma) = 0 the converged flow equations,

    dJ/dgamma = -lambda_T^T dR_T/dgamma  +  (dJ/du_c) du_c/dgamma,
    A^T lambda_T = dJ/dT,   dJ/du_c = -lambda_T^T dR_T/du_c,

and the second term is one flow adjoint solve with dJ/dw = (du_c/dw)^T
dJ/du_c as its right side (cps_simple_gradient_vjp), u_c being the
cell-centre average of the staggered velocity (output.cpp:98-107).
"""
import numpy as np

from . import configs
from . import _types as T
from .simple import SIMPLE
from .thermal import ThermalSolver3D


def base_weights(stack):
    """Cell volumes of the base layers (below the first channel layer) as an
    (nz, ny, nx) array, zero elsewhere (the T_avg_base region,
    thermal_output.cpp:161-170)."""
    nx, ny = stack["nx"], stack["ny"]
    dz = np.asarray(stack["dz"], dtype=np.float64)
    ch = np.asarray(stack["channel"])
    first = int(np.argmax(ch)) if ch.any() else len(ch)
    w = np.zeros((len(dz), ny, nx))
    for k in range(first):
        w[k] = stack.get("dx", 2e-5) * stack.get("dy", 2e-5) * dz[k]
    return w


def f0_smooth(Tf, stack, w_peak=1.0, w_uniformity=0.0, p=8.0):
    """(f0_smooth, dJ/dT, parts) for a temperature field (nz, ny, nx)."""
    w = base_weights(stack)
    W = w.sum()
    if W <= 0:
        raise ValueError("f0 needs at least one base layer below the channel")
    Tf = np.asarray(Tf, dtype=np.float64)
    avg = (w * Tf).sum() / W
    # the p-norms are formed on T / max and T / min over the base, which keeps
    # the powers in range for large p (the scale cancels)
    tb = Tf[w > 0]
    sx, sn = tb.max(), tb.min()
    mp = (w * (Tf / sx) ** p).sum() / W
    mm = (w * (Tf / sn) ** (-p)).sum() / W
    pmax, pmin = sx * mp ** (1.0 / p), sn * mm ** (-1.0 / p)
    var = (w * (Tf - avg) ** 2).sum() / W
    std = np.sqrt(var)
    J = w_peak * (avg + 0.5 * (pmax - pmin)) + w_uniformity * std / avg
    d_avg = w / W
    d_pmax = mp ** (1.0 / p - 1.0) * w * (Tf / sx) ** (p - 1.0) / W
    d_pmin = mm ** (-1.0 / p - 1.0) * w * (Tf / sn) ** (-p - 1.0) / W
    d_std = (w * (Tf - avg) / W) / std if std > 0 else np.zeros_like(Tf)
    g = w_peak * (d_avg + 0.5 * (d_pmax - d_pmin))
    if w_uniformity:
        g = g + w_uniformity * (d_std / avg - std / avg ** 2 * d_avg)
    parts = dict(t_avg_base=avg, pnorm_max=pmax, pnorm_min=pmin, range_smooth=pmax - pmin, std_base=std)
    return float(J), g, parts


def staggered_velocity_vjp(dJ_duc, dJ_dvc, nx, ny):
    """(dJ/du, dJ/dv) in the staggered field shapes from dJ/d(cell-centred
    velocity): the transpose of configs.cell_centre_velocity."""
    du = np.zeros(T.field_shape("u", nx, ny))
    dv = np.zeros(T.field_shape("v", nx, ny))
    du[1:ny + 1, 0:nx] += 0.5 * dJ_duc
    du[1:ny + 1, 1:nx + 1] += 0.5 * dJ_duc
    dv[0:ny, 1:nx + 1] += 0.5 * dJ_dvc
    dv[1:ny + 1, 1:nx + 1] += 0.5 * dJ_dvc
    return du, dv


class CoupledObjective:
    """f0_smooth(gamma) through the device flow and thermal solvers.

    flow: (cfg, inlet) of the SIMPLE solve, run `flow_steps` iterations from
    rest (or from `flow_init`, a dict of u, v, p) to its fixed point; stack:
    the layer description of configs.thermal_stack (its nx, ny match the
    flow grid's cells)."""

    def __init__(self, flow_cfg, inlet, stack, flow_steps=1500, thermal_rtol=1e-12, w_peak=1.0, w_uniformity=0.0,
                 p=8.0, adjoint_rtol=1e-10, device=0):
        if (stack["nx"], stack["ny"]) != (flow_cfg.nx, flow_cfg.ny):
            raise ValueError("the thermal stack must have the flow grid's nx x ny cells")
        self.cfg, self.inlet, self.stack = flow_cfg.copy(), inlet, stack
        self.flow_steps, self.thermal_rtol = flow_steps, thermal_rtol
        self.w_peak, self.w_uniformity, self.p = w_peak, w_uniformity, p
        self.adjoint_rtol, self.device = adjoint_rtol, device
        self.flow = self.thermal = None

    def close(self):
        for s in (self.flow, self.thermal):
            if s is not None:
                s.close()
        self.flow = self.thermal = None

    def evaluate(self, gamma, flow_init=None):
        """f0_smooth and its parts; keeps the solvers for gradient()."""
        self.close()
        self.gamma = np.ascontiguousarray(gamma, dtype=np.float64)
        self.flow = SIMPLE(self.cfg, self.gamma, self.inlet, device=self.device)
        if flow_init is not None:
            for k, a in flow_init.items():
                self.flow.set(k, a)
        self.flow.step(self.flow_steps)
        uc, vc = configs.cell_centre_velocity(self.flow.get("u"), self.flow.get("v"))
        self.thermal = ThermalSolver3D(self.stack, self.gamma, uc, vc, rtol=self.thermal_rtol, max_iter=20000,
                                       device=self.device)
        self.thermal.solve()
        self.T = self.thermal.T
        self.J, self.dJdT, self.parts = f0_smooth(self.T, self.stack, self.w_peak, self.w_uniformity, self.p)
        return self.J

    def gradient(self):
        """(dJ/dgamma, dict of its thermal and flow parts) at the last evaluate()."""
        lam = self.thermal.solve_adjoint(self.dJdT)
        g_th, duc, dvc = self.thermal.sensitivities(lam)
        du, dv = staggered_velocity_vjp(duc, dvc, self.cfg.nx, self.cfg.ny)
        g_flow = self.flow.gradient_vjp(du, dv, None, rtol=self.adjoint_rtol, max_iter=4000)
        return g_th + g_flow, dict(thermal=g_th, flow=g_flow, dJ_du_centre=duc, dJ_dv_centre=dvc,
                                   flow_adjoint=self.flow.last_adjoint)


def _torch_function():
    import torch

    class CoupledF0(torch.autograd.Function):
        """f0_smooth(gamma) (0-d tensor); backward = thermal adjoint + flow
        adjoint (vector-Jacobian product through the velocity)."""

        @staticmethod
        def forward(ctx, gamma, obj):
            g = gamma.detach().to("cpu", torch.float64).numpy()
            J = obj.evaluate(g)
            grad, _ = obj.gradient()
            ctx.save_for_backward(torch.from_numpy(grad).to(gamma.device, gamma.dtype))
            return gamma.new_tensor(J)

        @staticmethod
        def backward(ctx, gout):
            (grad,) = ctx.saved_tensors
            return gout * grad, None

    return CoupledF0


_CF = None


def coupled_f0(gamma, obj):
    """Differentiable f0_smooth(gamma) through a CoupledObjective."""
    global _CF
    if _CF is None:
        _CF = _torch_function()
    return _CF.apply(gamma, obj)
