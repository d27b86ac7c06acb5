"""torch.autograd.Function wrappers of the HIP adjoint (the backward passes run the library's
vector-Jacobian-product kernels, not torch's autograd of anything).

States are float64 CUDA tensors viewing the complex128 state [C ; phi(:, 1..M)] as (re, im)
pairs, the layout of the library's device entry points.  torch's gradient of such a tensor is
the Euclidean one of the real view, while the library works in the metric
<a, b> = Re(sum_C conj(a) b + sum_orbitals w conj(a) b); the wrappers convert between the two
(a cotangent's orbital part is divided by the grid weights on the way in, a gradient's orbital
part multiplied by them on the way out).
"""
from __future__ import annotations

import numpy as np
import torch

from . import hip


def _metric(ctx: hip.Context, device) -> torch.Tensor:
    """Per-real-entry weights of the metric: 1 on C, w(x) on the orbitals (both of re, im)."""
    w = np.concatenate([np.ones(ctx.nC), np.tile(ctx.w, ctx.nO)])
    return torch.from_numpy(np.repeat(w, 2)).to(device)


_SIDE = {}


def _run_ordered(device, fn):
    """Runs fn(stream handle) on a dedicated (default-priority) torch stream ordered after the
    current stream's work, and orders the current stream after it.  The library's own stream is
    non-blocking, so it would not order with torch's legacy default stream by itself."""
    side = _SIDE.get(device)
    if side is None:
        side = _SIDE[device] = torch.cuda.Stream(device=device)
    cur = torch.cuda.current_stream(device)
    side.wait_stream(cur)
    fn(side.cuda_stream)
    cur.wait_stream(side)


def laser_basis(n: int, T: float, t: float) -> np.ndarray:
    """dE(t)/dc_k = sin(k pi t / T) sin^2(pi t / T), k = 1..n (zero outside [0, T])."""
    if t < 0.0 or t > T:
        return np.zeros(n)
    env = np.sin(np.pi * t / T) ** 2
    return np.sin(np.arange(1, n + 1) * np.pi * t / T) * env


class RightHandSide(torch.autograd.Function):
    """dstate = f(state, t) for one part of the right-hand side (PART_FULL / PART_LIN /
    PART_NONLIN) under the laser field of `coeffs` (period T); the backward pass is
    cody_time_derivative_vjp_dev, whose <cotangent, df/dE> gives the coefficients' gradient."""

    @staticmethod
    def forward(actx, state: torch.Tensor, coeffs: torch.Tensor, ctx: hip.Context, t: float, part: int, T: float):
        if not state.is_cuda or state.dtype != torch.float64 or not state.is_contiguous():
            raise ValueError("state must be a contiguous float64 CUDA tensor")
        ctx.set_laser(coeffs.detach().cpu().numpy().astype(np.float64), T)
        out = torch.empty_like(state)
        mode = {hip.PART_FULL: "full", hip.PART_LIN: "lin", hip.PART_NONLIN: "nonlin"}[part]
        _run_ordered(state.device,
                     lambda st: ctx.time_derivative_dev(state.data_ptr(), out.data_ptr(), float(t), st, mode=mode))
        actx.save_for_backward(state, coeffs)
        actx.hctx, actx.t, actx.part, actx.T = ctx, float(t), part, float(T)
        return out

    @staticmethod
    def backward(actx, gout: torch.Tensor):
        state, coeffs = actx.saved_tensors
        ctx = actx.hctx
        ctx.set_laser(coeffs.detach().cpu().numpy().astype(np.float64), actx.T)
        m = _metric(ctx, state.device)
        lam = (gout.contiguous() / m).contiguous()
        g = torch.empty_like(state)
        eb = torch.zeros(1, dtype=torch.float64, device=state.device)
        _run_ordered(state.device, lambda st: ctx.vjp_dev(actx.part, state.data_ptr(), lam.data_ptr(), g.data_ptr(),
                                                          actx.t, eb.data_ptr(), st))
        basis = torch.from_numpy(laser_basis(coeffs.numel(), actx.T, actx.t)).to(coeffs.device)
        return g * m, eb.to(coeffs.device) * basis, None, None, None, None


class DipoleAfterPropagation(torch.autograd.Function):
    """J(c) = <x>(t0 + n dt) after n Chebyshev-Strang steps from `state0` under the laser field
    E(t) = sum_k c_k sin(k pi t / T) sin^2(pi t / T); the backward pass is the library's
    checkpointed discrete adjoint (cody_dipole_gradient)."""

    @staticmethod
    def forward(actx, coeffs: torch.Tensor, ctx: hip.Context, state0: np.ndarray, T: float, t0: float, dt: float,
                n_steps: int, checkpoint_every: int = 50):
        c = coeffs.detach().cpu().numpy().astype(np.float64)
        ctx.set_laser(c, T)
        ctx.upload_state(state0)
        J, g = ctx.dipole_gradient(t0, dt, n_steps, checkpoint_every)
        actx.save_for_backward(torch.from_numpy(g).to(coeffs.device))
        return torch.tensor(J, dtype=torch.float64, device=coeffs.device)

    @staticmethod
    def backward(actx, gJ: torch.Tensor):
        (g,) = actx.saved_tensors
        return gJ * g, None, None, None, None, None, None, None
