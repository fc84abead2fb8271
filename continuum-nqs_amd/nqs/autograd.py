"""log psi as a differentiable torch function over the hand-written HIP kernels.

    y = psiformer_logpsi(model, theta, R, S)     # (W, 2) float64: [Re log psi, Im log psi]
    y.backward(g)                                # theta.grad, R.grad

The reference differentiates its flax model with jax (`jax.grad` per walker in `per_sample_O`,
src/nqs/minsr.py:56-65, and `jax.grad` / `jax.hessian` in the local energy,
src/nqs/local_energy.py:52-81).  Here the forward is nqs_psiformer_fwd and the backward is the
library's own:

  d/dtheta : nqs_psiformer_vjp, a reverse sweep that recomputes each transformer block's
             activations from its input instead of keeping them for every walker, with the
             cotangent (g_re, g_im) folded in at the determinant head and the weight gradients
             summed over walkers and tokens on the device: sum_w g_re d Re L_w + g_im d Im L_w,
             without forming the per-walker rows O.
  d/dR     : the forward-mode tangent channels of nqs_eloc (called without Rashba term or fields,
             so no spin-flipped configuration is evaluated) give grad log psi per coordinate; the
             product is g_re Re grad + g_im Im grad.

S (integer spins) is not differentiable.  theta must be the flat real_ravel vector (float64).
"""
import ctypes

import torch

from . import _lib
from .psiformer import LogPsi, as_walkers


def vjp_theta(logpsi, R, S, cot, ws_tag="vjp"):
    """sum_w [cot_w.re d Re log psi_w / d theta + cot_w.im d Im log psi_w / d theta] (P,) float64,
    by nqs_psiformer_vjp; cot: (W,) complex or (W, 2) real."""
    lib = _lib.load()
    Rw, Sw, _ = as_walkers(R, S)
    W = Rw.shape[0]
    cot = torch.as_tensor(cot, device=Rw.device)
    if cot.is_complex():
        cot = torch.view_as_real(cot.to(torch.complex128))
    cot = cot.reshape(W, 2).to(torch.float64).contiguous()
    m = logpsi.model
    d = m.desc()
    g = torch.empty(m.n_params(), dtype=torch.float64, device=Rw.device)
    nb = lib.nqs_workspace_bytes(ctypes.byref(d), _lib.OP_VJP, W)
    ws = _lib.workspace(nb, ws_tag)
    _lib.check(lib.nqs_psiformer_vjp(ctypes.byref(d), _lib.ptr(logpsi.packed), _lib.ptr(Rw), _lib.ptr(Sw), W,
                                     _lib.ptr(cot), _lib.ptr(g), None, _lib.ptr(ws), ws.numel(), _lib.stream()),
               "nqs_psiformer_vjp")
    return g


class PsiformerLogPsi(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, R, S, model):
        lp = LogPsi(model, theta.detach())
        Rw, Sw, _ = as_walkers(R.detach(), S)
        y = torch.view_as_real(lp(Rw, Sw)).clone()
        ctx.model = model
        ctx.lp = lp
        ctx.save_for_backward(Rw, Sw)
        ctx.r_shape = R.shape
        return y

    @staticmethod
    def backward(ctx, g):
        from .local_energy import LocalEnergy
        Rw, Sw = ctx.saved_tensors
        W = Rw.shape[0]
        g = g.reshape(W, 2).to(torch.float64)
        d_theta = d_R = None
        if ctx.needs_input_grad[0]:
            d_theta = vjp_theta(ctx.lp, Rw, Sw, g)
        if ctx.needs_input_grad[1]:
            # lambda_R = 0 and no fields: the library runs only the forward-mode pass (value, 2N
            # tangents, Laplacian) and skips the spin-flipped configurations
            _, _, grad, _ = LocalEnergy(ctx.lp, lambda_r=0.0).parts(Rw, Sw)    # (W, 2N) complex
            d = g[:, :1] * grad.real + g[:, 1:] * grad.imag
            d_R = d.reshape(ctx.r_shape)
        return d_theta, d_R, None, None


def psiformer_logpsi(model, theta, R, S):
    """[Re, Im] log psi of the walkers (W, 2), differentiable in theta and R."""
    if not torch.is_tensor(R):
        R = torch.as_tensor(R, dtype=torch.float64, device=_lib.device())
    return PsiformerLogPsi.apply(theta, R, S, model)
