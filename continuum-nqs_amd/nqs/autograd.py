"""log psi as a differentiable torch function over the hand-written HIP kernels.

    y = psiformer_logpsi(model, theta, R, S)     # (W, 2) float64: [Re log psi, Im log psi]
    y.backward(g)                                # theta.grad, R.grad

The reference differentiates its flax model with jax (`jax.grad` per walker in `per_sample_O`,
src/nqs/minsr.py:56-65, and `jax.grad` / `jax.hessian` in the local energy,
src/nqs/local_energy.py:52-81).  Here the forward is nqs_psiformer_fwd and the backward is the
library's own:

  d/dtheta : nqs_psiformer_grad writes the per-walker rows O = [d Re log psi / d theta;
             d Im log psi / d theta] (a reverse sweep that recomputes each transformer block's
             activations from its input instead of keeping them for every walker); the vector-
             Jacobian product is g_re^T O_re + g_im^T O_im.
  d/dR     : the forward-mode tangent channels of nqs_eloc (called without Rashba term or fields,
             so no spin-flipped configuration is evaluated) give grad log psi per coordinate; the
             product is g_re Re grad + g_im Im grad.

S (integer spins) is not differentiable.  theta must be the flat real_ravel vector (float64).
"""
import torch

from . import _lib
from .psiformer import LogPsi, as_walkers


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
        from .minsr import grad_params
        from .local_energy import LocalEnergy
        Rw, Sw = ctx.saved_tensors
        W = Rw.shape[0]
        g = g.reshape(W, 2).to(torch.float64)
        d_theta = d_R = None
        if ctx.needs_input_grad[0]:
            O, _ = grad_params(ctx.lp, Rw, Sw)
            P = O.P
            coef = torch.cat([g[:, 0], g[:, 1]]).to(O.data.dtype)
            d_theta = (coef @ O.data)[:P].to(torch.float64)
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
