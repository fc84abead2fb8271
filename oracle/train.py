"""Oracle first-order training step (test infrastructure): src/nqs/train.py:28-65 restated.

  surrogate_gradient  d/dtheta of 2 mean_w Re[stopgrad(conj(e_w - <e>)) log psi_theta(w)] by torch
                      reverse mode over the oracle's log psi, as the real gradient per real-ravel
                      coordinate (train.py:48-58)
  adam_step           optax.adam (scale_by_adam then scale(-lr), optax 0.2.x) over the parameter
                      leaves: a complex leaf's elements are complex for Adam (complex first moment,
                      second moment |g|^2); the complex gradient is g_re - i g_im ("jax": what
                      jax.grad of a real loss returns for a complex input, per the JAX autodiff
                      convention) or g_re + i g_im ("descent")
Pins (tests/test_oracle.py): the surrogate gradient equals (2/W) sum_w Re[conj(e_w - <e>) O_w]
from the oracle's per-sample O and central differences of the surrogate; Adam on real leaves
equals torch.optim.Adam; a constant gradient gives the closed-form step -lr g/(|g| + eps) at every
count; a complex element with g_im = 0 moves like a real one.
"""
import numpy as np
import torch

from . import psiformer as opf


def surrogate(arch, theta, R, S, e):
    ec = np.conj(np.asarray(e, complex) - np.mean(e))
    W = len(R)
    tot = 0.0
    for w in range(W):
        lp = opf.logpsi(arch, theta, torch.as_tensor(R[w]), torch.as_tensor(S[w]))
        tot = tot + (complex(ec[w]) * lp).real
    return 2.0 * tot / W


def surrogate_gradient(arch, theta, R, S, e):
    th = torch.as_tensor(theta, dtype=torch.float64).clone().requires_grad_(True)
    s = surrogate(arch, th, R, S, e)
    s.backward()
    return th.grad.detach()


def adam_step(arch, theta, g, state, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, eps_root=0.0, complex_grads="jax"):
    """One optax.adam step; state = (count, mu list, nu list) per leaf (None before the first).
    Returns (new theta, new state)."""
    layout, P = opf.ravel_layout(arch)
    theta = np.asarray(theta, float).copy()
    g = np.asarray(g, float)
    count, mus, nus = state if state is not None else (0, None, None)
    count += 1
    new_mu, new_nu = [], []
    for k, (_, shape, cplx, off) in enumerate(layout):
        n = int(np.prod(shape)) if len(shape) else 1
        if cplx:
            gl = g[off:off + n] + (-1j if complex_grads == "jax" else 1j) * g[off + n:off + 2 * n]
            p = theta[off:off + n] + 1j * theta[off + n:off + 2 * n]
            mu0 = np.zeros(n, complex) if mus is None else mus[k]
        else:
            gl = g[off:off + n]
            p = theta[off:off + n]
            mu0 = np.zeros(n) if mus is None else mus[k]
        nu0 = np.zeros(n) if nus is None else nus[k]
        mu = (1 - b1) * gl + b1 * mu0
        nu = (1 - b2) * (np.conj(gl) * gl).real + b2 * nu0
        mhat = mu / (1 - b1 ** count)
        vhat = nu / (1 - b2 ** count)
        u = mhat / (np.sqrt(vhat + eps_root) + eps)
        p = p + (-lr) * u
        if cplx:
            theta[off:off + n], theta[off + n:off + 2 * n] = p.real, p.imag
        else:
            theta[off:off + n] = p
        new_mu.append(mu)
        new_nu.append(nu)
    return theta, (count, new_mu, new_nu)
