"""First-order VMC training on the HIP path (call surface of src/nqs/train.py:28-98).

One step: `n_sweeps` Metropolis sweeps of the persistent walkers (nqs_mcmc_sweeps), E_loc
(+ Coulomb) on them (nqs_eloc), the O rows (nqs_psiformer_grad), the surrogate-loss gradient
g = (2/W) sum_w Re[conj(e_w - <e>) O_w] (nqs_vmc_gradient: one column GEMV over O), and the
optax-style Adam update in place on the flat real parameter vector (nqs_adam_update).

Parameters are the flat real ravel theta (a torch float64 vector; `real_ravel` / `ravel_params`
give it from a parameter tree).  `Adam(complex_grads="jax")` reproduces what the reference's
optax.adam receives for a complex leaf (jax.grad of a real loss: g_re - i g_im) and
`complex_grads="descent"` steps downhill in both parts.
"""
import ctypes
from typing import NamedTuple

import numpy as np
import torch

from . import _lib
from .cell import as_cell
from .psiformer import LogPsi
from .sampler import Chain
from .local_energy import make_local_energy
from .minsr import grad_params, lib_minsr_desc, real_ravel


class AdamState(NamedTuple):
    count: int
    mu: torch.Tensor
    nu: torch.Tensor


class Adam:
    """optax.adam(lr, b1, b2, eps, eps_root) on the flat real ravel (complex leaves paired)."""

    def __init__(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, eps_root=0.0, complex_grads="jax"):
        if complex_grads not in ("jax", "descent"):
            raise ValueError("complex_grads must be 'jax' or 'descent'")
        self.lr, self.b1, self.b2, self.eps, self.eps_root = float(lr), float(b1), float(b2), float(eps), float(eps_root)
        self.mode = 0 if complex_grads == "jax" else 1
        self._partner = {}

    def init(self, theta):
        z = torch.zeros_like(torch.as_tensor(theta, dtype=torch.float64).to(_lib.device()))
        return AdamState(0, z, z.clone())

    def partner(self, model):
        """Index map pairing the real and imaginary part of every complex leaf element."""
        key = id(model)
        if key not in self._partner:
            P = model.n_params()
            pt = np.full(P, -1, np.int64)
            for _, shape, cplx, off in model.layout():
                if cplx:
                    n = int(np.prod(shape)) if shape else 1
                    pt[off:off + n] = np.arange(off + n, off + 2 * n)
                    pt[off + n:off + 2 * n] = -2
            self._partner[key] = torch.tensor(pt, device=_lib.device())
        return self._partner[key]

    def update_(self, theta, g, state, partner=None):
        """theta -= lr * Adam(g) in place; returns the new state (moments updated in place)."""
        lib = _lib.load()
        count = state.count + 1
        _lib.check(lib.nqs_adam_update(_lib.ptr(theta), _lib.ptr(g), _lib.ptr(state.mu), _lib.ptr(state.nu),
                                       _lib.ptr(partner), theta.shape[0], count, self.lr, self.b1, self.b2, self.eps,
                                       self.eps_root, self.mode, _lib.stream()), "nqs_adam_update")
        return AdamState(count, state.mu, state.nu)


def vmc_gradient(O, e):
    """g = (2/W) sum_w Re[conj(e_w - <e>) O_w] (train.py:48-58) from an OMatrix and E_loc (W,)."""
    lib = _lib.load()
    e = torch.as_tensor(e).to(_lib.device(), torch.complex128).contiguous()
    d = lib_minsr_desc(O.W, O.P, O.ldo, O.data.dtype == torch.float64, "fp64", 0, 0.0)
    nb = lib.nqs_vmc_gradient_workspace_bytes(ctypes.byref(d))
    if nb < 0:
        _lib.check(-1, "nqs_vmc_gradient_workspace_bytes")
    ws = _lib.workspace(nb, "vmc_gradient")
    g = torch.empty(O.P, dtype=torch.float64, device=e.device)
    _lib.check(lib.nqs_vmc_gradient(ctypes.byref(d), _lib.ptr(O.data), _lib.ptr(torch.view_as_real(e)), _lib.ptr(g),
                                    _lib.ptr(ws), ws.numel(), _lib.stream()), "nqs_vmc_gradient")
    return g


def make_train_step(model, L, n_elec, m_star=1.0, lambda_r=1.0, h_z=0.0, coulomb=None, optimizer=None, n_sweeps=10,
                    step_size=0.4, n_spin_flips=2):
    """Return (optimizer, step) with step(theta, opt_state, R, S, key) ->
    (theta, opt_state, R, S, E, Eerr, acc_pos, acc_spin) (train.py:28-65).  theta is updated in
    place (a device float64 vector); R, S are device walkers advanced in place."""
    if optimizer is None:
        optimizer = Adam(1e-3)
    cache = {}

    def step(theta, opt_state, R, S, key):
        theta = theta if (torch.is_tensor(theta) and theta.device.type == "cuda" and theta.dtype == torch.float64) \
            else torch.as_tensor(theta, dtype=torch.float64).to(_lib.device()).contiguous()
        lp = LogPsi(model, theta)
        ch = Chain(lp, R, S)
        ap, asp = ch.run(n_sweeps, step_size, n_spin_flips, key=key)
        e = make_local_energy(lp, m_star, lambda_r, h_z)(ch.R, ch.S)
        if coulomb is not None:
            e = e + coulomb(ch.R)
        O, _ = grad_params(lp, ch.R, ch.S, cache.get("O"))
        cache["O"] = O
        g = vmc_gradient(O, e)
        opt_state = optimizer.update_(theta, g, opt_state, optimizer.partner(model))
        Ebar = e.mean()
        Eerr = e.real.std(unbiased=False) / np.sqrt(e.shape[0])
        return theta, opt_state, ch.R, ch.S, Ebar.real, Eerr, ap.mean(), asp.mean()

    return optimizer, step


def init_walkers(key, n_walkers, n_elec, L):
    """Uniform positions in the cell and random spins (train.py:68-72), seeded torch generator."""
    dev = _lib.device()
    g = torch.Generator(device=dev)
    g.manual_seed(int(key))
    A = torch.tensor(as_cell(L).T, device=dev)
    R = torch.rand(n_walkers, n_elec, 2, generator=g, dtype=torch.float64, device=dev) @ A
    S = torch.randint(0, 2, (n_walkers, n_elec), generator=g, device=dev).to(torch.int32)
    return R, S


def train(model, params, L, n_elec, key, *, m_star=1.0, lambda_r=1.0, h_z=0.0, coulomb=None, n_walkers=256,
          n_steps=200, lr=1e-3, n_sweeps=10, step_size=0.4, burn=50, optimizer=None, log_every=20, history=None):
    """VMC training with Adam (train.py:75-98); returns (params, history).  params: a parameter
    tree (returned as a tree) or a flat theta (returned flat)."""
    flat = torch.is_tensor(params) and params.dim() == 1
    if flat:
        theta, unravel = params.to(torch.float64), None
    else:
        theta, unravel = real_ravel(params)
    theta = theta.to(_lib.device()).contiguous().clone()
    optimizer, step = make_train_step(model, L, n_elec, m_star, lambda_r, h_z, coulomb, optimizer or Adam(lr),
                                      n_sweeps, step_size)
    opt_state = optimizer.init(theta)
    R, S = init_walkers(key, n_walkers, n_elec, L)
    # burn-in at the initial parameters: the reference runs (burn // n_sweeps + 1) steps and keeps
    # only the walkers, i.e. that many blocks of n_sweeps sweeps
    ch = Chain(LogPsi(model, theta), R, S)
    for i in range(burn // n_sweeps + 1):
        ch.run(n_sweeps, step_size, 2, key=int(key) * 7919 + i)
    R, S = ch.R, ch.S
    hist = history if history is not None else {"E": [], "Eerr": [], "acc_pos": []}
    for it in range(n_steps):
        theta, opt_state, R, S, E, Eerr, ap, asp = step(theta, opt_state, R, S, int(key) * 7919 + 1000 + it)
        hist["E"].append(float(E))
        hist["Eerr"].append(float(Eerr))
        hist["acc_pos"].append(float(ap))
        if log_every and (it % log_every == 0 or it == n_steps - 1):
            print(f"    step {it:4d}: E={float(E):+.6f} +/- {float(Eerr):.1e}  acc_pos={float(ap):.2f} "
                  f"acc_spin={float(asp):.2f}", flush=True)
    return (theta if flat else unravel(theta.cpu())), hist
