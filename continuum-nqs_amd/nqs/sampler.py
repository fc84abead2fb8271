"""Metropolis MCMC over (positions, spins) on the HIP path (call surface of src/nqs/sampler.py:19-98).

A sweep = one joint all-electron Gaussian move, wrapped into the cell, then n_spin_flips
single-spin flips, each accepted on |Psi|^2 = exp(2 Re log Psi).  nqs_mcmc_sweeps runs all
sweeps for all walkers on the GPU with the current log|Psi|^2 cached per walker (one forward
per proposal).  Random streams: `streams=` passes explicit (xi, u_pos, flip_i, u_flip) arrays
(see make_streams) so that decisions can be compared with any other implementation; otherwise
they are drawn on the GPU from a torch generator seeded by `key`.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .cell import as_cell
from .psiformer import LogPsi


def make_streams(n_sweeps, W, N, n_flip, key=0, device=None):
    """Device streams: xi (n,W,N,2) normal, u_pos (n,W), flip_i (n,n_flip,W) int32, u_flip (n,n_flip,W)."""
    dev = device or _lib.device()
    g = torch.Generator(device=dev)
    g.manual_seed(int(key))
    return dict(xi=torch.randn(n_sweeps, W, N, 2, generator=g, dtype=torch.float64, device=dev),
                u_pos=torch.rand(n_sweeps, W, generator=g, dtype=torch.float64, device=dev),
                flip_i=torch.randint(0, N, (n_sweeps, n_flip, W), generator=g, device=dev).to(torch.int32),
                u_flip=torch.rand(n_sweeps, n_flip, W, generator=g, dtype=torch.float64, device=dev))


def _dev_streams(st, dev):
    f = lambda a, dt: torch.as_tensor(np.asarray(a) if not torch.is_tensor(a) else a).to(device=dev, dtype=dt).contiguous()
    return (f(st["xi"], torch.float64), f(st["u_pos"], torch.float64), f(st["flip_i"], torch.int32),
            f(st["u_flip"], torch.float64))


class Chain:
    """Persistent walker state (R, S, log|Psi|^2) advanced by nqs_mcmc_sweeps."""

    def __init__(self, logpsi, R, S):
        if not isinstance(logpsi, LogPsi):
            raise TypeError("the sampler needs a logpsi from nqs.psiformer.make_logpsi / slater_reference")
        dev = _lib.device()
        self.logpsi = logpsi
        self.R = torch.as_tensor(np.asarray(R) if not torch.is_tensor(R) else R).to(device=dev, dtype=torch.float64).contiguous()
        self.S = torch.as_tensor(np.asarray(S) if not torch.is_tensor(S) else S).to(device=dev, dtype=torch.int32).contiguous()
        self.logp = (2.0 * logpsi(self.R, self.S).real).contiguous()

    def run(self, n_sweeps, step, n_spin_flips=2, key=0, streams=None):
        lib = _lib.load()
        W, N = self.R.shape[0], self.R.shape[1]
        dev = self.R.device
        st = streams if streams is not None else make_streams(n_sweeps, W, N, n_spin_flips, key, dev)
        xi, up, fi, uf = _dev_streams(st, dev)
        assert xi.shape == (n_sweeps, W, N, 2) and up.shape == (n_sweeps, W)
        assert fi.shape == (n_sweeps, n_spin_flips, W) and uf.shape == (n_sweeps, n_spin_flips, W)
        acc_p = torch.zeros(W, dtype=torch.int32, device=dev)
        acc_s = torch.zeros(W, dtype=torch.int32, device=dev)
        m = self.logpsi.model
        d = m.desc()
        nb = lib.nqs_workspace_bytes(ctypes.byref(d), _lib.OP_MCMC, W)
        ws = _lib.workspace(nb)
        _lib.check(lib.nqs_mcmc_sweeps(ctypes.byref(d), _lib.ptr(self.logpsi.packed), _lib.ptr(self.R),
                                       _lib.ptr(self.S), _lib.ptr(self.logp), W, int(n_sweeps), int(n_spin_flips),
                                       float(step), _lib.ptr(xi), _lib.ptr(up), _lib.ptr(fi), _lib.ptr(uf),
                                       _lib.ptr(acc_p), _lib.ptr(acc_s), _lib.ptr(ws), ws.numel(), _lib.stream()),
                   "nqs_mcmc_sweeps")
        ap = acc_p.double() / max(n_sweeps, 1)
        asp = acc_s.double() / max(n_sweeps * n_spin_flips, 1)
        return ap, asp


def make_mcmc(logpsi, cell, n_elec, n_spin_flips=2):
    """sweep(key, R, S, step) -> (R', S', acc_pos, acc_spin) for a batch of walkers (W, N, 2)."""
    def sweep(key, R, S, step, streams=None):
        ch = Chain(logpsi, R, S)
        ap, asp = ch.run(1, step, n_spin_flips, key=key, streams=streams)
        return ch.R, ch.S, ap, asp
    return sweep


def run_chain(logpsi, cell, n_elec, key, n_walkers=256, n_sweeps=200, burn=100, step=0.4, n_spin_flips=2,
              R0=None, S0=None):
    """Batched MCMC; returns (R, S, accept_pos, accept_spin) after burn-in (sampler.py:62-91)."""
    dev = _lib.device()
    A = as_cell(cell)
    g = torch.Generator(device=dev)
    g.manual_seed(int(key))
    if R0 is None:
        R = torch.rand(n_walkers, n_elec, 2, generator=g, dtype=torch.float64, device=dev) @ torch.tensor(A.T, device=dev)
    else:
        R = torch.as_tensor(np.asarray(R0), dtype=torch.float64, device=dev).expand(n_walkers, n_elec, 2).contiguous()
    if S0 is None:
        S = torch.randint(0, 2, (n_walkers, n_elec), generator=g, device=dev).to(torch.int32)
    else:
        S = torch.as_tensor(np.asarray(S0), device=dev).to(torch.int32).expand(n_walkers, n_elec).contiguous()
    ch = Chain(logpsi, R, S)
    seed = int(torch.randint(0, 2 ** 31 - 1, (1,), generator=g, device=dev).item())
    block = 50
    for b0 in range(0, burn, block):
        ch.run(min(block, burn - b0), step, n_spin_flips, key=seed + b0)
    aps, asps, tot = 0.0, 0.0, 0
    for b0 in range(0, n_sweeps, block):
        nb = min(block, n_sweeps - b0)
        ap, asp = ch.run(nb, step, n_spin_flips, key=seed + 10 ** 6 + b0)
        aps += float(ap.mean()) * nb
        asps += float(asp.mean()) * nb
        tot += nb
    return ch.R, ch.S, aps / max(tot, 1), asps / max(tot, 1)


def vmc_energy(E_loc, R, S):
    """Mean local energy + standard error over a walker ensemble (sampler.py:94-98)."""
    e = E_loc(R, S)
    er = e.real
    return float(er.mean()), float(er.std(unbiased=False) / np.sqrt(er.shape[0])), e
