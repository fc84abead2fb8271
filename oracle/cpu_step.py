"""One VMC step of the oracle on the CPU (test infrastructure; the CPU baseline of bench.py).

The reference's step (src/nqs/minsr.py:139-152) restated with the oracle's pieces: n_sweeps
Metropolis sweeps (sampler.py:36-57, log psi recomputed for current and proposed states as the
reference does), dense-Hessian E_loc (local_energy.py:58-115) + Ewald Coulomb, O by reverse-mode
autodiff (minsr.py:56-65), median/MAD clip, MinSR with an LU solve (minsr.py:76-81), update.
"""
import time

import numpy as np
import torch
from torch.func import vmap

from . import psiformer as opf, local_energy as ole, minsr as oms, sampler as osa


def vmc_step(arch, theta, R, S, *, lambda_r, h_z, coulomb, n_sweeps=10, step=0.9, lam=1e-4, lr=0.05, seed=0,
             timings=None):
    p = opf.unravel(arch, theta)
    single = lambda r, s: opf.logpsi_from_params(arch, p, r, s)
    batched = vmap(single)
    logpsi_b = lambda R_, S_: batched(torch.as_tensor(R_), torch.as_tensor(S_)).numpy()
    W, N = R.shape[0], R.shape[1]
    st = osa.make_streams(n_sweeps, W, N, 2, seed)
    t = [time.perf_counter()]
    for s in range(n_sweeps):
        R, S, _, _, _ = osa.sweep(logpsi_b, R, S, arch.A, step, st["xi"][s], st["u_pos"][s], st["flip_i"][s],
                                  st["u_flip"][s])
    t.append(time.perf_counter())
    E = ole.make_local_energy(single, 1.0, lambda_r, h_z)
    e = np.array([E(torch.tensor(R[w]), torch.tensor(S[w])) for w in range(W)])
    if coulomb is not None:
        e = e + np.array([coulomb(R[w]) for w in range(W)])
    t.append(time.perf_counter())
    O = oms.per_sample_O(arch, theta, R, S)
    t.append(time.perf_counter())
    theta2, Em, Eerr, d = oms.step_update(theta, O, e, lam, lr)
    t.append(time.perf_counter())
    if timings is not None:
        timings.update(sample=t[1] - t[0], eloc=t[2] - t[1], grad=t[3] - t[2], minsr=t[4] - t[3])
    return theta2, R, S, Em
