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


def vmc_step_detail(arch, theta, R, S, *, lambda_r, h_z, coulomb, streams, step, lam=1e-4, lr=0.05, clip_k=5.0,
                    m_star=1.0, timings=None):
    """One step from (theta, R, S) with caller-supplied Metropolis streams (sampler.make_streams
    layout, n_sweeps = streams["xi"].shape[0]).  Returns a dict: theta (updated), R, S, acc_pos and
    acc_spin (accepted moves per walker), e (E_loc + Coulomb, complex (W,)), E (unclipped mean),
    Eerr, dtheta."""
    p = opf.unravel(arch, theta)
    single = lambda r, s: opf.logpsi_from_params(arch, p, r, s)
    batched = vmap(single)
    logpsi_b = lambda R_, S_: batched(torch.as_tensor(R_), torch.as_tensor(S_)).numpy()
    W = R.shape[0]
    n_sweeps = streams["xi"].shape[0]
    acc_pos = np.zeros(W, np.int64)
    acc_spin = np.zeros(W, np.int64)
    t = [time.perf_counter()]
    for s in range(n_sweeps):
        R, S, a, asp, _ = osa.sweep(logpsi_b, R, S, arch.A, step, streams["xi"][s], streams["u_pos"][s],
                                    streams["flip_i"][s], streams["u_flip"][s])
        acc_pos += a
        acc_spin += asp.sum(0)
    t.append(time.perf_counter())
    E = ole.make_local_energy(single, m_star, lambda_r, h_z)
    e = np.array([E(torch.tensor(R[w]), torch.tensor(S[w])) for w in range(W)])
    if coulomb is not None:
        e = e + np.array([coulomb(R[w]) for w in range(W)])
    t.append(time.perf_counter())
    O = oms.per_sample_O(arch, theta, R, S)
    t.append(time.perf_counter())
    theta2, Em, Eerr, d = oms.step_update(theta, O, e, lam, lr, clip_k)
    t.append(time.perf_counter())
    if timings is not None:
        timings.update(sample=t[1] - t[0], eloc=t[2] - t[1], grad=t[3] - t[2], minsr=t[4] - t[3])
    return dict(theta=theta2, R=R, S=S, acc_pos=acc_pos, acc_spin=acc_spin, e=e, E=Em, Eerr=Eerr, dtheta=d)


def vmc_step(arch, theta, R, S, *, lambda_r, h_z, coulomb, n_sweeps=10, step=0.9, lam=1e-4, lr=0.05, seed=0,
             timings=None):
    """vmc_step_detail with seeded streams; returns (theta', R, S, E)."""
    W, N = R.shape[0], R.shape[1]
    st = osa.make_streams(n_sweeps, W, N, 2, seed)
    out = vmc_step_detail(arch, theta, R, S, lambda_r=lambda_r, h_z=h_z, coulomb=coulomb, streams=st, step=step,
                          lam=lam, lr=lr, timings=timings)
    return out["theta"], out["R"], out["S"], out["E"]
