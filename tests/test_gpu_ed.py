"""Two electrons against exact diagonalisation (oracle/ed.py), on the GPU path end to end:
MinSR training (sampler, forward-Laplacian E_loc, Ewald, per-walker O, Gram, Cholesky) of the
Psiformer of the CPU-runnable configuration (N = 2, rs = 2, F = 8, d = 32, 2 layers, 4 heads, K = 1,
zero-initialised backflow, plane-wave envelope), with the smooth periodic cusp Jastrow.

test_two_electron_vmc_reaches_the_exact_diagonalisation_energy: lambda_R = 0 with the smooth
periodic cusp Jastrow (a project extension), 2048 walkers, 300 steps.

test_two_electron_vmc_at_the_rashba_coupling: lambda_R = 0.5, the configuration's coupling, no
Jastrow.  The plane-wave start (both electrons in the k = 0 orbitals, pure up and pure down) is a
spin singlet with exactly zero amplitude on parallel-spin configurations, and a multiplicative
backflow keeps it so; the sampler never visits them and the sampled energy gradient has no
component that leaves the singlet, so training from it stays at the lambda_R = 0 energy.  The
reference model's own knobs break this: learnable_env = True with an envelope of five plane waves
(k = 0 and the first shell, the extra ones with zero weight at the start) and the learnable
correction dW drawn small and random (N(0, 0.05^2) per real component) instead of zero.  MinSR then
runs 3000 steps with shift 0.1 and lr 0.05, then 3000 with shift 0.03 and lr 0.02, 4096 walkers;
the energy is averaged over 20 blocks of 32768 walkers.  It takes about 6 minutes of GPU time, so
it runs only when NQS_LONG_TESTS=1.

Bound: |E_VMC - E_ED| < 0.7 mHa (the figure the reference quotes for its two-electron benchmark).
"""
import os
import math

import numpy as np
import pytest
import torch

from oracle import ed
from nqs.rashba_reference import noninteracting_gs, occupied_orbital_matrix, planewave_envelope
from nqs.psiformer import ContinuumSpinorPsiformer, ravel_params, LogPsi
from nqs.local_energy import make_local_energy
from nqs.ewald import make_ewald
from nqs.sampler import Chain
from nqs import minsr as nms

pytestmark = pytest.mark.gpu


def test_two_electron_vmc_reaches_the_exact_diagonalisation_energy():
    N, RS, LAM = 2, 2.0, 0.0
    L = RS * math.sqrt(math.pi * N)
    _, E_ed = ed.extrapolated_energy(L, [10, 12, 14], power=2.0, lambda_r=LAM, alpha=1.0)
    g = noninteracting_gs(N, L, 1.0, LAM, 0.0, n_cut=10, twist=(0.0, 0.0))
    Gs, Wenv = planewave_envelope(*occupied_orbital_matrix(g))
    m = ContinuumSpinorPsiformer(n_elec=N, Gs=Gs, W=Wenv, L=L, embed_dim=32, depth=2, n_heads=4, n_freq=8, n_dets=1,
                                 use_jastrow=True, jastrow_kind="cusp_periodic")
    _, _, coul = make_ewald(L, 1.0, n_real=4, n_recip=8)
    params, hist, R, S = nms.train_minsr(m, m.init(0), L, N, 0, m_star=1.0, lambda_r=LAM, h_z=0.0, coulomb=coul,
                                         n_walkers=2048, n_steps=300, lr=0.05, lam=1e-3, n_sweeps=10,
                                         step_size=0.3 * RS, burn=50, log_every=0, return_walkers=True)
    lp = LogPsi(m, ravel_params(params).cuda())
    E = make_local_energy(lp, 1.0, LAM, 0.0)
    ch = Chain(lp, R.repeat(8, 1, 1).contiguous(), S.repeat(8, 1).contiguous())
    ch.run(50, 0.3 * RS, 2, key=99)
    blocks = []
    for b in range(20):
        ch.run(5, 0.3 * RS, 2, key=1000 + b)
        blocks.append(float((E(ch.R, ch.S) + coul(ch.R)).real.mean()))
    blocks = np.array(blocks)
    E_vmc, err = blocks.mean(), blocks.std() / np.sqrt(len(blocks))
    print(f"E_VMC = {E_vmc:.6f} +- {err:.6f}, E_ED = {E_ed:.6f}, diff = {1e3 * (E_vmc - E_ed):.3f} mHa", flush=True)
    assert abs(E_vmc - E_ed) < 7e-4
    assert E_vmc > E_ed - 5 * err - 1e-4          # variational, within noise and the ED extrapolation


@pytest.mark.skipif(os.environ.get("NQS_LONG_TESTS") != "1", reason="about 6 GPU minutes: set NQS_LONG_TESTS=1")
def test_two_electron_vmc_at_the_rashba_coupling():
    N, RS, LAM = 2, 2.0, 0.5
    L = RS * math.sqrt(math.pi * N)
    _, E_ed = ed.extrapolated_energy(L, [10, 12, 14], power=2.0, lambda_r=LAM, alpha=1.0)
    b = 2 * math.pi / L
    Gs = np.array([[0, 0], [b, 0], [-b, 0], [0, b], [0, -b]], float)
    Wenv = np.zeros((N, 5, 2), complex)
    Wenv[0, 0, 0] = Wenv[1, 0, 1] = 1.0
    m = ContinuumSpinorPsiformer(n_elec=N, Gs=Gs, W=Wenv, L=L, embed_dim=32, depth=2, n_heads=4, n_freq=8, n_dets=1,
                                 learnable_env=True)
    params = m.init(0)
    rng = np.random.default_rng(2)
    shp = params["params"]["dW"].shape
    params["params"]["dW"] = torch.tensor(0.05 * (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)))
    _, _, coul = make_ewald(L, 1.0, n_real=4, n_recip=8)
    params, _, R, S = nms.train_minsr(m, params, L, N, 1, m_star=1.0, lambda_r=LAM, h_z=0.0, coulomb=coul,
                                      n_walkers=4096, n_steps=3000, lr=0.05, lam=0.1, n_sweeps=10, step_size=0.3 * RS,
                                      burn=50, log_every=500, return_walkers=True)
    params, hist, R, S = nms.train_minsr(m, params, L, N, 1, m_star=1.0, lambda_r=LAM, h_z=0.0, coulomb=coul,
                                         n_walkers=4096, n_steps=3000, lr=0.02, lam=0.03, n_sweeps=10,
                                         step_size=0.3 * RS, burn=0, log_every=500, return_walkers=True, R0=R, S0=S)
    assert hist["failed_steps"][-1] == 0
    lp = LogPsi(m, ravel_params(params).cuda())
    E = make_local_energy(lp, 1.0, LAM, 0.0)
    ch = Chain(lp, R.repeat(8, 1, 1).contiguous(), S.repeat(8, 1).contiguous())
    ch.run(50, 0.3 * RS, 2, key=99)
    blocks = []
    for k in range(20):
        ch.run(5, 0.3 * RS, 2, key=1000 + k)
        blocks.append(float((E(ch.R, ch.S) + coul(ch.R)).real.mean()))
    blocks = np.array(blocks)
    E_vmc, err = blocks.mean(), blocks.std() / np.sqrt(len(blocks))
    par = float((ch.S.sum(1) != 1).double().mean())
    print(f"E_VMC = {E_vmc:.6f} +- {err:.6f}, E_ED = {E_ed:.6f}, diff = {1e3 * (E_vmc - E_ed):.3f} mHa, "
          f"parallel-spin fraction {par:.3f}", flush=True)
    assert par > 0.01                              # left the singlet sector
    assert abs(E_vmc - E_ed) < 7e-4
