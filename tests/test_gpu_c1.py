"""The CPU-runnable configuration, whole MinSR steps on the GPU against the oracle's trajectory.

Configuration and trajectory: tests/golden/make_c1_minsr_steps.py (N = 2, rs = 2, lambda_R = 0.5,
h_z = 0, d = 32, 2 layers, 4 heads, K = 1, no Jastrow, n_freq = 8, zero-initialised backflow; Ewald
eta = sqrt(pi)/L with 3 x 3 images and |m| <= 6; 256 walkers; 80 burn sweeps; 10 MinSR steps of 10
sweeps, shift 1e-4, lr 0.05, clip k = 5), written by the float64 torch oracle
(oracle/cpu_step.py: sweeps with log psi recomputed per proposal, dense-Hessian E_loc + Ewald,
reverse-mode O, median/MAD clip, LU solve of the sample-space system; src/nqs/minsr.py:139-156).

The float64 GPU path (nqs_mcmc_sweeps, nqs_eloc + nqs_ewald, nqs_clip_energies, nqs_psiformer_grad,
nqs_minsr_direction with the float64 Gram and Cholesky) runs the burn sweeps from the same uniform
start and then each step from the oracle's (theta, R, S) with the oracle's random streams:
  * every accept/reject decision equal (accept counts per walker, positions and spins bit for bit);
  * local energies 1e-9, step energy and error bar 1e-10 relative;
  * dtheta and theta' within 1e-8 relative (2-norm) of the oracle's.
"""
import importlib.util
import math
import os

import numpy as np
import pytest
import torch

from oracle import sampler as osa
from nqs.psiformer import ContinuumSpinorPsiformer, make_logpsi
from nqs.ewald import make_ewald
from nqs.sampler import Chain
from nqs import minsr as nms

pytestmark = pytest.mark.gpu

_HERE = os.path.dirname(os.path.abspath(__file__))
_spec = importlib.util.spec_from_file_location("make_c1", os.path.join(_HERE, "golden", "make_c1_minsr_steps.py"))
c1 = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(c1)


@pytest.fixture(scope="module")
def setup():
    gold = dict(np.load(os.path.join(_HERE, "golden", "c1_minsr_steps.npz")))
    a = c1.arch()
    model = ContinuumSpinorPsiformer(n_elec=c1.N, Gs=a.Gs, W=a.W, L=c1.L, embed_dim=32, depth=2, n_heads=4,
                                     n_freq=8, n_dets=1, precision="fp64")
    _, _, coul = make_ewald(c1.L, 1.0, ewald_alpha=math.sqrt(math.pi) / c1.L, n_real=1, n_recip=6)
    return gold, a, model, coul


def _theta(a, gold, k):
    from oracle import psiformer as opf
    th = opf.init_theta(a, 0)
    for j in range(k):
        th = th - c1.LR * torch.as_tensor(gold["step_dtheta"][j])
    return th


def test_c1_burn_sweeps_bit_exact(setup):
    gold, a, model, _ = setup
    R0, S0 = c1.initial_walkers()
    ch = Chain(make_logpsi(model, _theta(a, gold, 0).cuda()), R0, S0)
    st = osa.make_streams(c1.BURN, c1.W, c1.N, 2, c1.BURN_SEED)
    ap, asp = ch.run(c1.BURN, c1.STEP, 2, streams=st)
    assert np.array_equal(np.rint(ap.cpu().numpy() * c1.BURN).astype(int), gold["burn_acc_pos"])
    assert np.array_equal(np.rint(asp.cpu().numpy() * c1.BURN * 2).astype(int), gold["burn_acc_spin"])
    assert np.array_equal(ch.S.cpu().numpy(), gold["burn_S"])
    assert np.array_equal(ch.R.cpu().numpy(), gold["burn_R"])


@pytest.mark.parametrize("k", range(10))
def test_c1_minsr_step_matches_oracle(setup, k):
    gold, a, model, coul = setup
    th = _theta(a, gold, k)
    R = gold["burn_R"] if k == 0 else gold["step_R"][k - 1]
    S = gold["burn_S"] if k == 0 else gold["step_S"][k - 1]
    vs = nms.VMCStep(model, th, R, S, m_star=1.0, lambda_r=c1.LAM_R, h_z=c1.H_Z, coulomb=coul, lr=c1.LR,
                     lam=c1.SHIFT, n_sweeps=c1.N_SWEEPS, step_size=c1.STEP, clip_k=c1.CLIP_K, gram="fp64")
    st = osa.make_streams(c1.N_SWEEPS, c1.W, c1.N, 2, c1.STEP_SEED0 + k)
    stats = vs.step(streams=st).cpu().numpy()
    last = vs.last
    # decisions
    assert np.array_equal(np.rint(last["acc"].cpu().numpy() * c1.N_SWEEPS).astype(int), gold["step_acc_pos"][k])
    assert np.array_equal(np.rint(last["acc_spin"].cpu().numpy() * c1.N_SWEEPS * 2).astype(int),
                          gold["step_acc_spin"][k])
    assert np.array_equal(vs.chain.S.cpu().numpy(), gold["step_S"][k])
    assert np.array_equal(vs.chain.R.cpu().numpy(), gold["step_R"][k])
    # local energies, step energy
    e = last["e"].cpu().numpy()
    eg = gold["step_e"][k]
    err_e = np.max(np.abs(e - eg)) / max(1.0, np.max(np.abs(eg)))
    assert err_e < 1e-9, err_e
    assert abs(stats[0] - gold["step_E"][k]) < 1e-10 * max(1.0, abs(gold["step_E"][k]))
    assert abs(stats[1] - gold["step_Eerr"][k]) < 1e-10 * max(1.0, abs(gold["step_Eerr"][k]))
    # update
    d = last["dtheta"].cpu()
    dg = torch.as_tensor(gold["step_dtheta"][k])
    rel_d = float(torch.linalg.norm(d - dg) / torch.linalg.norm(dg))
    th_ref = th - c1.LR * dg
    rel_t = float(torch.linalg.norm(vs.theta.cpu() - th_ref) / torch.linalg.norm(th_ref))
    print(f"step {k}: |de| {err_e:.1e}  E {stats[0]:+.10f}  dtheta rel {rel_d:.2e}  theta' rel {rel_t:.2e}", flush=True)
    assert vs.failed_steps == 0
    assert rel_d < 1e-8, rel_d
    assert rel_t < 1e-8, rel_t
