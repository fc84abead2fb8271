"""Writes tests/golden/c1_minsr_steps.npz: the oracle's trajectory of the CPU-runnable configuration.

Configuration (two electrons): N = 2, rs = 2 (L = rs sqrt(pi N)), Rashba lambda_R = 0.5, h_z = 0;
Psiformer n_freq = 8, d = 32, 2 layers, 4 heads, K = 1, no Jastrow, zero-initialised backflow,
plane-wave spinor envelope of the non-interacting ground state (twist 0); Ewald with
eta = sqrt(pi)/L, 3 x 3 real-space images (n_real = 1), reciprocal box |m| <= 6; 256 walkers drawn
uniform in the cell with random spins (numpy seed 0); Metropolis step 0.3 rs with 2 spin flips per
sweep; 80 burn sweeps, then 10 MinSR steps of 10 sweeps each with shift 1e-4, lr 0.05, clip k = 5
(src/nqs/minsr.py:92-156).  Initial parameters: oracle.psiformer.init_theta(arch, 0).
Random streams: oracle.sampler.make_streams(n, 256, 2, 2, seed) with seed 7 for the burn sweeps and
1000 + k for step k.

Stored (float64 unless noted): burn_R, burn_S (int8), burn_acc_pos, burn_acc_spin (int16); per step
k (leading axis 10): R, S (the walkers after the step's sweeps), acc_pos, acc_spin, e (complex local
energies incl. Coulomb), E, Eerr, dtheta.  theta_k = theta_0 - 0.05 sum_{j<k} dtheta_j.

Run from the repository root:  python tests/golden/make_c1_minsr_steps.py   (~6 min on 8 cores)
"""
import math
import os
import sys

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import bands, cpu_step, ewald as oew, psiformer as opf, sampler as osa  # noqa: E402

N, RS, LAM_R, H_Z = 2, 2.0, 0.5, 0.0
L = RS * math.sqrt(math.pi * N)
W, N_STEPS, N_SWEEPS, BURN = 256, 10, 10, 80
STEP = 0.3 * RS
SHIFT, LR, CLIP_K = 1e-4, 0.05, 5.0
BURN_SEED, STEP_SEED0 = 7, 1000


def arch():
    g = bands.rashba_gs(L, N, 1.0, LAM_R, H_Z, n_cut=10, twist=(0.0, 0.0))
    return opf.Arch(N, g["Gs"], g["W"], L, embed_dim=32, depth=2, n_heads=4, n_freq=8, n_dets=1)


def coulomb():
    return oew.make_ewald(L, 1.0, ewald_alpha=math.sqrt(math.pi) / L, n_real=1, n_recip=6)[2]


def initial_walkers():
    rng = np.random.default_rng(0)
    R = rng.uniform(0.0, 1.0, (W, N, 2)) @ (np.eye(2) * L).T
    S = rng.integers(0, 2, (W, N))
    return R, S


def main(out=os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_minsr_steps.npz")):
    a = arch()
    theta = opf.init_theta(a, 0)
    ce = coulomb()
    R, S = initial_walkers()
    p = opf.unravel(a, theta)
    lpb = torch.func.vmap(lambda r, s: opf.logpsi_from_params(a, p, r, s))
    logpsi_b = lambda R_, S_: lpb(torch.as_tensor(R_), torch.as_tensor(S_)).numpy()
    st = osa.make_streams(BURN, W, N, 2, BURN_SEED)
    bap, bas = np.zeros(W, np.int64), np.zeros(W, np.int64)
    with np.errstate(invalid="ignore"):
        for s in range(BURN):
            R, S, acc, asp, _ = osa.sweep(logpsi_b, R, S, a.A, STEP, st["xi"][s], st["u_pos"][s], st["flip_i"][s],
                                          st["u_flip"][s])
            bap += acc
            bas += asp.sum(0)
    rec = dict(burn_R=R.copy(), burn_S=S.astype(np.int8), burn_acc_pos=bap.astype(np.int16),
               burn_acc_spin=bas.astype(np.int16))
    keys = ("R", "S", "acc_pos", "acc_spin", "e", "E", "Eerr", "dtheta")
    steps = {k: [] for k in keys}
    for k in range(N_STEPS):
        st = osa.make_streams(N_SWEEPS, W, N, 2, STEP_SEED0 + k)
        with np.errstate(invalid="ignore"):
            o = cpu_step.vmc_step_detail(a, theta, R, S, lambda_r=LAM_R, h_z=H_Z, coulomb=ce, streams=st, step=STEP,
                                         lam=SHIFT, lr=LR, clip_k=CLIP_K)
        R, S, theta = o["R"], o["S"], o["theta"]
        for key in keys:
            steps[key].append(o[key].numpy() if torch.is_tensor(o[key]) else o[key])
        print(f"step {k}: E = {o['E']:+.8f} +- {o['Eerr']:.2e}  |dtheta| = {float(torch.linalg.norm(o['dtheta'])):.4e}",
              flush=True)
    for key in keys:
        v = np.array(steps[key])
        rec["step_" + key] = v.astype(np.int8) if key == "S" else (v.astype(np.int16) if "acc" in key else v)
    np.savez_compressed(out, **rec)
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
