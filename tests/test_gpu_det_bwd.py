"""The determinant-head backward of the O rows (nqs_psiformer_grad) on its row-per-lane path (N <= 32,
no learnable envelope: one determinant per group of 16 or 32 lanes, Gauss-Jordan inverses,
log-sum-exp weights exchanged through LDS) against the oracle's reverse-mode autodiff of log psi (the restatement of
per_sample_O, src/nqs/minsr.py:56-65), float64, 1e-9 relative to max |O|, and the log psi the
backward returns against the forward's (1e-12).

Determinant counts that fill the workgroup differently: K = 1 (four configurations per wave),
K = 3 (four configurations over three waves, the last group idle), K = 2 with the cusp Jastrow,
K = 3 with the RBF Jastrow (its O entries and its share of log psi), and K = 16 (one configuration
per workgroup), at N = 5 electrons (16-lane groups with identity padding); and the 32-lane groups
of 16 < N <= 32 at N = 20 (padded, K = 3, RBF Jastrow) and N = 32 (K = 4).
"""
import numpy as np
import pytest
import torch

from oracle import bands, psiformer as opf, minsr as oms
from nqs.psiformer import ContinuumSpinorPsiformer, LogPsi
from nqs import minsr as nms

pytestmark = pytest.mark.gpu

ARCHS = {  # name: (N, architecture)
    "k1": (5, dict(embed_dim=16, depth=2, n_heads=4, n_freq=4)),
    "k3": (5, dict(embed_dim=16, depth=1, n_heads=2, n_freq=3, n_dets=3)),
    "k2_cusp": (5, dict(embed_dim=8, depth=1, n_heads=2, n_freq=2, n_dets=2, use_jastrow=True, jastrow_kind="cusp")),
    "k3_rbf": (5, dict(embed_dim=8, depth=1, n_heads=2, n_freq=2, n_dets=3, use_jastrow=True, n_knots=6)),
    "k16": (5, dict(embed_dim=8, depth=1, n_heads=2, n_freq=2, n_dets=16)),
    "n20_k3_rbf": (20, dict(embed_dim=8, depth=1, n_heads=2, n_freq=2, n_dets=3, use_jastrow=True, n_knots=6)),
    "n32_k4": (32, dict(embed_dim=8, depth=1, n_heads=2, n_freq=2, n_dets=4)),
}


@pytest.mark.parametrize("name", list(ARCHS))
def test_det_backward_rows_match_oracle(name):
    N, kw = ARCHS[name]
    L, W, seed = (7.0 if N <= 16 else 2.0 * np.sqrt(np.pi * N)), (5 if N <= 16 else 2), 11
    g = bands.rashba_gs(L, N, 1.0, 1.0, 0.5, n_cut=10, twist=(0.1, 0.17))
    arch = opf.Arch(N, g["Gs"], g["W"], L, **kw)
    theta = opf.init_theta(arch, seed, backflow_scale=0.05, scale_all=0.02)
    model = ContinuumSpinorPsiformer(n_elec=N, Gs=g["Gs"], W=g["W"], L=L, precision="fp64", **kw)
    lp = LogPsi(model, theta.cuda())
    rng = np.random.default_rng(seed)
    R = rng.uniform(0, L, (W, N, 2))
    S = rng.integers(0, 2, (W, N))
    Rg = torch.tensor(R, device="cuda")
    Sg = torch.tensor(S, device="cuda", dtype=torch.int32)
    O, lpb = nms.grad_params(lp, Rg, Sg)
    Og = O.complex().cpu().numpy()
    Oo = oms.per_sample_O(arch, theta, R, S).numpy()
    err = np.max(np.abs(Og - Oo)) / np.max(np.abs(Oo))
    lpf = lp(Rg, Sg)
    dlp = float((lpb - lpf).abs().max())
    print(f"{name}: |O - O_oracle| / max|O| = {err:.2e}; |log psi(backward) - log psi(forward)| = {dlp:.2e}",
          flush=True)
    assert err < 1e-9
    assert dlp < 1e-12
