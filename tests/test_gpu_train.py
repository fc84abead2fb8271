"""First-order training on the GPU path (src/nqs/train.py:28-98).

* nqs_vmc_gradient (the surrogate-loss gradient from the O rows) against the oracle's reverse-mode
  gradient of the surrogate (1e-10 relative).
* nqs_adam_update against the oracle's optax.adam restatement over several steps, both complex
  conventions (1e-14).
* The reference's own certification (test_nqs_train.py:36-77): perturb the backflow of a
  Psiformer at zero backflow (the exact non-interacting ground state), train with Adam, recover
  at least 95% of the energy excess without going below E_gs.  Run with both complex-gradient
  conventions; each result is printed.
"""
import numpy as np
import pytest
import torch

from oracle import bands, psiformer as opf, train as otr
from nqs.psiformer import ContinuumSpinorPsiformer, LogPsi, make_logpsi
from nqs.rashba_reference import noninteracting_gs, occupied_orbital_matrix, planewave_envelope
from nqs.local_energy import make_local_energy
from nqs.sampler import run_chain, vmc_energy
from nqs.minsr import grad_params
from nqs import train as ntr

pytestmark = pytest.mark.gpu


def _small():
    g = bands.rashba_gs(7.0, 4, 1.0, 1.0, 0.5, n_cut=8, twist=(0.1, 0.17))
    kw = dict(embed_dim=8, depth=1, n_heads=2, n_freq=2)
    arch = opf.Arch(4, g["Gs"], g["W"], 7.0, **kw)
    theta = opf.init_theta(arch, 1, backflow_scale=0.05, scale_all=0.02)
    return arch, theta, ContinuumSpinorPsiformer(n_elec=4, Gs=g["Gs"], W=g["W"], L=7.0, **kw)


def test_vmc_gradient_matches_oracle_surrogate():
    arch, theta, model = _small()
    rng = np.random.default_rng(3)
    R, S = rng.uniform(0, 7.0, (6, 4, 2)), rng.integers(0, 2, (6, 4))
    e = rng.normal(size=6) + 1j * rng.normal(size=6)
    O, _ = grad_params(LogPsi(model, theta.cuda()), R, S)
    g = ntr.vmc_gradient(O, torch.tensor(e)).cpu().numpy()
    ref = otr.surrogate_gradient(arch, theta, R, S, e).numpy()
    assert np.max(np.abs(g - ref)) < 1e-10 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("mode", ["jax", "descent"])
def test_adam_update_matches_oracle(mode):
    arch, theta, model = _small()
    P = theta.shape[0]
    opt = ntr.Adam(lr=0.02, b1=0.8, b2=0.99, eps=1e-7, eps_root=1e-12, complex_grads=mode)
    th = theta.cuda().clone()
    st = opt.init(th)
    th_o, st_o = theta.numpy(), None
    rng = np.random.default_rng(9)
    for it in range(4):
        g = rng.normal(size=P)
        st = opt.update_(th, torch.tensor(g, device="cuda"), st, opt.partner(model))
        th_o, st_o = otr.adam_step(arch, th_o, g, st_o, lr=0.02, b1=0.8, b2=0.99, eps=1e-7, eps_root=1e-12,
                                   complex_grads=mode)
        assert np.max(np.abs(th.cpu().numpy() - th_o)) < 1e-14
    assert st.count == 4


def _perturb_backflow(params, scale, seed):
    """Complex noise on the backflow leaves only (test_nqs_train.py:25-33)."""
    rng = np.random.default_rng(seed)
    bf = params["params"]["backflow"]
    for k in list(bf):
        x = bf[k]
        noise = rng.standard_normal(tuple(x.shape)) + 1j * rng.standard_normal(tuple(x.shape))
        bf[k] = x + torch.tensor(scale * noise, dtype=x.dtype)
    return params


@pytest.mark.parametrize("mode", ["descent", "jax"])
def test_training_descends_to_Egs(mode):
    N, L, lr_soc, hz = 5, 7.0, 1.0, 0.5
    g = noninteracting_gs(N, L, 1.0, lr_soc, hz, n_cut=10, twist=(0.1, 0.17))
    Gs, W = planewave_envelope(*occupied_orbital_matrix(g))
    model = ContinuumSpinorPsiformer(n_elec=N, Gs=Gs, W=W, L=L, embed_dim=16, depth=2, n_heads=4, n_freq=4)
    params = _perturb_backflow(model.init(0), 0.15, 2)

    def energy_now(p):
        lp = make_logpsi(model, p)
        R, S, _, _ = run_chain(lp, L, N, 7, n_walkers=256, n_sweeps=80, burn=80, step=0.5)
        return vmc_energy(make_local_energy(lp, 1.0, lr_soc, hz), R, S)[0]

    E_init = energy_now(params)
    params, hist = ntr.train(model, params, L, N, 3, m_star=1.0, lambda_r=lr_soc, h_z=hz, n_walkers=256, n_steps=120,
                             lr=3e-3, n_sweeps=8, step_size=0.5, burn=40, log_every=30,
                             optimizer=ntr.Adam(3e-3, complex_grads=mode))
    E_final = energy_now(params)
    excess = E_init - g["E_gs"]
    frac_remaining = (E_final - g["E_gs"]) / excess
    print(f"complex_grads={mode}: E_gs={g['E_gs']:+.6f} E_init={E_init:+.6f} E_final={E_final:+.6f} "
          f"recovered={100 * (1 - frac_remaining):.1f}%", flush=True)
    assert excess > 1e-2
    assert E_final < E_init - 1e-3
    assert frac_remaining < 0.05
    assert E_final > g["E_gs"] - 0.02
