"""The mid-size configuration on the GPU: N = 8, rs = 2, lambda = 0.5, h_z = 0.1, d = 64, 3 layers,
4 heads, K = 4 determinants, RBF pair Jastrow with 16 knots, learnable envelope (P = 107,240),
walkers drawn by the kernel's own Metropolis chain.

* The differentiable log psi (nqs.autograd.psiformer_logpsi, forward nqs_psiformer_fwd, backward
  the hand-written reverse sweep / forward-mode tangents) against central finite differences of
  the same forward, along random parameter directions (h = 1e-5, float64) and per coordinate.
* O rows against the oracle's reverse-mode autodiff (the restatement of per_sample_O,
  src/nqs/minsr.py:56-65) to 1e-9 relative; O columns against central differences.
* One full VMC step (20 sweeps, E_loc, O, MinSR) on the plain FP64 path and on the mixed path at
  4096 walkers.
"""
import math

import numpy as np
import pytest
import torch

from oracle import psiformer as opf, minsr as oms
from nqs.rashba_reference import noninteracting_gs, occupied_orbital_matrix, planewave_envelope
from nqs.psiformer import ContinuumSpinorPsiformer, ravel_params, LogPsi
from nqs.sampler import Chain
from nqs.autograd import psiformer_logpsi
from nqs import minsr as nms
from nqs.ewald import make_ewald

pytestmark = pytest.mark.gpu

N_E, RS, LAM, HZ = 8, 2.0, 0.5, 0.1
ARCH = dict(embed_dim=64, depth=3, n_heads=4, n_freq=8, n_dets=4, use_jastrow=True, jastrow_kind="rbf",
            n_knots=16, learnable_env=True)


def _c2(precision="fp64", seed=0):
    L = RS * math.sqrt(math.pi * N_E)
    g = noninteracting_gs(N_E, L, 1.0, LAM, HZ, n_cut=10, twist=(0.1, 0.17))
    Gs, Wenv = planewave_envelope(*occupied_orbital_matrix(g))
    m = ContinuumSpinorPsiformer(n_elec=N_E, Gs=Gs, W=Wenv, L=L, precision=precision, **ARCH)
    params = m.init(seed)
    rng = np.random.default_rng(seed + 1)
    kr = params["params"]["backflow"]["kernel"]
    z = torch.tensor(rng.standard_normal(tuple(kr.shape)) * 0.05)
    params["params"]["backflow"]["kernel"] = (z + 1j * z).to(torch.complex128)
    theta = ravel_params(params)
    return dict(L=L, Gs=Gs, Wenv=Wenv, model=m, theta=theta)


@pytest.fixture(scope="module")
def c2():
    c = _c2()
    W = 256
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    R = torch.rand(W, N_E, 2, generator=g, dtype=torch.float64, device="cuda") * c["L"]
    S = torch.randint(0, 2, (W, N_E), generator=g, device="cuda").to(torch.int32)
    ch = Chain(LogPsi(c["model"], c["theta"].cuda()), R, S)
    ch.run(20, 0.3 * RS, 2, key=5)
    c["R"], c["S"] = ch.R.clone(), ch.S.clone()
    return c


def test_c2_parameter_count(c2):
    assert c2["model"].n_params() == 107240


def test_autograd_theta_matches_central_differences(c2):
    theta = c2["theta"].cuda().clone().requires_grad_(True)
    R, S = c2["R"], c2["S"]
    W = R.shape[0]
    rng = np.random.default_rng(11)
    a = torch.tensor(rng.normal(size=(W, 2)), device="cuda")
    y = psiformer_logpsi(c2["model"], theta, R, S)
    (y * a).sum().backward()
    grad = theta.grad.detach()
    h = 1e-5
    for k in range(3):
        v = torch.tensor(rng.normal(size=theta.shape[0]), device="cuda")
        v /= v.norm()
        with torch.no_grad():
            lp = torch.view_as_real(LogPsi(c2["model"], theta.detach() + h * v)(R, S))
            lm = torch.view_as_real(LogPsi(c2["model"], theta.detach() - h * v)(R, S))
            fd = float(((lp - lm) * a).sum() / (2 * h))
        an = float(grad @ v)
        assert abs(an - fd) < 1e-6 * max(1.0, abs(fd)), (an, fd)


def test_autograd_positions_match_central_differences(c2):
    R = c2["R"][:32].clone().requires_grad_(True)
    S = c2["S"][:32]
    theta = c2["theta"].cuda()
    rng = np.random.default_rng(12)
    a = torch.tensor(rng.normal(size=(32, 2)), device="cuda")
    y = psiformer_logpsi(c2["model"], theta, R, S)
    (y * a).sum().backward()
    gR = R.grad.detach()
    lpf = LogPsi(c2["model"], theta)
    h = 1e-5
    for (i, c) in [(0, 0), (3, 1), (7, 0)]:
        e = torch.zeros_like(R)
        e[:, i, c] = h
        with torch.no_grad():
            fd = (torch.view_as_real(lpf(R.detach() + e, S)) - torch.view_as_real(lpf(R.detach() - e, S))) / (2 * h)
            fd = (fd * a).sum(1)
        assert float((gR[:, i, c] - fd).abs().max()) < 1e-6 * max(1.0, float(fd.abs().max()))


def test_O_rows_match_oracle_autodiff(c2):
    arch = opf.Arch(N_E, c2["Gs"], c2["Wenv"], c2["L"], **ARCH)
    R = c2["R"][:3].cpu().numpy()
    S = c2["S"][:3].cpu().numpy()
    O = nms.per_sample_O(c2["model"], c2["theta"].cuda(), None, R, S).cpu().numpy()
    Oo = oms.per_sample_O(arch, c2["theta"], R, S).numpy()
    assert np.max(np.abs(O - Oo)) < 1e-9 * np.max(np.abs(Oo))


def test_O_columns_match_central_differences(c2):
    R, S = c2["R"][:64], c2["S"][:64]
    theta = c2["theta"].cuda()
    O = nms.per_sample_O(c2["model"], theta, None, R, S)
    P = theta.shape[0]
    h = 1e-5
    rng = np.random.default_rng(13)
    for k in list(rng.choice(P, size=16, replace=False)) + [0, P - 1]:
        e = torch.zeros_like(theta)
        e[int(k)] = h
        fd = (LogPsi(c2["model"], theta + e)(R, S) - LogPsi(c2["model"], theta - e)(R, S)) / (2 * h)
        col = O[:, int(k)]
        assert float((col - fd).abs().max()) < 1e-6 * max(1.0, float(fd.abs().max())), int(k)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_c2_vmc_step(precision):
    c = _c2(precision)
    W = 4096
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    R = torch.rand(W, N_E, 2, generator=g, dtype=torch.float64, device="cuda") * c["L"]
    S = torch.randint(0, 2, (W, N_E), generator=g, device="cuda").to(torch.int32)
    _, _, coul = make_ewald(c["L"], 1.0, n_real=4, n_recip=8)
    vs = nms.VMCStep(c["model"], c["theta"], R, S, m_star=1.0, lambda_r=LAM, h_z=HZ, coulomb=coul, lr=0.05, lam=1e-4,
                     n_sweeps=20, step_size=0.3 * RS, clip_k=5.0, gram="auto", refine=1)
    vs.chain.run(20, 0.3 * RS, 2, key=1)
    st = vs.step(key=2).cpu().numpy()
    assert np.all(np.isfinite(st))
    assert int(vs.last_info.item()) == 0 and vs.failed_steps == 0
    assert bool(torch.isfinite(vs.last["dtheta"]).all()) and float(vs.last["dtheta"].abs().max()) > 0
