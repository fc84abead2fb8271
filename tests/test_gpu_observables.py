"""Spin observables, overlaps and rotation eigenvalues on the GPU path.

* nqs_spin_estimators against the oracle's flip ratios on a Psiformer (1e-10).
* The analytic case of the reference's own test (test_nqs_observables.py:17-39): a spin-independent
  state (every orbital's spinor (1, 1)/sqrt 2) is the +x product state, so sigma_x = 1 and
  Re sigma_y = 0 for every electron and m_x = 1, m_y = 0 per electron; S_perp peaks at q = 0.
* nqs_spin_texture and nqs_spin_fourier against the oracle's numpy sums on the same spins (bin
  counts bit for bit, sums 1e-12); nqs_ratio_stats against numpy (1e-12).
* mc_overlap of two plane-wave determinants against det(Wa^* Wb^T) with walkers from the GPU
  sampler (test_manybody_chern.py:32-50's tolerance).
* The C4 / C2 rotation eigenvalue of an exact C4 eigenstate: zero variance and equal to det(m)
  (test_nqs_topology.py:80-122); a non-symmetric state within the MC error of det(m).
"""
import numpy as np
import pytest
import torch

from oracle import bands, psiformer as opf, observables as oob
from nqs.psiformer import ContinuumSpinorPsiformer, make_logpsi
from nqs.slater_reference import make_logpsi_general
from nqs.rashba_reference import planewave_envelope
from nqs.sampler import run_chain
from nqs import observables as nob, manybody_chern as nmc, topology as ntp

pytestmark = pytest.mark.gpu


def test_spin_estimators_match_oracle():
    N, L = 5, 7.0
    g = bands.rashba_gs(L, N, 1.0, 1.0, 0.5, n_cut=10, twist=(0.1, 0.17))
    kw = dict(embed_dim=16, depth=2, n_heads=4, n_freq=4, n_dets=2)
    arch = opf.Arch(N, g["Gs"], g["W"], L, **kw)
    theta = opf.init_theta(arch, 3, backflow_scale=0.05, scale_all=0.02)
    lp = make_logpsi(ContinuumSpinorPsiformer(n_elec=N, Gs=g["Gs"], W=g["W"], L=L, **kw), theta.cuda())
    rng = np.random.default_rng(5)
    R, S = rng.uniform(0, L, (4, N, 2)), rng.integers(0, 2, (4, N))
    sx, sy, sz = [t.cpu().numpy() for t in nob.make_spin_estimators(lp)(R, S)]
    olp = opf.make_logpsi(arch, theta)
    for w in range(4):
        ox, oy, oz = oob.spin_estimators(olp, R[w], S[w])
        scale = max(1.0, np.max(np.abs(ox)))
        assert np.max(np.abs(sx[w] - ox)) < 1e-10 * scale
        assert np.max(np.abs(sy[w] - oy)) < 1e-10 * scale
        assert np.array_equal(sz[w], oz)


def _x_polarised(N, L, seed=0):
    rng = np.random.default_rng(seed)
    B = 2 * np.pi / L
    Gs = np.array([(i * B, j * B) for i in range(-1, 2) for j in range(-1, 2)], float)
    f = rng.normal(size=(N, len(Gs))) + 1j * rng.normal(size=(N, len(Gs)))
    W = np.stack([f, f], -1) / np.sqrt(2.0)
    return make_logpsi_general(Gs, W, L)


def test_spin_independent_state_is_x_polarised():
    N, L, B = 8, 5.0, 1500
    lp = _x_polarised(N, L)
    rng = np.random.default_rng(0)
    R = rng.uniform(0, L, (B, N, 2))
    S = rng.integers(0, 2, (B, N))
    A = np.diag([L, L])
    sx, sy, sz = nob.spin_estimates(lp, R, S)
    assert np.max(np.abs(sx - 1.0)) < 1e-12 and np.max(np.abs(sy)) < 1e-12
    tex = nob.spin_texture(lp, R, S, A, ngrid=10)
    assert abs(tex["mbar"][0] - 1.0) < 1e-9 and abs(tex["mbar"][1]) < 1e-9 and abs(tex["mbar"][2]) < 0.12
    assert abs(nob.berg_luscher(tex["m"])) < 0.15
    sf = nob.structure_factor(lp, R, S, A, nshell=3)
    qn = np.linalg.norm(sf["q"], axis=1)
    assert sf["Sperp"][qn < 1e-9][0] > 3.0 * sf["Sperp"][qn > 1e-9].max()


def test_texture_fourier_and_ratio_kernels_match_numpy():
    rng = np.random.default_rng(4)
    W, N = 777, 6
    A = np.array([[6.0, 1.5], [0.0, 5.0]])
    R = rng.uniform(-3, 9, (W, N, 2))
    sp = rng.normal(size=(W, N, 3))
    Rd = torch.tensor(R, device="cuda")
    spd = torch.tensor(sp, device="cuda")
    for ng in (1, 7, 12):
        tex = nob._texture(spd, Rd, A, ng)
        m, rho = oob.texture_sums(R, sp, A, ng)
        assert np.max(np.abs(tex["rho"] * W - rho)) < 1e-9          # same count in every bin
        assert np.max(np.abs(tex["m"] * W - m)) < 1e-11
    Bm = 2 * np.pi * np.linalg.inv(A).T
    qs = np.array([p * Bm[:, 0] + q * Bm[:, 1] for p in range(-2, 3) for q in range(-2, 3)])
    mq, szz, sperp = nob._fourier(spd, Rd, qs)
    om, oszz, osperp = oob.spin_fourier(R, sp, qs)
    assert np.max(np.abs(mq - om)) < 1e-12 and np.max(np.abs(szz - oszz)) < 1e-11
    assert np.max(np.abs(sperp - osperp)) < 1e-11
    la = rng.normal(size=W) * 3 + 1j * rng.normal(size=W)
    lb = rng.normal(size=W) * 3 + 1j * rng.normal(size=W)
    ph = np.exp(1j * rng.uniform(0, 6, W))
    c = lambda x: torch.tensor(x, device="cuda", dtype=torch.complex128)
    for shift, phase in ((True, None), (False, ph)):
        st = nmc.ratio_stats(c(la), c(lb), None if phase is None else c(phase), shift=shift)
        ref = oob.ratio_stats(la, lb, phase, shift)
        assert np.max(np.abs(st - ref) / np.maximum(1.0, np.abs(ref))) < 1e-12


def test_mc_overlap_matches_exact_determinant():
    rng = np.random.default_rng(7)
    L = 6.0
    B = 2 * np.pi / L
    Gs = np.array([[i * B, j * B] for i in range(-2, 3) for j in range(-2, 3)])
    N, nG = 3, len(Gs)

    def orbitals():
        M = rng.normal(size=(N, nG)) + 1j * rng.normal(size=(N, nG))
        q, _ = np.linalg.qr(M.conj().T)
        return q.conj().T[:N]
    Wa, Wb = orbitals(), orbitals()
    exact = np.linalg.det(Wa.conj() @ Wb.T)
    spinor = lambda Wx: np.stack([Wx, Wx], -1) / np.sqrt(2.0)
    la, lb = make_logpsi_general(Gs, spinor(Wa), L), make_logpsi_general(Gs, spinor(Wb), L)
    R, S, _, _ = run_chain(la, L, N, 0, n_walkers=3072, n_sweeps=40, burn=200, step=0.8)
    O, absO, ess = nmc.mc_overlap(la, lb, R, S)
    print(f"overlap: MC {O:.4f} exact {exact:.4f} ess {ess:.3f}", flush=True)
    assert ess > 0.05
    assert abs(O - exact) < 0.08 * max(0.2, abs(exact)) + 0.03


_U4 = np.array([[np.exp(-1j * np.pi / 4), 0], [0, np.exp(1j * np.pi / 4)]])


def _c4_state(L, chi0, symmetric=True):
    g = 2 * np.pi / L
    Gs = g * np.array([[1, 0], [0, 1], [-1, 0], [0, -1]], float)
    if symmetric:
        Chi = np.stack([np.linalg.matrix_power(_U4, k) @ chi0 for k in range(4)])
    else:
        rng = np.random.default_rng(2)
        Chi = rng.normal(size=(4, 2)) + 1j * rng.normal(size=(4, 2))
        Chi = Chi / np.linalg.norm(Chi, axis=1, keepdims=True)
    return Gs, Chi


def test_rotation_c4_and_c2_of_an_exact_eigenstate():
    L = 7.0
    c0 = np.array([1.0, 0.4 + 0.2j]) / np.linalg.norm([1.0, 0.4 + 0.2j])
    Gs, Chi = _c4_state(L, c0, True)
    lp = make_logpsi_general(*planewave_envelope(Gs, Chi), L)
    rng = np.random.default_rng(0)
    R = rng.uniform(0, L, (40, 4, 2))
    S = rng.integers(0, 2, (40, 4))
    for n in (4, 2):
        exact = oob.rotation_overlap_planewaves(Gs, Chi, n)
        vmc, err = ntp.rotation_eigenvalue_vmc(lp, R, S, n, L)
        assert abs(exact) > 0.999
        assert err.real < 1e-7 and err.imag < 1e-7
        assert abs(vmc - exact) < 1e-7, (n, vmc, exact)


def test_rotation_general_state_matches_exact_by_mcmc():
    L = 7.0
    Gs, Chi = _c4_state(L, None, symmetric=False)
    exact = oob.rotation_overlap_planewaves(Gs, Chi, 4)
    lp = make_logpsi_general(*planewave_envelope(Gs, Chi), L)
    R, S, _, _ = run_chain(lp, L, 4, 0, n_walkers=4000, n_sweeps=200, burn=150, step=0.6)
    vmc, err = ntp.rotation_eigenvalue_vmc(lp, R, S, 4, L)
    print(f"<R4>: MC {vmc:.4f} +- {abs(err):.1e}, exact {exact:.4f}", flush=True)
    assert abs(vmc - exact) < 5 * abs(err) + 5e-3
