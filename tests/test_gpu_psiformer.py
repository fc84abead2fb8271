"""GPU parity of log psi and E_loc (HIP path through the C ABI) against the CPU oracle.

Mirrors the reference's hot-path tests (tests/test_nqs_psiformer.py, test_nqs_local_energy.py,
test_nqs_spiral.py, test_nqs_production.py, test_nqs_folx.py) and adds value parity of the
perturbed network against the oracle's dense-autodiff E_loc.
Tolerances: float64 path, log psi 1e-10 (mod 2 pi i), E_loc 1e-8 relative; mixed path (float32
backbone) log psi 2e-4, E_loc 5e-3 relative.
"""
import numpy as np
import pytest
import torch

from oracle import bands, psiformer as opf, local_energy as ole
from nqs.psiformer import ContinuumSpinorPsiformer, make_logpsi
from nqs.local_energy import make_local_energy, ExtField
from nqs import spiral_reference as nsp
from nqs.slater_reference import make_logpsi_general
from nqs.cell import rect_cell

pytestmark = pytest.mark.gpu

ARCHS = {
    "base": dict(embed_dim=16, depth=2, n_heads=4, n_freq=4),
    "kdet_jastrow_env": dict(embed_dim=16, depth=2, n_heads=2, n_freq=3, n_dets=3, use_jastrow=True,
                             learnable_env=True, n_knots=6),
    "cusp": dict(embed_dim=8, depth=1, n_heads=2, n_freq=2, n_dets=2, use_jastrow=True, jastrow_kind="cusp"),
    "cusp_periodic": dict(embed_dim=8, depth=1, n_heads=2, n_freq=2, n_dets=1, use_jastrow=True,
                          jastrow_kind="cusp_periodic"),
}


def _system(N=5, L=7.0, lr=1.0, hz=0.5):
    g = bands.rashba_gs(L, N, 1.0, lr, hz, n_cut=10, twist=(0.1, 0.17))
    return g, g["Gs"], g["W"]


def _pair(name, N=5, L=7.0, precision="fp64", seed=3, bf=0.05, pert=0.02):
    g, Gs, W = _system(N, L)
    kw = ARCHS[name]
    arch = opf.Arch(N, Gs, W, L, **kw)
    theta = opf.init_theta(arch, seed, backflow_scale=bf, scale_all=pert)
    model = ContinuumSpinorPsiformer(n_elec=N, Gs=Gs, W=W, L=L, precision=precision, **kw)
    return g, arch, theta, model


def _configs(W, N, L, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0, L, (W, N, 2)), rng.integers(0, 2, (W, N))


def _mod2pi(x):
    return (x + np.pi) % (2 * np.pi) - np.pi


@pytest.mark.parametrize("name", list(ARCHS))
def test_logpsi_matches_oracle(name):
    g, arch, theta, model = _pair(name)
    lp = make_logpsi(model, theta.cuda())
    R, S = _configs(6, arch.N, 7.0, 11)
    v = lp(R, S).cpu().numpy()
    ref = np.array([complex(opf.logpsi(arch, theta, torch.tensor(R[w]), torch.tensor(S[w]))) for w in range(len(R))])
    assert np.max(np.abs(v.real - ref.real)) < 1e-10
    assert np.max(np.abs(_mod2pi(v.imag - ref.imag))) < 1e-10


@pytest.mark.parametrize("name", list(ARCHS))
def test_eloc_matches_dense_autodiff_oracle(name):
    g, arch, theta, model = _pair(name)
    lp = make_logpsi(model, theta.cuda())
    E = make_local_energy(lp, 1.0, 1.0, 0.5)
    R, S = _configs(4, arch.N, 7.0, 12)
    e, l0, grad, lap = [t.cpu().numpy() for t in E.parts(R, S)]
    Eo = ole.make_local_energy(opf.make_logpsi(arch, theta), 1.0, 1.0, 0.5)
    for w in range(len(R)):
        p = Eo.parts(torch.tensor(R[w]), torch.tensor(S[w]))
        scale = max(1.0, abs(p["kin"]))
        assert abs(e[w] - (p["kin"] + p["zeeman"] + p["soc"])) < 1e-8 * scale
        assert np.max(np.abs(grad[w] - p["grad"].numpy())) < 1e-9 * max(1.0, np.max(np.abs(p["grad"].numpy())))
        assert abs(lap[w] - p["lap"]) < 1e-8 * max(1.0, abs(p["lap"]))


def test_mixed_precision_close_to_fp64():
    g, arch, theta, model = _pair("base", precision="mixed")
    lp = make_logpsi(model, theta.cuda())
    R, S = _configs(8, arch.N, 7.0, 13)
    v = lp(R, S).cpu().numpy()
    ref = np.array([complex(opf.logpsi(arch, theta, torch.tensor(R[w]), torch.tensor(S[w]))) for w in range(len(R))])
    assert np.max(np.abs(v.real - ref.real)) < 2e-4
    E = make_local_energy(lp, 1.0, 1.0, 0.5)(R[:3], S[:3]).cpu().numpy()
    Eo = ole.make_local_energy(opf.make_logpsi(arch, theta), 1.0, 1.0, 0.5)
    Er = np.array([Eo(torch.tensor(R[w]), torch.tensor(S[w])) for w in range(3)])
    assert np.max(np.abs(E - Er)) < 5e-3 * max(1.0, np.max(np.abs(Er)))


@pytest.mark.parametrize("N,L,lr,hz", [(5, 7.0, 1.0, 0.5), (7, 8.0, 1.0, 0.5), (9, 9.0, 1.0, 0.0)])
def test_zero_variance_slater(N, L, lr, hz):
    """test_nqs_local_energy.py:38-46 on the HIP path: E_loc == E_gs at every configuration."""
    g = bands.rashba_gs(L, N, 1.0, lr, hz, n_cut=10, twist=(0.1, 0.17))
    lp = make_logpsi_general(g["Gs"], g["W"], L)
    R, S = _configs(24, N, L, 0)
    e = make_local_energy(lp, 1.0, lr, hz)(R, S).cpu().numpy()
    assert np.std(e.real) < 1e-7
    assert abs(np.mean(e.real) - g["E_gs"]) < 1e-7
    assert np.max(np.abs(e.imag)) < 1e-7


def test_zeeman_only_polarized_limit():
    """test_nqs_local_energy.py:49-66: lambda=0, h_z=4, all spins up; exercises the singular-flip guard."""
    N, L, hz = 5, 9.0, 4.0
    g = bands.rashba_gs(L, N, 1.0, 0.0, hz, n_cut=10, twist=(0.1, 0.17))
    assert all(b == 0 for b in g["bands"])
    lp = make_logpsi_general(g["Gs"], g["W"], L)
    rng = np.random.default_rng(1)
    R = rng.uniform(0.0, L, size=(1, N, 2))
    S = np.zeros((1, N), dtype=int)
    e = complex(make_local_energy(lp, 1.0, 0.0, hz)(R, S).cpu().numpy()[0])
    ekin = float(np.sum(np.sum(g["Gs"] ** 2, axis=1) / 2.0))
    assert abs(e.real - (ekin - 0.5 * hz * N)) < 1e-7
    assert abs(e.real - g["E_gs"]) < 1e-7


def test_psiformer_zero_backflow_is_slater_and_zero_variance():
    """test_nqs_psiformer.py:36-57."""
    g, arch, theta, model = _pair("base", bf=0.0, pert=0.0)
    lp = make_logpsi(model, theta.cuda())
    ls = make_logpsi_general(g["Gs"], g["W"], 7.0)
    R, S = _configs(8, 5, 7.0, 3)
    d = lp(R, S).cpu().numpy() - ls(R, S).cpu().numpy()
    assert np.max(np.abs(d.real)) < 1e-9 and np.max(np.abs(_mod2pi(d.imag))) < 1e-9
    e = make_local_energy(lp, 1.0, 1.0, 0.5)(R, S).cpu().numpy()
    assert np.std(e.real) < 1e-6 and abs(e.real.mean() - g["E_gs"]) < 1e-6


def test_antisymmetry():
    """test_nqs_psiformer.py:60-70 with the backflow switched on."""
    g, arch, theta, model = _pair("kdet_jastrow_env")
    lp = make_logpsi(model, theta.cuda())
    R, S = _configs(1, 5, 7.0, 5)
    R2, S2 = R.copy(), S.copy()
    R2[0, [0, 1]] = R[0, [1, 0]]
    S2[0, [0, 1]] = S[0, [1, 0]]
    ratio = np.exp(lp(R2, S2).cpu().numpy()[0] - lp(R, S).cpu().numpy()[0])
    assert abs(ratio + 1.0) < 1e-9


def test_spiral_zero_variance_with_ext_field():
    """test_nqs_spiral.py:137-149 / test_nqs_folx.py:380-391: non-collinear ext_field branch."""
    res = nsp.spiral_gs(N=7, L=8.0, Q_int=(1, 0), B=0.3, lambda_r=1.0, h_z=0.0, n_cut=10, twist=(0.1, 0.17))
    assert res["cutoff_ok"]
    lp = make_logpsi_general(res["Gs"], res["W"], 8.0)
    E = make_local_energy(lp, 1.0, 1.0, 0.0, ext_field=nsp.spiral_field(res["B"], res["Q"]))
    R, S = _configs(16, 7, 8.0, 0)
    e = E(R, S).cpu().numpy()
    assert np.std(e.real) < 1e-7
    assert abs(e.real.mean() - res["E_gs"]) < 1e-7
    assert np.max(np.abs(e.imag)) < 1e-7


def test_rect_cell_envelope_zero_variance():
    """test_nqs_production.py:22-37: general (rectangular) cell."""
    A = rect_cell(8.0, 8.0 * np.sqrt(3))
    g = bands.rashba_gs(A, 6, 1.0, 1.0, 0.5, n_cut=8, twist=(0.1, 0.17))
    lp = make_logpsi_general(g["Gs"], g["W"], A)
    rng = np.random.default_rng(0)
    R = rng.uniform(0, 1, (12, 6, 2)) @ A.T
    S = rng.integers(0, 2, (12, 6))
    e = make_local_energy(lp, 1.0, 1.0, 0.5)(R, S).cpu().numpy()
    assert np.std(e.real) < 1e-7 and abs(e.real.mean() - g["E_gs"]) < 1e-7


def test_psiformer_general_cell_matches_oracle():
    A = rect_cell(7.0, 7.0 * np.sqrt(3))
    g = bands.rashba_gs(A, 4, 1.0, 1.0, 0.3, n_cut=8, twist=(0.1, 0.17))
    kw = ARCHS["kdet_jastrow_env"]
    arch = opf.Arch(4, g["Gs"], g["W"], A, **kw)
    theta = opf.init_theta(arch, 9, backflow_scale=0.05, scale_all=0.02)
    model = ContinuumSpinorPsiformer(n_elec=4, Gs=g["Gs"], W=g["W"], L=A, **kw)
    lp = make_logpsi(model, theta.cuda())
    rng = np.random.default_rng(2)
    R = rng.uniform(0, 1, (3, 4, 2)) @ A.T
    S = rng.integers(0, 2, (3, 4))
    e = make_local_energy(lp, 1.0, 1.0, 0.3)(R, S).cpu().numpy()
    Eo = ole.make_local_energy(opf.make_logpsi(arch, theta), 1.0, 1.0, 0.3)
    er = np.array([Eo(torch.tensor(R[w]), torch.tensor(S[w])) for w in range(3)])
    assert np.max(np.abs(e - er)) < 1e-8 * max(1.0, np.max(np.abs(er)))


@pytest.mark.parametrize("N,d,heads", [(20, 32, 2), (24, 64, 2), (27, 32, 2)])
def test_attention_beyond_16_tokens(N, d, heads):
    """16 < N <= 32 tokens (the register / MFMA attention with two key blocks for head width 16,
    and for width 32 in float32): float64 log psi and E_loc against the oracle, mixed against
    float64 (log psi within 5e-4, the bound of the other mixed-precision log psi checks: one of the
    three walkers sits near a singular determinant and shows 2e-4 where the others show 1e-6)."""
    L = 2.0 * np.sqrt(np.pi * N)
    g = bands.rashba_gs(L, N, 1.0, 0.7, 0.2, n_cut=10, twist=(0.1, 0.17))
    kw = dict(embed_dim=d, depth=1, n_heads=heads, n_freq=2, n_dets=2)
    arch = opf.Arch(N, g["Gs"], g["W"], L, **kw)
    theta = opf.init_theta(arch, 5, backflow_scale=0.05, scale_all=0.02)
    R, S = _configs(3, N, L, 21)
    lp64 = make_logpsi(ContinuumSpinorPsiformer(n_elec=N, Gs=g["Gs"], W=g["W"], L=L, **kw), theta.cuda())
    v = lp64(R, S).cpu().numpy()
    ref = np.array([complex(opf.logpsi(arch, theta, torch.tensor(R[w]), torch.tensor(S[w]))) for w in range(3)])
    assert np.max(np.abs(v.real - ref.real)) < 1e-10
    e64 = make_local_energy(lp64, 1.0, 0.7, 0.2)(R[:1], S[:1]).cpu().numpy()
    Eo = ole.make_local_energy(opf.make_logpsi(arch, theta), 1.0, 0.7, 0.2)
    er = Eo(torch.tensor(R[0]), torch.tensor(S[0]))
    assert abs(e64[0] - er) < 1e-8 * max(1.0, abs(er))
    lpm = make_logpsi(ContinuumSpinorPsiformer(n_elec=N, Gs=g["Gs"], W=g["W"], L=L, precision="mixed", **kw),
                      theta.cuda())
    assert np.max(np.abs(lpm(R, S).cpu().numpy().real - v.real)) < 5e-4
    em = make_local_energy(lpm, 1.0, 0.7, 0.2)(R[:1], S[:1]).cpu().numpy()
    assert abs(em[0] - e64[0]) < 5e-3 * max(1.0, abs(e64[0]))


@pytest.mark.parametrize("N,extra", [(17, dict(n_dets=3, use_jastrow=True, learnable_env=True, n_knots=6)),
                                     (32, dict(n_dets=2, use_jastrow=True, jastrow_kind="cusp"))])
def test_register_det_head_beyond_16(N, extra):
    """The register determinant head for 16 < N <= 32 (2 x 2 blocks of 16 x 16, padded with an
    identity block at N = 17) with several determinants, a Jastrow factor and a learnable envelope:
    float64 log psi (three walkers) and E_loc (one walker, kinetic + SOC + Zeeman) against the
    oracle."""
    L = 2.0 * np.sqrt(np.pi * N)
    g = bands.rashba_gs(L, N, 1.0, 0.6, 0.3, n_cut=10, twist=(0.1, 0.17))
    kw = dict(embed_dim=16, depth=1, n_heads=2, n_freq=2, **extra)
    arch = opf.Arch(N, g["Gs"], g["W"], L, **kw)
    theta = opf.init_theta(arch, 7, backflow_scale=0.05, scale_all=0.02)
    R, S = _configs(3, N, L, 33)
    lp64 = make_logpsi(ContinuumSpinorPsiformer(n_elec=N, Gs=g["Gs"], W=g["W"], L=L, **kw), theta.cuda())
    v = lp64(R, S).cpu().numpy()
    ref = np.array([complex(opf.logpsi(arch, theta, torch.tensor(R[w]), torch.tensor(S[w]))) for w in range(3)])
    assert np.max(np.abs(v.real - ref.real)) < 1e-10
    dim = np.angle(np.exp(1j * (v.imag - ref.imag)))
    assert np.max(np.abs(dim)) < 1e-10
    e64 = make_local_energy(lp64, 1.0, 0.6, 0.3)(R[:1], S[:1]).cpu().numpy()
    Eo = ole.make_local_energy(opf.make_logpsi(arch, theta), 1.0, 0.6, 0.3)
    er = Eo(torch.tensor(R[0]), torch.tensor(S[0]))
    assert abs(e64[0] - er) < 1e-8 * max(1.0, abs(er))
