"""Pinning on the GPU E_loc path: every texture factory of nqs.pinning (the reference's pinning
targets, src/nqs/pinning.py:20-119, coupled as ext_field) and Gaussian pinning centres (scalar
and spin, no reference counterpart) against the oracle's dense-autodiff E_loc with the oracle's
own restatement of the same field, at the same walkers (1e-8 relative)."""
import numpy as np
import pytest
import torch

from oracle import bands, psiformer as opf, local_energy as ole, pinning as opn
from nqs import pinning as npn
from nqs.psiformer import ContinuumSpinorPsiformer, make_logpsi
from nqs.local_energy import make_local_energy, pin_table

pytestmark = pytest.mark.gpu

N, L = 5, 7.0
KW = dict(embed_dim=16, depth=2, n_heads=4, n_freq=4, n_dets=2)
Q1 = np.array([2 * np.pi / L, 0.0])
Q2 = np.array([-np.pi / L, np.sqrt(3) * np.pi / L])
Q = np.array([0.4, -0.3])
KEYS = [(0, 0), (1, 0), (0, 1), (-1, -1)]
VALS = [(0.1, 0.0, 0.5), (0.2j, -0.1, 0.3 - 0.1j), (0.05, 0.4j, 1.0), (-0.3j, 0.1 + 0.2j, 0.7)]
TEXTURES = {
    "skx": lambda m: m.make_pin("skx", Q1, Q2, 0.3),
    "skx_bloch": lambda m: m.make_pin("skx_bloch", Q1, Q2, 0.3, chirality=-1.0),
    "fm": lambda m: m.make_pin("fm", None, None, 0.2, axis=(1.0, -2.0, 0.5)),
    "spiral": lambda m: m.make_pin("spiral", Q, None, 0.2),
    "spiral_xy": lambda m: m.make_pin("spiral_xy", Q, None, 0.2),
    "overhauser": lambda m: m.make_pin("overhauser", Q, None, 0.2),
    "hf_skx": lambda m: m.make_hf_skx_pin(KEYS, VALS, Q1, Q2, 0.25),
}


@pytest.fixture(scope="module")
def sysm():
    g = bands.rashba_gs(L, N, 1.0, 1.0, 0.5, n_cut=10, twist=(0.1, 0.17))
    arch = opf.Arch(N, g["Gs"], g["W"], L, **KW)
    theta = opf.init_theta(arch, 3, backflow_scale=0.05, scale_all=0.02)
    model = ContinuumSpinorPsiformer(n_elec=N, Gs=g["Gs"], W=g["W"], L=L, **KW)
    rng = np.random.default_rng(7)
    R, S = rng.uniform(0, L, (3, N, 2)), rng.integers(0, 2, (3, N))
    return dict(arch=arch, theta=theta, lp=make_logpsi(model, theta.cuda()), R=R, S=S)


def _check(sysm, gpu_kw, oracle_kw):
    e = make_local_energy(sysm["lp"], 1.0, 1.0, 0.5, **gpu_kw)(sysm["R"], sysm["S"]).cpu().numpy()
    Eo = ole.make_local_energy(opf.make_logpsi(sysm["arch"], sysm["theta"]), 1.0, 1.0, 0.5, **oracle_kw)
    er = np.array([Eo(torch.tensor(sysm["R"][w]), torch.tensor(sysm["S"][w])) for w in range(len(e))])
    assert np.max(np.abs(e - er)) < 1e-8 * max(1.0, np.max(np.abs(er)))
    return e


@pytest.mark.parametrize("name", list(TEXTURES))
def test_texture_eloc_matches_oracle(sysm, name):
    e = _check(sysm, dict(ext_field=TEXTURES[name](npn)), dict(ext_field=TEXTURES[name](opn)))
    e0 = make_local_energy(sysm["lp"], 1.0, 1.0, 0.5)(sysm["R"], sysm["S"]).cpu().numpy()
    assert np.max(np.abs(e - e0)) > 1e-3          # the field changes E_loc


def test_gaussian_pins_eloc_matches_oracle(sysm):
    pins = [npn.GaussianPin((1.0, 2.0), 1.3, 0.4, (0.1, -0.2, 0.3)),
            npn.GaussianPin((5.5, 6.0), 2.0, -0.2, (0.0, 0.15, 0.0)),
            npn.GaussianPin((3.0, 0.5), 0.8, 0.0, (0.0, 0.0, -0.25))]
    rows = pin_table(pins)[:, :7]
    _check(sysm, dict(pinning=pins), dict(pinning=rows, cell=np.diag([L, L])))
    # texture and pins together
    _check(sysm, dict(ext_field=TEXTURES["skx"](npn), pinning=pins),
           dict(ext_field=TEXTURES["skx"](opn), pinning=rows, cell=np.diag([L, L])))
