"""Host-side (numpy) parts of the observables, overlap and rotation modules, no GPU: the FHS
assembly on the QWZ model's known Chern numbers under random gauges (test_manybody_chern.py:53-80),
the vectorised Berg-Luscher sum against the oracle's plaquette loops, and the exact Slater spin
density of a spin-polarised and an x-polarised determinant."""
import numpy as np
import pytest

from oracle import observables as oob
from nqs import manybody_chern as nmc, observables as nob


def _qwz_u(kx, ky, mgap):
    d = np.array([np.sin(kx), np.sin(ky), mgap + np.cos(kx) + np.cos(ky)])
    nrm = np.linalg.norm(d)
    u = np.array([d[0] - 1j * d[1], -(d[2] + nrm)], complex)
    if np.linalg.norm(u) < 1e-12:
        u = np.array([-(d[2] - nrm), d[0] + 1j * d[1]], complex)
    return u / np.linalg.norm(u)


@pytest.mark.parametrize("mgap,C_expect", [(-1.0, -1), (-3.0, 0)])
def test_fhs_on_qwz_known_chern(mgap, C_expect):
    nt = 12
    th = np.linspace(0, 2 * np.pi, nt, endpoint=False)
    rng = np.random.default_rng(3)
    U = np.array([[_qwz_u(a, b, mgap) * np.exp(1j * rng.uniform(0, 6.3)) for b in th] for a in th])
    lx = np.array([[np.vdot(U[i, j], U[(i + 1) % nt, j]) for j in range(nt)] for i in range(nt)])
    ly = np.array([[np.vdot(U[i, j], U[i, (j + 1) % nt]) for j in range(nt)] for i in range(nt)])
    C, _ = nmc.fhs_chern(lx / np.abs(lx), ly / np.abs(ly))
    assert abs(C - C_expect) < 1e-9
    rep = nmc.chern_report(lx / np.abs(lx), ly / np.abs(ly), np.abs(lx), np.abs(ly))
    assert rep["C_int"] == C_expect




def test_berg_luscher_vectorised_equals_loops():
    rng = np.random.default_rng(8)
    m = rng.normal(size=(9, 9, 3))
    assert abs(nob.berg_luscher(m) - oob.berg_luscher(m)) < 1e-12


def test_slater_spin_density_closed_forms():
    L = 5.0
    B = 2 * np.pi / L
    Gs = np.array([(0.0, 0.0), (B, 0.0), (0.0, B)])
    f = np.eye(3, dtype=complex) / np.sqrt(L * L)
    up = nob.slater_spin_density(Gs, np.stack([f, 0 * f], -1), np.diag([L, L]), ngrid=6)
    assert np.max(np.abs(up[..., 2] - 3 / (L * L))) < 1e-15 and np.max(np.abs(up[..., :2])) < 1e-15
    xp = nob.slater_spin_density(Gs, np.stack([f, f], -1) / np.sqrt(2), np.diag([L, L]), ngrid=6)
    assert np.max(np.abs(xp[..., 0] - 3 / (L * L))) < 1e-15 and np.max(np.abs(xp[..., 1:])) < 1e-15
