"""Spin-resolved observables on the HIP path (call surface of src/nqs/observables.py:31-220).

Per-walker work runs in the library: the spin estimators (nqs_spin_estimators: log psi of every
walker and of its N single-spin-flipped configurations, then sigma_x = ratio_i,
sigma_y = -i (1 - 2 s_i) ratio_i, sigma_z = 1 - 2 s_i), the spin Fourier sums behind the structure
factor and the coherent order parameter (nqs_spin_fourier), and the texture bin sums
(nqs_spin_texture).  What remains on the host is grid-sized: the Berg-Luscher charge of an
ngrid x ngrid texture, the Fourier reconstruction of a texture from a few m(G), the exact Slater
spin density and the classification heuristic, in numpy as the reference has them.

`logpsi` is a LogPsi (nqs.psiformer.make_logpsi or nqs.slater_reference); R (W, N, 2), S (W, N).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .psiformer import LogPsi, as_walkers


def _check_logpsi(logpsi):
    if not isinstance(logpsi, LogPsi):
        raise TypeError("observables need a LogPsi (nqs.psiformer.make_logpsi / nqs.slater_reference)")


def make_spin_estimators(logpsi):
    """Return f(R, S) -> (sx, sy, sz), each (W, N) complex device tensors (observables.py:31-47;
    batched over walkers instead of per configuration)."""
    _check_logpsi(logpsi)

    def estimators(R, S):
        sp = _spin_tensor(logpsi, R, S)
        c = torch.view_as_complex(sp)
        return c[..., 0], c[..., 1], c[..., 2]
    return estimators


def _spin_tensor(logpsi, R, S):
    """(W, N, 3, 2) float64: the complex estimators (x, y, z) of every electron."""
    lib = _lib.load()
    R, S, _ = as_walkers(R, S)
    W, N = R.shape[0], R.shape[1]
    d = logpsi.model.desc()
    out = torch.empty(W, N, 3, 2, dtype=torch.float64, device=R.device)
    nb = lib.nqs_workspace_bytes(ctypes.byref(d), _lib.OP_SPIN, W)
    ws = _lib.workspace(nb)
    _lib.check(lib.nqs_spin_estimators(ctypes.byref(d), _lib.ptr(logpsi.packed), _lib.ptr(R), _lib.ptr(S), W,
                                       _lib.ptr(out), _lib.ptr(ws), ws.numel(), _lib.stream()),
               "nqs_spin_estimators")
    return out


def _real_spins(logpsi, R, S):
    """Device (W, N, 3) real parts of the estimators, plus the device walkers."""
    R, S, _ = as_walkers(R, S)
    return _spin_tensor(logpsi, R, S)[..., 0].contiguous(), R, S


def spin_estimates(logpsi, R_batch, S_batch):
    """(sx, sy, sz) real parts, each (W, N) numpy (observables.py:50-53)."""
    _check_logpsi(logpsi)
    sp, _, _ = _real_spins(logpsi, R_batch, S_batch)
    sp = sp.cpu().numpy()
    return sp[..., 0], sp[..., 1], sp[..., 2]


def _texture(sp, R, A, ngrid):
    lib = _lib.load()
    W, N = R.shape[0], R.shape[1]
    Ai = np.linalg.inv(np.asarray(A, float))
    ai = (ctypes.c_double * 4)(*[float(x) for x in Ai.reshape(-1)])
    m = torch.empty(ngrid, ngrid, 3, dtype=torch.float64, device=R.device)
    rho = torch.empty(ngrid, ngrid, dtype=torch.float64, device=R.device)
    _lib.check(lib.nqs_spin_texture(_lib.ptr(R), _lib.ptr(sp), W, N, ai, int(ngrid), _lib.ptr(m), _lib.ptr(rho),
                                    _lib.stream()), "nqs_spin_texture")
    m = m.cpu().numpy() / W
    rho = rho.cpu().numpy() / W
    return dict(m=m, rho=rho, mbar=m.reshape(-1, 3).sum(0) / N, mabs=float(np.mean(np.linalg.norm(m, axis=-1))))


def spin_texture(logpsi, R_batch, S_batch, A, ngrid=12):
    """Magnetisation density on an ngrid x ngrid grid over the cell A (observables.py:56-76):
    dict(m (ngrid, ngrid, 3), rho (ngrid, ngrid), mbar (3,) net moment per electron, mabs)."""
    _check_logpsi(logpsi)
    sp, R, _ = _real_spins(logpsi, R_batch, S_batch)
    return _texture(sp, R, A, ngrid)


def berg_luscher(m):
    """Skyrmion number of a periodic texture m (ngrid, ngrid, 3) from lattice solid angles
    (observables.py:79-91), vectorised over the plaquettes in the same order of terms."""
    m = np.asarray(m, float)
    n = m / (np.linalg.norm(m, axis=-1, keepdims=True) + 1e-12)
    a = n
    b = np.roll(n, -1, axis=0)
    c = np.roll(n, -1, axis=1)
    d = np.roll(b, -1, axis=1)

    def omega(p, q, w):
        num = np.einsum("...i,...i->...", p, np.cross(q, w))
        den = 1.0 + (p * q).sum(-1) + (q * w).sum(-1) + (w * p).sum(-1)
        return 2.0 * np.arctan2(num, den)
    return float((omega(a, b, d) + omega(a, d, c)).sum() / (4.0 * np.pi))


def _fourier(sp, R, qs):
    lib = _lib.load()
    W, N = R.shape[0], R.shape[1]
    q = torch.tensor(np.ascontiguousarray(qs, dtype=float).reshape(-1, 2), device=R.device)
    out = torch.empty(q.shape[0], 8, dtype=torch.float64, device=R.device)
    _lib.check(lib.nqs_spin_fourier(_lib.ptr(R), _lib.ptr(sp), W, N, _lib.ptr(q), q.shape[0], _lib.ptr(out),
                                    _lib.stream()), "nqs_spin_fourier")
    o = out.cpu().numpy()
    mq = np.stack([o[:, 0] + 1j * o[:, 1], o[:, 2] + 1j * o[:, 3], o[:, 4] + 1j * o[:, 5]])   # (3, Q)
    return mq, o[:, 6], o[:, 7]


def _recip_grid(A, nshell):
    Bmat = 2.0 * np.pi * np.linalg.inv(np.asarray(A, float)).T
    mn = [(p, q) for p in range(-nshell, nshell + 1) for q in range(-nshell, nshell + 1)]
    qs = np.array([p * Bmat[:, 0] + q * Bmat[:, 1] for (p, q) in mn])
    return mn, qs


def structure_factor(logpsi, R_batch, S_batch, A, nshell=4):
    """S_zz(q) and S_perp(q) on the reciprocal grid m b1 + n b2, |m|, |n| <= nshell
    (observables.py:94-109): dict(q, qfrac, Szz, Sperp)."""
    _check_logpsi(logpsi)
    sp, R, _ = _real_spins(logpsi, R_batch, S_batch)
    mn, qs = _recip_grid(A, nshell)
    _, szz, sperp = _fourier(sp, R, qs)
    return dict(q=qs, qfrac=mn, Szz=szz, Sperp=sperp)


def spin_order_parameter(logpsi, R_batch, S_batch, A, Gset):
    """m_a(G) = (1/N) <sum_i sigma_a^i exp(-i G.r_i)> for each G (observables.py:112-121): (3, nG)."""
    _check_logpsi(logpsi)
    sp, R, _ = _real_spins(logpsi, R_batch, S_batch)
    mq, _, _ = _fourier(sp, R, np.asarray(Gset, float))
    return mq


def skx_arms(q1, q2, nharm=1):
    """{0} and +-h q_j for the arms q1, q2, q3 = -(q1 + q2), h = 1..nharm (observables.py:124-131)."""
    q1 = np.asarray(q1, float)
    q2 = np.asarray(q2, float)
    q3 = -(q1 + q2)
    Gs = [np.zeros(2)]
    for h in range(1, nharm + 1):
        for q in (q1, q2, q3):
            Gs += [h * q, -h * q]
    return np.array(Gs)


def _reconstruct(mG, G, A, ngrid):
    fr = (np.arange(ngrid) + 0.5) / ngrid
    u, v = np.meshgrid(fr, fr, indexing="ij")
    r = np.stack([u, v], -1) @ np.asarray(A, float).T                  # (ng, ng, 2)
    return np.real(np.exp(1j * (r @ G.T)) @ mG.T)                       # (ng, ng, 3)


def skyrmion_number_fourier(logpsi, R_batch, S_batch, A, q1, q2, nharm=1, ngrid=24):
    """Skyrmion number of the texture rebuilt from the coherent m(G) at the skyrmion arms
    (observables.py:134-151): dict(Q, order_amp, m_recon, mG, m0)."""
    G = skx_arms(q1, q2, nharm)
    mG = spin_order_parameter(logpsi, R_batch, S_batch, A, G)
    m = _reconstruct(mG, G, A, ngrid)
    arm_amp = float(np.mean(np.linalg.norm(mG[:, 1:], axis=0)))
    return dict(Q=float(berg_luscher(m)), order_amp=arm_amp, m_recon=m, mG=mG, m0=np.real(mG[:, 0]))


def slater_spin_density(Gs, W, A, ngrid=24):
    """Exact m(r) = sum_occ phi_a^dagger sigma phi_a of the determinant with orbitals
    phi_a(r, s) = sum_G W[a, G, s] exp(i G.r) (observables.py:154-169): (ngrid, ngrid, 3)."""
    Gs = np.asarray(Gs, float)
    W = np.asarray(W, complex)
    fr = (np.arange(ngrid) + 0.5) / ngrid
    u, v = np.meshgrid(fr, fr, indexing="ij")
    r = np.stack([u, v], -1) @ np.asarray(A, float).T
    ph = np.exp(1j * (r @ Gs.T))                                        # (ng, ng, nG)
    phi = np.einsum("xyg,ags->xyas", ph, W)                             # (ng, ng, N, 2)
    up, dn = phi[..., 0], phi[..., 1]
    cud = np.sum(np.conj(up) * dn, axis=-1)
    mz = np.sum(np.abs(up) ** 2 - np.abs(dn) ** 2, axis=-1)
    return np.stack([2 * np.real(cud), 2 * np.imag(cud), mz], -1)
