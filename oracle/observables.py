"""Oracle spin observables and ratio estimators (test infrastructure).

Restates, for one configuration at a time or in plain numpy:
  spin_estimators  src/nqs/observables.py:31-47 (flip ratios of a logpsi(R, S) callable)
  texture_sums     observables.py:56-76 (bin sums before the division by the walker count)
  spin_fourier     observables.py:94-121 (rho_a(q) per walker, m_a(q), S_zz, S_perp)
  berg_luscher     observables.py:79-91 (explicit plaquette loops)
  ratio_stats      manybody_chern.py:23-34 and topology.py:114-122 (mean, mean |r|^2, ess, std)
  rotation_overlap_planewaves  the exact <R_n> of a single-plane-wave spinor Slater determinant, the
                   closed form topology.py:125-150 computes, derived here from the orbital overlaps
Pins (tests/test_oracle.py): a spin-independent trial is exactly +x polarised (m_x = 1, m_y = 0);
uncorrelated spins give a flat S_zz; a triple-Q skyrmion texture has Berg-Luscher charge 1 and a
uniform one 0; the overlap of two plane-wave determinants is det(Wa^* Wb^T).
"""
import numpy as np
import torch


def spin_estimators(logpsi, R, S):
    """(sx, sy, sz) complex (N,) of one configuration; non-finite ratios -> 0."""
    R = torch.as_tensor(R, dtype=torch.float64)
    S = torch.as_tensor(S).long()
    N = R.shape[0]
    L0 = complex(logpsi(R, S))
    ratio = np.zeros(N, complex)
    for i in range(N):
        Sf = S.clone()
        Sf[i] = 1 - Sf[i]
        with np.errstate(all="ignore"):
            r = np.exp(complex(logpsi(R, Sf)) - L0)
        ratio[i] = r if np.isfinite(r) else 0.0
    sz = 1.0 - 2.0 * S.numpy().astype(float)
    return ratio, -1j * sz * ratio, sz.astype(complex)


def texture_sums(R, spins, A, ngrid):
    """m (ngrid, ngrid, 3) = sum of spins per bin, rho (ngrid, ngrid) = counts; spins (W, N, 3) real."""
    Ai = np.linalg.inv(np.asarray(A, float))
    R = np.asarray(R, float)
    fx = np.mod(R[..., 0] * Ai[0, 0] + R[..., 1] * Ai[0, 1], 1.0)
    fy = np.mod(R[..., 0] * Ai[1, 0] + R[..., 1] * Ai[1, 1], 1.0)
    ix = np.clip((fx * ngrid).astype(int), 0, ngrid - 1).ravel()
    iy = np.clip((fy * ngrid).astype(int), 0, ngrid - 1).ravel()
    m = np.zeros((ngrid, ngrid, 3))
    rho = np.zeros((ngrid, ngrid))
    sp = np.asarray(spins, float).reshape(-1, 3)
    for c in range(3):
        np.add.at(m[..., c], (ix, iy), sp[:, c])
    np.add.at(rho, (ix, iy), 1.0)
    return m, rho


def spin_fourier(R, spins, qs):
    """m_a(q) (3, Q) complex, S_zz (Q,), S_perp (Q,) over walkers R (W, N, 2), spins (W, N, 3)."""
    R = np.asarray(R, float)
    sp = np.asarray(spins, float)
    W, N = R.shape[:2]
    ph = np.exp(-1j * np.einsum("wnd,qd->wnq", R, np.asarray(qs, float)))
    rho = np.einsum("wna,wnq->waq", sp, ph)                               # (W, 3, Q)
    m = rho.mean(0) / N
    szz = np.mean(np.abs(rho[:, 2]) ** 2, axis=0) / N
    sperp = np.mean(np.abs(rho[:, 0]) ** 2 + np.abs(rho[:, 1]) ** 2, axis=0) / N
    return m, szz, sperp


def berg_luscher(m):
    ng = m.shape[0]
    n = m / (np.linalg.norm(m, axis=-1, keepdims=True) + 1e-12)
    tot = 0.0
    for i in range(ng):
        for j in range(ng):
            a, b = n[i, j], n[(i + 1) % ng, j]
            c, d = n[i, (j + 1) % ng], n[(i + 1) % ng, (j + 1) % ng]
            for p, q, w in ((a, b, d), (a, d, c)):
                num = np.dot(p, np.cross(q, w))
                den = 1.0 + np.dot(p, q) + np.dot(q, w) + np.dot(w, p)
                tot += 2.0 * np.arctan2(num, den)
    return tot / (4.0 * np.pi)


def ratio_stats(la, lb, phase=None, shift=True):
    """[mean Re r, mean Im r, mean |r|^2, ess, shift, std r], r = phase exp(lb - la - shift)."""
    d = np.asarray(lb, complex) - np.asarray(la, complex)
    s = float(d.real.max()) if shift else 0.0
    r = np.exp(d - s)
    if phase is not None:
        r = np.asarray(phase, complex) * r
    a = np.abs(r)
    return np.array([r.real.mean(), r.imag.mean(), (a ** 2).mean(), a.sum() ** 2 / (a ** 2).sum() / len(r), s,
                     np.std(r)])


def rotation_overlap_planewaves(Gs, Chi, n):
    """<Psi| R_n |Psi> for the Slater determinant of phi_a(r, s) = exp(i G_a.r) Chi[a, s] (orthonormal
    orbitals spanning an R_n-invariant space).  R_n maps phi_b to the plane wave at Rot(2 pi/n) G_b
    with the spinor U Chi[b], U = diag(exp(-i pi/n), exp(+i pi/n)) (spin s = 0 first), so the orbital
    overlap matrix is m[a, b] = [G_a == Rot G_b] Chi[a]^* . U Chi[b] and <R_n> = det m."""
    Gs = np.asarray(Gs, float)
    Chi = np.asarray(Chi, complex)
    t = 2.0 * np.pi / n
    rot = Gs @ np.array([[np.cos(t), np.sin(t)], [-np.sin(t), np.cos(t)]])      # row b = Rot(t) G_b
    same = np.linalg.norm(Gs[:, None, :] - rot[None, :, :], axis=-1) < 1e-8 * (1.0 + np.abs(Gs).max())
    spin = np.conj(Chi) @ (np.array([np.exp(-1j * t / 2), np.exp(1j * t / 2)])[:, None] * Chi.T)
    return complex(np.linalg.det(np.where(same, spin, 0.0)))
