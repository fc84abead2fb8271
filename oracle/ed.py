"""Exact diagonalisation of two electrons in the periodic 2D cell (test infrastructure: only tests
use it, as the reference for the interacting two-electron energy).

The reference has no ED code (its README quotes a 0.7 mHa agreement at N = 2 without the code or
data); this is the project's own, built from the reference's Hamiltonian:

  H = sum_i [ k_i^2/(2 m*) + lambda_R (sigma_x k_y + sigma_y k_x) - (h_z/2) sigma_z ]   (rashba_reference.py:36-39)
      + alpha phi(r_1 - r_2) + alpha zeta_M                                         (ewald.py:26-65, N = 2)

phi is the converged Ewald pair potential: its Fourier coefficients are 2 pi/(area |q|) for q != 0
and 0 for q = 0 (the real-space erfc part contributes (2 pi/(area q)) erf(q/2 eta), the reciprocal
part (2 pi/(area q)) erfc(q/2 eta), and V_G0 cancels the q = 0 term), so in the plane-wave basis
the interaction scatters a pair (k1, k2) -> (k1 + q, k2 - q) with amplitude alpha 2 pi/(area |q|),
spin unchanged.  zeta_M (the self-image energy per electron pair of images) comes from the Ewald
restatement with sums large enough to be converged to machine precision.

Basis: single-particle plane-wave spinors k = B (n + twist/2pi) (B = 2 pi A^-T) with
|n| <= n_cut (a disc of integer vectors); two-particle states of total momentum index Ntot
(n1 + n2 = Ntot), antisymmetrised.  The lowest eigenvalue converges from above as n_cut grows
(variational in the basis); `extrapolated_energy` fits E(n_cut) = E_inf + a / n_cut^p over a
sequence of cutoffs.
"""
import numpy as np

from . import bands
from . import ewald


def disc_vectors(n_cut):
    r = np.arange(-n_cut, n_cut + 1)
    v = np.array([(i, j) for i in r for j in r if i * i + j * j <= n_cut * n_cut], dtype=int)
    return v


def madelung(A, alpha=1.0):
    """zeta_M of ewald.py:52-58 with converged image and reciprocal sums."""
    A = np.asarray(A, float) if np.ndim(A) else np.array([[float(A), 0.0], [0.0, float(A)]])
    _, zeta, _ = ewald.make_ewald_lattice(A, alpha=alpha, n_real=7, n_recip=14)
    return zeta


def sector_hamiltonian(A, n_cut, Ntot=(0, 0), m_star=1.0, lambda_r=0.0, h_z=0.0, alpha=1.0, twist=(0.0, 0.0)):
    """Antisymmetrised two-electron Hamiltonian of the momentum sector Ntot (without zeta_M).
    Returns (H (D, D) complex Hermitian, description)."""
    A = np.asarray(A, float) if np.ndim(A) else np.array([[float(A), 0.0], [0.0, float(A)]])
    area = float(abs(np.linalg.det(A)))
    B = 2.0 * np.pi * np.linalg.inv(A).T
    kt = np.linalg.inv(A).T @ np.asarray(twist, float)
    vecs = disc_vectors(n_cut)
    index = {tuple(v): i for i, v in enumerate(vecs)}
    Ntot = np.asarray(Ntot, int)
    pairs = [(i, index[tuple(Ntot - v)]) for i, v in enumerate(vecs) if tuple(Ntot - v) in index]
    npair = len(pairs)
    # ordered-pair basis (p, s1, s2) -> p * 4 + 2 s1 + s2
    D0 = 4 * npair
    H = np.zeros((D0, D0), dtype=complex)
    ks = vecs @ B.T + kt
    h1 = np.array([bands.bloch_h(k[0], k[1], m_star, lambda_r, h_z) for k in ks])      # (nk, 2, 2)
    eye = np.eye(2)
    for p, (i, j) in enumerate(pairs):
        H[4 * p:4 * p + 4, 4 * p:4 * p + 4] = np.kron(h1[i], eye) + np.kron(eye, h1[j])
    # Coulomb: (n1, n2) -> (n1 + m, n2 - m), m != 0
    n1 = np.array([vecs[i] for i, _ in pairs])                                   # (npair, 2)
    m = n1[None, :, :] - n1[:, None, :]                                          # [p, p'] = n1' - n1
    q = m @ B.T
    qn = np.linalg.norm(q, axis=-1)
    with np.errstate(divide="ignore"):
        v = np.where(qn > 0, alpha * 2.0 * np.pi / (area * np.where(qn > 0, qn, 1.0)), 0.0)
    for s in range(4):
        H[s::4, s::4] += v
    # antisymmetric subspace: exchange X (p, s1, s2) -> (p~, s2, s1) with p~ = (j, i)
    pidx = {pr: p for p, pr in enumerate(pairs)}
    cols = []
    seen = set()
    for p, (i, j) in enumerate(pairs):
        for s1 in range(2):
            for s2 in range(2):
                a = 4 * p + 2 * s1 + s2
                b = 4 * pidx[(j, i)] + 2 * s2 + s1
                if a == b or (min(a, b), max(a, b)) in seen:
                    continue
                seen.add((min(a, b), max(a, b)))
                u = np.zeros(D0)
                u[a], u[b] = 1.0 / np.sqrt(2.0), -1.0 / np.sqrt(2.0)
                cols.append(u)
    U = np.array(cols).T
    HA = U.T @ H @ U
    return 0.5 * (HA + HA.conj().T), dict(npair=npair, dim=HA.shape[0], area=area)


def ground_energy(A, n_cut, sectors=((0, 0),), m_star=1.0, lambda_r=0.0, h_z=0.0, alpha=1.0, twist=(0.0, 0.0)):
    """Lowest two-electron energy over the given momentum sectors, including alpha zeta_M."""
    A = np.asarray(A, float) if np.ndim(A) else np.array([[float(A), 0.0], [0.0, float(A)]])
    best = np.inf
    for Nt in sectors:
        HA, _ = sector_hamiltonian(A, n_cut, Nt, m_star, lambda_r, h_z, alpha, twist)
        e = float(np.linalg.eigvalsh(HA)[0])
        best = min(best, e)
    zeta = madelung(A, alpha) if alpha != 0.0 else 0.0
    return best + zeta


def extrapolated_energy(A, cuts, power=3.0, **kw):
    """E(n_cut) for each cutoff and the fit E_inf + a / n_cut^power on the last three."""
    Es = np.array([ground_energy(A, c, **kw) for c in cuts])
    x = 1.0 / np.asarray(cuts, float) ** power
    coef = np.polyfit(x[-3:], Es[-3:], 1)
    return Es, float(coef[1])
