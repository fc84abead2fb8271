"""Oracle 2D Ewald Coulomb (test infrastructure): src/nqs/ewald.py:26-73 restated in numpy.

V(r) = alpha [ sum_{|n|<=n_real box} erfc(eta |r + A n|)/|r + A n|
               + sum_{G != 0, |m|<=n_recip box} (2 pi/area) erfc(|G|/2 eta)/|G| cos(G.r) + V_G0 ]
V_G0 = -2 sqrt(pi) / (area eta);   eta = sqrt(pi)/sqrt(area) unless given
zeta_M = alpha (sum_{R != 0} erfc(eta R)/R - 2 eta/sqrt(pi) + sum_G R_G + V_G0)
E(R) = sum_{i<j} V(r_i - r_j) (raw differences, no minimum image) + N zeta_M / 2
"""
import math
import numpy as np
from scipy.special import erfc


def make_ewald_lattice(A, alpha=1.0, ewald_alpha=None, n_real=5, n_recip=10):
    A = np.asarray(A, dtype=float)
    area = float(abs(np.linalg.det(A)))
    B = 2.0 * np.pi * np.linalg.inv(A).T
    eta = math.sqrt(math.pi) / math.sqrt(area) if ewald_alpha is None else float(ewald_alpha)
    n = np.arange(-n_real, n_real + 1)
    cells = np.array([(i, j) for i in n for j in n], dtype=float)
    R_lat = cells @ A.T
    m = np.arange(-n_recip, n_recip + 1)
    mc = np.array([(i, j) for i in m for j in m if (i, j) != (0, 0)], dtype=float)
    G = mc @ B.T
    Gm = np.linalg.norm(G, axis=-1)
    RG = (2.0 * np.pi / area) * erfc(Gm / (2.0 * eta)) / Gm
    VG0 = -2.0 * math.sqrt(math.pi) / (area * eta)

    def ewald_pair(r):
        r = np.asarray(r, float)
        d = np.linalg.norm(r[None, :] + R_lat, axis=-1)
        return alpha * (np.sum(erfc(eta * d) / d) + np.sum(RG * np.cos(G @ r)) + VG0)

    Rs = np.linalg.norm(R_lat, axis=-1)
    real_nz = float(sum(erfc(eta * d) / d for d in Rs if d > 1e-10))
    zeta_M = float(alpha * (real_nz - 2.0 * eta / math.sqrt(math.pi) + float(np.sum(RG)) + VG0))

    def coulomb_energy(R):
        R = np.asarray(R, float)
        N = R.shape[0]
        tot = 0.0
        for i in range(N):
            for j in range(i + 1, N):
                tot += ewald_pair(R[i] - R[j])
        return tot + N * zeta_M / 2.0

    coulomb_energy.table = dict(R_lat=R_lat, G=G, RG=RG, VG0=VG0, eta=eta, alpha=alpha, zeta_M=zeta_M)
    return ewald_pair, zeta_M, coulomb_energy


def make_ewald(L, alpha=1.0, ewald_alpha=None, n_real=4, n_recip=8):
    """ewald.py:70-73 -- square special case."""
    return make_ewald_lattice(np.array([[float(L), 0.0], [0.0, float(L)]]), alpha=alpha,
                              ewald_alpha=ewald_alpha, n_real=n_real, n_recip=n_recip)
