"""Exact non-interacting spin-spiral ground state (same names as src/nqs/spiral_reference.py:24-75).

Host-side set-up: the Rashba Bloch blocks plus a helical in-plane field -B(sigma_x cos Q.r +
sigma_y sin Q.r) couple |k, up> to |k+Q, down>; the lowest N eigenvectors of the plane-wave
Hamiltonian are the occupied combs.  spiral_field returns the matching ext_field as an ExtField
(Fourier series) so the local-energy kernel can evaluate it.
"""
import numpy as np

from .local_energy import ExtField
from .rashba_reference import h0_k


def spiral_gs(N, L, Q_int, B, m_star=1.0, lambda_r=1.0, h_z=0.0, n_cut=6, twist=(0.0, 0.0)):
    qx, qy = Q_int
    Qx, Qy = 2.0 * np.pi * qx / L, 2.0 * np.pi * qy / L
    n = np.arange(-n_cut, n_cut + 1)
    grid = [(a, b) for a in n for b in n]
    where = {g: i for i, g in enumerate(grid)}
    npw = len(grid)
    kx = np.array([(2 * np.pi * a + twist[0]) / L for a, _ in grid])
    ky = np.array([(2 * np.pi * b + twist[1]) / L for _, b in grid])
    H = np.zeros((2 * npw, 2 * npw), dtype=complex)
    for i in range(npw):
        H[2 * i:2 * i + 2, 2 * i:2 * i + 2] = h0_k(kx[i], ky[i], m_star, lambda_r, h_z)
    for i, (a, b) in enumerate(grid):
        j = where.get((a + qx, b + qy))
        if j is not None:
            H[2 * j + 1, 2 * i] += -B
            H[2 * i, 2 * j + 1] += -B
    E, V = np.linalg.eigh(H)
    if 2 * npw < N + 1:
        raise ValueError("basis too small for N")
    occ = V[:, :N]
    W = np.ascontiguousarray(np.transpose(occ.reshape(npw, 2, N), (2, 0, 1)))
    edge = np.array([(abs(a) == n_cut or abs(b) == n_cut) for a, b in grid])
    bw = float(np.sum(np.abs(occ.reshape(npw, 2, N)[edge]) ** 2))
    return dict(E_gs=float(np.sum(E[:N])), gap=float(E[N] - E[N - 1]), Gs=np.stack([kx, ky], -1), W=W, N=N, L=L,
                Q=(Qx, Qy), Q_int=Q_int, B=B, m_star=m_star, lambda_r=lambda_r, h_z=h_z, twist=twist,
                boundary_weight=bw, cutoff_ok=bool(bw < 1e-13))


def spiral_field(B, Q):
    """ext_field(r) = (B cos Q.r, B sin Q.r, 0)."""
    return ExtField(q=[Q], c=[[B, 0.0, 0.0]], s=[[0.0, B, 0.0]])
