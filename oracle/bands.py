"""Oracle single-particle references (test infrastructure).

Restates src/nqs/rashba_reference.py:36-160 (Rashba 2DEG on a torus: Bloch Hamiltonian, bands,
lowest-N occupation, plane-wave spinor envelope) and src/nqs/spiral_reference.py:24-75 (exact
spin-spiral ground state in a plane-wave basis, its ext_field).

Conventions kept from the reference (rashba_reference.py:1-30):
  h0(k) = k^2/(2 m*) I + lambda_R (sigma_x k_y + sigma_y k_x) - (h_z/2) sigma_z
  E_-+(k) = k^2/(2m*) -+ sqrt(lambda_R^2 k^2 + h_z^2/4)
"""
import numpy as np

SX = np.array([[0, 1], [1, 0]], dtype=complex)
SY = np.array([[0, -1j], [1j, 0]], dtype=complex)
SZ = np.array([[1, 0], [0, -1]], dtype=complex)


def bloch_h(kx, ky, m_star=1.0, lambda_r=1.0, h_z=0.0):
    """rashba_reference.py:36-39."""
    return ((kx * kx + ky * ky) / (2.0 * m_star)) * np.eye(2) + lambda_r * (SX * ky + SY * kx) - 0.5 * h_z * SZ


def band_energies(kx, ky, m_star=1.0, lambda_r=1.0, h_z=0.0):
    """Closed form (rashba_reference.py:49-54)."""
    k2 = kx * kx + ky * ky
    d = np.sqrt(lambda_r ** 2 * k2 + 0.25 * h_z ** 2)
    return k2 / (2.0 * m_star) - d, k2 / (2.0 * m_star) + d


def _levels(Ks, m_star, lambda_r, h_z):
    """All spin-orbital levels for a list of momenta, ordered as the reference enumerates them
    (momentum-major, lower band first) and then stably sorted by energy
    (rashba_reference.py:68-80, 146-153)."""
    out = []
    for k in Ks:
        w, v = np.linalg.eigh(bloch_h(k[0], k[1], m_star, lambda_r, h_z))
        out.append((float(w[0]), np.array(k, float), v[:, 0].copy(), 0))
        out.append((float(w[1]), np.array(k, float), v[:, 1].copy(), 1))
    out.sort(key=lambda t: t[0])
    return out


def rashba_gs(A, N, m_star=1.0, lambda_r=1.0, h_z=0.0, n_cut=8, twist=(0.0, 0.0)):
    """Exact non-interacting ground state on a general cell (rashba_reference.py:136-160; the
    square case A = L*I reproduces noninteracting_gs, :83-108, whose momenta are
    (2 pi n + twist)/L).  Returns dict(E_gs, gap, Gs (N,2), W (N,N,2), occ_bands)."""
    A = np.array([[float(A), 0.0], [0.0, float(A)]]) if np.ndim(A) == 0 else np.asarray(A, float)
    Ainv = np.linalg.inv(A)
    B = 2.0 * np.pi * Ainv.T
    kt = Ainv.T @ np.asarray(twist, float)
    js = np.arange(-n_cut, n_cut + 1)
    Ks = [B @ np.array([float(nx), float(ny)]) + kt for nx in js for ny in js]
    lv = _levels(Ks, m_star, lambda_r, h_z)
    if len(lv) < N + 1:
        raise ValueError("n_cut too small for N levels")
    occ = lv[:N]
    Gs = np.array([t[1] for t in occ])
    W = np.zeros((N, N, 2), dtype=complex)
    for a, t in enumerate(occ):
        W[a, a, :] = t[2]
    # the energy from the closed-form bands (independent of eigh)
    E_closed = 0.0
    for t in occ:
        lo, hi = band_energies(t[1][0], t[1][1], m_star, lambda_r, h_z)
        E_closed += lo if t[3] == 0 else hi
    return dict(E_gs=float(sum(t[0] for t in occ)), E_closed=float(E_closed),
                gap=float(lv[N][0] - lv[N - 1][0]), Gs=Gs, W=W,
                bands=[t[3] for t in occ], A=A, N=N)


def spiral_gs(N, L, Q_int, B, m_star=1.0, lambda_r=1.0, h_z=0.0, n_cut=6, twist=(0.0, 0.0)):
    """Exact spiral ground state (spiral_reference.py:24-65): plane-wave basis |k, s>, Bloch
    blocks on the diagonal, -B coupling (k, up) <-> (k+Q, down)."""
    qx, qy = Q_int
    Qx, Qy = 2.0 * np.pi * qx / L, 2.0 * np.pi * qy / L
    js = np.arange(-n_cut, n_cut + 1)
    grid = [(nx, ny) for nx in js for ny in js]
    pos = {nm: p for p, nm in enumerate(grid)}
    npw = len(grid)
    kx = np.array([(2 * np.pi * nx + twist[0]) / L for (nx, ny) in grid])
    ky = np.array([(2 * np.pi * ny + twist[1]) / L for (nx, ny) in grid])
    H = np.zeros((2 * npw, 2 * npw), dtype=complex)
    for p in range(npw):
        H[2 * p:2 * p + 2, 2 * p:2 * p + 2] = bloch_h(kx[p], ky[p], m_star, lambda_r, h_z)
    for p, (nx, ny) in enumerate(grid):
        p2 = pos.get((nx + qx, ny + qy))
        if p2 is not None:
            H[2 * p2 + 1, 2 * p] += -B
            H[2 * p, 2 * p2 + 1] += -B
    E, V = np.linalg.eigh(H)
    occ = V[:, :N]
    W = np.transpose(occ.reshape(npw, 2, N), (2, 0, 1)).copy()
    bmask = np.array([(abs(nx) == n_cut or abs(ny) == n_cut) for (nx, ny) in grid])
    bweight = float(np.sum(np.abs(occ.reshape(npw, 2, N)[bmask]) ** 2))
    return dict(E_gs=float(np.sum(E[:N])), gap=float(E[N] - E[N - 1]),
                Gs=np.stack([kx, ky], -1), W=W, N=N, L=L, Q=(Qx, Qy), B=B,
                lambda_r=lambda_r, h_z=h_z, boundary_weight=bweight)


def spiral_field(B, Q):
    """spiral_reference.py:68-75 -- ext_field(r) = (B cos Q.r, B sin Q.r, 0); works on torch
    tensors r (..., 2)."""
    import torch

    def f(r):
        ph = Q[0] * r[..., 0] + Q[1] * r[..., 1]
        return torch.stack([B * torch.cos(ph), B * torch.sin(ph), torch.zeros_like(ph)], -1)
    return f
