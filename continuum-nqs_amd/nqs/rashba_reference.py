"""Non-interacting periodic Rashba 2DEG: exact ground state and the warm-start spinor envelope.

Host-side set-up (numpy), same names and return values as src/nqs/rashba_reference.py:36-160.
Bloch Hamiltonian (the reference's convention, rashba_reference.py:1-30):
    h0(k) = k^2/(2 m*) I + lambda_R (sigma_x k_y + sigma_y k_x) - (h_z/2) sigma_z
The N lowest spin-orbitals of the twisted momentum grid form the ground-state determinant; each
orbital is one plane wave times its band spinor (from a 2x2 eigh, as the reference does).
"""
import numpy as np

SX = np.array([[0, 1], [1, 0]], dtype=complex)
SY = np.array([[0, -1j], [1j, 0]], dtype=complex)
SZ = np.array([[1, 0], [0, -1]], dtype=complex)
I2 = np.eye(2, dtype=complex)


def h0_k(kx, ky, m_star=1.0, lambda_r=1.0, h_z=0.0):
    return (kx * kx + ky * ky) / (2.0 * m_star) * I2 + lambda_r * (ky * SX + kx * SY) - 0.5 * h_z * SZ


def bands_k(kx, ky, m_star=1.0, lambda_r=1.0, h_z=0.0):
    w, v = np.linalg.eigh(h0_k(kx, ky, m_star, lambda_r, h_z))
    return float(w[0]), float(w[1]), v[:, 0].copy(), v[:, 1].copy()


def allowed_k(L, n_cut, twist=(0.0, 0.0)):
    n = np.arange(-n_cut, n_cut + 1)
    nx, ny = np.meshgrid(n, n, indexing="ij")
    return ((2.0 * np.pi * nx + twist[0]) / L).ravel(), ((2.0 * np.pi * ny + twist[1]) / L).ravel()


def _sorted_levels(ks, m_star, lambda_r, h_z, labels=None):
    lv = []
    for q, (x, y) in enumerate(ks):
        El, Eu, cl, cu = bands_k(x, y, m_star, lambda_r, h_z)
        lab = labels[q] if labels is not None else None
        lv.append(dict(E=El, kx=x, ky=y, band=0, n=lab, chi=cl))
        lv.append(dict(E=Eu, kx=x, ky=y, band=1, n=lab, chi=cu))
    lv.sort(key=lambda t: t["E"])       # stable: ties keep (momentum-major, lower band first)
    return lv


def single_particle_levels(L, m_star=1.0, lambda_r=1.0, h_z=0.0, n_cut=8, twist=(0.0, 0.0)):
    kx, ky = allowed_k(L, n_cut, twist)
    n = np.arange(-n_cut, n_cut + 1)
    nx, ny = [a.ravel() for a in np.meshgrid(n, n, indexing="ij")]
    return _sorted_levels(list(zip(kx, ky)), m_star, lambda_r, h_z,
                          labels=[(int(a), int(b)) for a, b in zip(nx, ny)])


def noninteracting_gs(N, L, m_star=1.0, lambda_r=1.0, h_z=0.0, n_cut=8, twist=(0.0, 0.0)):
    levels = single_particle_levels(L, m_star, lambda_r, h_z, n_cut, twist)
    if len(levels) < N + 1:
        raise ValueError("n_cut too small for N levels")
    occ = levels[:N]
    k2_occ = max(d["kx"] ** 2 + d["ky"] ** 2 for d in occ)
    k2_edge = (2.0 * np.pi * n_cut / L) ** 2
    E_gs = float(sum(d["E"] for d in occ))
    return dict(E_gs=E_gs, E_per=E_gs / N, gap=float(levels[N]["E"] - levels[N - 1]["E"]), occ=occ,
                E_homo=float(occ[-1]["E"]), E_lumo=float(levels[N]["E"]), density=N / (L * L), N=N, L=L,
                m_star=m_star, lambda_r=lambda_r, h_z=h_z, twist=twist, k2_occ_max=k2_occ, k2_cut=k2_edge,
                cutoff_ok=bool(k2_occ < 0.5 * k2_edge))


def occupied_orbital_matrix(gs):
    occ = gs["occ"]
    return (np.array([[d["kx"], d["ky"]] for d in occ], dtype=float),
            np.array([d["chi"] for d in occ], dtype=complex))


def planewave_envelope(Kvecs, Chi):
    """One plane wave per orbital: Gs = Kvecs, W[a, g, :] = Chi[a] delta_{ag}."""
    Kvecs = np.asarray(Kvecs, float)
    Chi = np.asarray(Chi, complex)
    n = Kvecs.shape[0]
    W = np.zeros((n, n, 2), dtype=complex)
    W[np.arange(n), np.arange(n), :] = Chi
    return Kvecs.copy(), W
