"""C_n rotation eigenvalue of a many-body state on the HIP path (call surface of
src/nqs/topology.py:84-150).

[R_n psi](r_1 s_1, ...) = prod_i exp(-i theta (1 - 2 s_i)/2) psi(R_n^-1 r_1, s_1; ...) with
theta = 2 pi/n, so <R_n> = < phase(S) psi(R_n^-1 R, S)/psi(R, S) >_{|psi|^2}: two batched log psi
evaluations (nqs_psiformer_fwd) and one nqs_ratio_stats reduction (mean and standard error).
The exact single-plane-wave Slater value det(m) and the position map are host numpy.
"""
import numpy as np
import torch

from . import _lib
from .psiformer import LogPsi, as_walkers
from .manybody_chern import ratio_stats


def _rot2d(theta):
    c, s = np.cos(theta), np.sin(theta)
    return np.array([[c, -s], [s, c]])


def rotate_positions(R, n, L, center=0.0):
    """R_n^-1 r = Rot(-2 pi/n)(r - center) + center, wrapped to [0, L)^2 (topology.py:100-105)."""
    th = 2.0 * np.pi / n
    Rm = _rot2d(-th)
    c = np.asarray(center, float) if np.ndim(center) else np.array([center, center], float)
    return np.mod((np.asarray(R, float) - c) @ Rm.T + c, L)


def spin_rotation_phase(S, n):
    """prod_i exp(-i (2 pi/n)(1 - 2 s_i)/2) per configuration (topology.py:108-111)."""
    th = 2.0 * np.pi / n
    return np.exp(-1j * th * np.sum(1.0 - 2.0 * np.asarray(S), axis=-1) / 2.0)


def rotation_eigenvalue_vmc(logpsi, R, S, n, L, center=0.0):
    """(<R_n>, standard error) over walkers R (W, N, 2), S (W, N) (topology.py:114-122);
    logpsi is a batched LogPsi."""
    if not isinstance(logpsi, LogPsi):
        raise TypeError("rotation_eigenvalue_vmc needs a LogPsi")
    R = R.cpu().numpy() if torch.is_tensor(R) else np.asarray(R, float)
    S = S.cpu().numpy() if torch.is_tensor(S) else np.asarray(S)
    Rrot = rotate_positions(R, n, L, center)
    sp = spin_rotation_phase(S, n)
    Rd, Sd, _ = as_walkers(R, S)
    Rr, _, _ = as_walkers(Rrot, S)
    st = ratio_stats(logpsi(Rd, Sd), logpsi(Rr, Sd), phase=torch.tensor(sp, device=Rd.device), shift=False)
    W = R.shape[0]
    return complex(st[0], st[1]), complex(st[5] / np.sqrt(W))


def rotation_eigenvalue_exact_singleG(Gs, Chi, n, L):
    """Exact <R_n> = det(m) for single-plane-wave spinor orbitals exp(i Gs[a].r) Chi[a] on a
    C_n-closed G set (topology.py:125-150)."""
    Gs = np.asarray(Gs, float)
    Chi = np.asarray(Chi, complex)
    N = len(Gs)
    th = 2.0 * np.pi / n
    Rp = _rot2d(th)
    phase = np.array([np.exp(-1j * th * (1 - 2 * s) / 2.0) for s in (0, 1)])
    m = np.zeros((N, N), dtype=complex)
    for b in range(N):
        RGb = Rp @ Gs[b]
        for a in range(N):
            if np.allclose(Gs[a], RGb, atol=1e-8):
                m[a, b] = np.sum(np.conj(Chi[a]) * phase * Chi[b])
    return complex(np.linalg.det(m))
