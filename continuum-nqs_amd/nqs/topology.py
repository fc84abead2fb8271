"""C_n rotation eigenvalue of a many-body state on the HIP path (call surface of
src/nqs/topology.py:100-122).

The rotation acts on positions and spins together: [R_n psi](X) = U(S) psi(R_n^-1 X), where every
electron's position is turned by -2 pi/n about `center` (then wrapped into the cell) and every spin
picks up exp(-i (pi/n) sigma_z), sigma_z = +1 for s = 0.  For a configuration the spin factor is
therefore exp(-i (pi/n) (n_up - n_down)), with n_up - n_down = N - 2 sum_i s_i.  The estimator
<R_n> = < U(S) psi(R_n^-1 R, S) / psi(R, S) >_{|psi|^2} is two batched log psi evaluations
(nqs_psiformer_fwd) and one nqs_ratio_stats reduction; the position map and the phase are computed
on the device in torch.
"""
import math

import numpy as np
import torch

from . import _lib
from .psiformer import LogPsi, as_walkers
from .manybody_chern import ratio_stats


def rotate_positions(R, n, L, center=0.0):
    """Positions R (..., 2) mapped by R_n^-1 (a turn by -2 pi/n about `center`), wrapped to [0, L)^2.
    Returns a tensor on R's device (numpy input goes to the current GPU)."""
    R = torch.as_tensor(np.asarray(R, float) if not torch.is_tensor(R) else R, dtype=torch.float64)
    if not R.is_cuda:
        R = R.to(_lib.device())
    c = torch.as_tensor(np.broadcast_to(np.asarray(center, float), (2,)).copy(), dtype=torch.float64,
                        device=R.device)
    z = torch.view_as_complex((R - c).contiguous()) * complex(math.cos(2 * math.pi / n), -math.sin(2 * math.pi / n))
    return torch.remainder(torch.view_as_real(z) + c, L)


def spin_rotation_phase(S, n):
    """exp(-i (pi/n)(N - 2 sum_i s_i)) per configuration, S (..., N) in {0, 1}."""
    S = torch.as_tensor(np.asarray(S) if not torch.is_tensor(S) else S)
    m = S.shape[-1] - 2 * S.to(torch.float64).sum(-1)                   # n_up - n_down
    return torch.polar(torch.ones_like(m), -(math.pi / n) * m)


def rotation_eigenvalue_vmc(logpsi, R, S, n, L, center=0.0):
    """(<R_n>, standard error) over walkers R (W, N, 2), S (W, N); logpsi is a batched LogPsi."""
    if not isinstance(logpsi, LogPsi):
        raise TypeError("rotation_eigenvalue_vmc needs a LogPsi")
    Rd, Sd, _ = as_walkers(R, S)
    Rr = rotate_positions(Rd, n, L, center).contiguous()
    st = ratio_stats(logpsi(Rd, Sd), logpsi(Rr, Sd), phase=spin_rotation_phase(Sd, n), shift=False)
    return complex(st[0], st[1]), complex(st[5] / math.sqrt(Rd.shape[0]))
