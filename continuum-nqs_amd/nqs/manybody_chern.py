"""Many-body overlaps on the HIP path (call surface of src/nqs/manybody_chern.py:23-75).

mc_overlap evaluates both wavefunctions on the walkers with nqs_psiformer_fwd and reduces the
ratios r = exp(log psi_b - log psi_a - max Re(.)) with nqs_ratio_stats, one launch each:
O_ab = mean r / sqrt(mean |r|^2), plus the effective-sample fraction.  The Fukui-Hatsugai-Suzuki
assembly of a Chern number from link overlaps is grid-sized host work (numpy).
"""
import numpy as np
import torch

from . import _lib
from .psiformer import LogPsi, as_walkers


def ratio_stats(la, lb, phase=None, shift=True):
    """[mean Re r, mean Im r, mean |r|^2, ess, shift, std r] of r = phase exp(lb - la - shift)
    (la, lb: (W,) complex device tensors) by one nqs_ratio_stats launch."""
    lib = _lib.load()
    la = torch.view_as_real(la.to(torch.complex128).contiguous())
    lb = torch.view_as_real(lb.to(torch.complex128).contiguous())
    ph = None if phase is None else torch.view_as_real(torch.as_tensor(phase, device=la.device)
                                                       .to(torch.complex128).contiguous())
    out = torch.empty(6, dtype=torch.float64, device=la.device)
    _lib.check(lib.nqs_ratio_stats(_lib.ptr(la), _lib.ptr(lb), _lib.ptr(ph), la.shape[0], 1 if shift else 0,
                                   _lib.ptr(out), _lib.stream()), "nqs_ratio_stats")
    return out.cpu().numpy()


def mc_overlap(logpsi_a, logpsi_b, R, S, batch=4096):
    """Normalised overlap O_ab from walkers drawn from |psi_a|^2 (manybody_chern.py:23-34):
    (O complex, |O|, effective-sample fraction).  `batch` is accepted for drop-in compatibility
    (the library chunks its own launches)."""
    for lp in (logpsi_a, logpsi_b):
        if not isinstance(lp, LogPsi):
            raise TypeError("mc_overlap needs LogPsi objects")
    R, S, _ = as_walkers(R, S)
    st = ratio_stats(logpsi_a(R, S), logpsi_b(R, S), shift=True)
    O = complex(st[0], st[1]) / np.sqrt(st[2])
    return O, float(abs(O)), float(st[3])


def fhs_chern(links_x, links_y):
    """FHS Chern number from link overlaps on a periodic nt x nt twist grid
    (manybody_chern.py:37-50): (C, plaquette fluxes F (nt, nt) in (-pi, pi])."""
    lx, ly = np.asarray(links_x), np.asarray(links_y)
    u = lx * np.roll(ly, -1, axis=0) * np.conj(np.roll(lx, -1, axis=1)) * np.conj(ly)
    F = np.angle(u)
    return float(F.sum() / (2 * np.pi)), F


def chern_report(links_x, links_y, absx, absy):
    """C and the grid diagnostics min |O| and max |F|/pi (manybody_chern.py:53-59)."""
    C, F = fhs_chern(links_x, links_y)
    return dict(C=C, C_int=int(np.rint(C)), F=F.tolist(), minO=float(min(np.min(absx), np.min(absy))),
                maxF_over_pi=float(np.max(np.abs(F)) / np.pi))
