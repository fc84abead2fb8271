"""Oracle MinSR (test infrastructure): src/nqs/minsr.py:56-88, 139-152 restated in torch.

  O[s,k]   = d Re logpsi(x_s)/d theta_k + i d Im logpsi(x_s)/d theta_k        (:56-65)
  Otil     = [Re dO; Im dO], dO = O - mean_s O;  etil = [Re ec; Im ec]         (:68-73)
  MinSR    dtheta = Otil^T (Otil Otil^T + lam I)^-1 etil   (LU solve)          (:76-81)
  paramSR  dtheta = (Otil^T Otil + lam I)^-1 Otil^T etil                       (:84-88)
  step     median/MAD clip (k=5, MAD + 1e-12) of Re e, Im e kept, theta -= lr dtheta  (:139-152)
No 1/W normalisation anywhere; lam is added to the unnormalised Gram.
"""
import numpy as np
import torch
from torch.func import grad

from . import psiformer as _pf


def per_sample_O(arch, theta, R, S):
    """(W, P) complex128 by reverse-mode autodiff, one walker at a time."""
    def lr(t, r, s):
        return _pf.logpsi(arch, t, r, s).real

    def li(t, r, s):
        return _pf.logpsi(arch, t, r, s).imag
    gr, gi = grad(lr), grad(li)
    rows = []
    for w in range(R.shape[0]):
        r = torch.as_tensor(R[w], dtype=torch.float64)
        s = torch.as_tensor(S[w]).long()
        rows.append(gr(theta, r, s) + 1j * gi(theta, r, s))
    return torch.stack(rows)


def otil_etil(O, eps):
    O = torch.as_tensor(O, dtype=torch.complex128)
    eps = torch.as_tensor(eps, dtype=torch.complex128)
    dO = O - O.mean(0, keepdim=True)
    ec = eps - eps.mean()
    return torch.cat([dO.real, dO.imag], 0), torch.cat([ec.real, ec.imag])


def minsr_direction(O, eps, lam):
    Ot, et = otil_etil(O, eps)
    T = Ot @ Ot.T
    y = torch.linalg.solve(T + lam * torch.eye(T.shape[0], dtype=torch.float64), et)
    return Ot.T @ y


def paramsr_direction(O, eps, lam):
    Ot, et = otil_etil(O, eps)
    P = Ot.shape[1]
    return torch.linalg.solve(Ot.T @ Ot + lam * torch.eye(P, dtype=torch.float64), Ot.T @ et)


def clip_energies(e, clip_k=5.0):
    """minsr.py:145-148 -- median/MAD clip of the real part, imaginary part kept.
    jnp.median of an even count averages the two middle values (numpy convention)."""
    e = np.asarray(e, complex)
    er = e.real
    med = np.median(er)
    mad = np.median(np.abs(er - med)) + 1e-12
    return np.clip(er, med - clip_k * mad, med + clip_k * mad) + 1j * e.imag


def step_update(theta, O, e, lam, lr, clip_k=5.0):
    """One parameter update from given O and E_loc (minsr.py:145-152); returns (theta', E, Eerr)."""
    e = np.asarray(e, complex)
    ecl = clip_energies(e, clip_k)
    d = minsr_direction(O, torch.as_tensor(ecl), lam)
    E = float(np.mean(e).real)
    Eerr = float(np.std(e.real) / np.sqrt(e.shape[0]))
    return theta - lr * d, E, Eerr, d
