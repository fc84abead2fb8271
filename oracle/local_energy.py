"""Oracle local energy (test infrastructure): src/nqs/local_energy.py:40-117 restated in torch.

E_loc = Psi^-1 H Psi for one configuration, generic in logpsi(R, S) -> complex128:
  kinetic  -1/(2 m*) (Lap L + grad L . grad L), with grad / Lap of Re L and Im L taken
           separately by dense autodiff (local_energy.py:52-81, the dense-Hessian branch)
  Zeeman   -(h_z/2) sum_i (1 - 2 s_i) - sum_i f_z(r_i) (1 - 2 s_i)       (:83-88)
  SOC      per electron: flip s_i, ratio = exp(L' - L), d_{x_i,y_i} L' of the flipped
           configuration, non-finite -> 0, lambda (sx (-i) ratio dy L' + sy (-i) ratio dx L')
           - (f_x sx + f_y sy) ratio                                       (:90-114)
  pinning  optional Gaussian centres (an extension: the reference has none), rows
           (cx, cy, width, v0, hx, hy, hz): periodic profile g(r) = sum over the 3 x 3 images
           around the minimum image of r - c of exp(-|.|^2 / (2 width^2)); adds the scalar
           potential v0 g(r_i) and the spin field h g(r_i), coupled like ext_field
"""
import numpy as np
import torch
from torch.func import grad, hessian

_SY = ((0.0, -1.0j), (1.0j, 0.0))


def _pin_rows(pinning):
    return np.asarray(pinning, float).reshape(-1, 7)


def pin_profiles(R, pinning, A):
    """g_p(r) for every centre p and position r: (n_pin, M) for R (M, 2)."""
    R = np.asarray(R, float).reshape(-1, 2)
    A = np.asarray(A, float)
    Ai = np.linalg.inv(A)
    out = []
    for cx, cy, wd, *_ in _pin_rows(pinning):
        u = R - np.array([cx, cy])
        u = u - np.rint(u @ Ai.T) @ A.T
        g = np.zeros(len(R))
        for nx in (-1, 0, 1):
            for ny in (-1, 0, 1):
                d = u - A @ np.array([nx, ny], float)
                g += np.exp(-(d * d).sum(1) / (2.0 * wd * wd))
        out.append(g)
    return np.array(out).reshape(-1, len(R))


def pinning_potential(R, pinning, A):
    """sum_i sum_p v0_p g_p(r_i)."""
    return float((_pin_rows(pinning)[:, 3:4] * pin_profiles(R, pinning, A)).sum())


def pinning_field(R, pinning, A):
    """Spin field of the centres, sum_p h_p g_p(r): (M, 3)."""
    return pin_profiles(R, pinning, A).T @ _pin_rows(pinning)[:, 4:7]


def make_local_energy(logpsi, m_star=1.0, lambda_r=1.0, h_z=0.0, ext_field=None, pinning=None, cell=None):
    """Returns E_loc(R (N,2), S (N,)) -> complex (python complex)."""

    def _lr(rf, S):
        return logpsi(rf.reshape(-1, 2), S).real

    def _li(rf, S):
        return logpsi(rf.reshape(-1, 2), S).imag

    g_r, g_i = grad(_lr), grad(_li)
    h_r, h_i = hessian(_lr), hessian(_li)

    def field(R):
        """ext_field (if any) plus the pinning spin field at the rows of R."""
        f = torch.zeros(R.shape[0], 3, dtype=torch.float64)
        if ext_field is not None:
            f = f + torch.as_tensor(ext_field(R), dtype=torch.float64).reshape(-1, 3)
        if pinning is not None:
            f = f + torch.as_tensor(pinning_field(R.detach().numpy(), pinning, cell))
        return f

    def parts(R, S):
        R = torch.as_tensor(R, dtype=torch.float64)
        S = torch.as_tensor(S).long()
        N = R.shape[0]
        rf = R.reshape(-1)
        L0 = logpsi(R, S)
        g = g_r(rf, S) + 1j * g_i(rf, S)
        lap = torch.trace(h_r(rf, S)) + 1j * torch.trace(h_i(rf, S))
        e_kin = -1.0 / (2.0 * m_star) * (lap + (g * g).sum())
        sgn = 1.0 - 2.0 * S.to(torch.float64)
        e_z = -0.5 * h_z * sgn.sum()
        fields = ext_field is not None or pinning is not None
        if fields:
            f_all = field(R)                                   # (N, 3)
            e_z = e_z - (f_all[:, 2] * sgn).sum()
        e_soc = 0.0 + 0.0j
        for i in range(N):
            s_i = int(S[i])
            s_p = 1 - s_i
            Sf = S.clone()
            Sf[i] = s_p
            Lf = logpsi(R, Sf)
            ratio = torch.exp(Lf - L0)
            try:
                gf = g_r(rf, Sf) + 1j * g_i(rf, Sf)
            except RuntimeError:
                # exactly singular flipped determinant: jax returns a non-finite gradient here,
                # torch raises; both end in the guard below
                gf = torch.full((2 * N,), complex("nan+nanj"), dtype=torch.complex128)
            ax = ratio * gf[2 * i]
            ay = ratio * gf[2 * i + 1]
            ax = ax if torch.isfinite(ax.real) and torch.isfinite(ax.imag) else torch.zeros_like(ax)
            ay = ay if torch.isfinite(ay.real) and torch.isfinite(ay.imag) else torch.zeros_like(ay)
            sx = 1.0
            sy = _SY[s_i][s_p]
            e = lambda_r * (sx * (-1j) * ay + sy * (-1j) * ax)
            if ext_field is not None or (pinning is not None and (f_all[i, 0] != 0 or f_all[i, 1] != 0)):
                f = f_all[i]
                e = e - (f[0] * sx + f[1] * sy) * ratio
            e_soc = e_soc + complex(e)
        e_pin = 0.0 if pinning is None else pinning_potential(R.detach().numpy(), pinning, cell)
        return dict(kin=complex(e_kin), zeeman=float(e_z), soc=complex(e_soc), pin=e_pin, logpsi=complex(L0),
                    grad=g.detach(), lap=complex(lap))

    def E_loc(R, S):
        p = parts(R, S)
        return p["kin"] + p["zeeman"] + p["soc"] + p["pin"]

    E_loc.parts = parts
    return E_loc


def vmc_energy(E_loc, R, S):
    """src/nqs/sampler.py:94-98 -- mean, std/sqrt(W), values (real parts)."""
    import numpy as np
    e = np.array([E_loc(R[w], S[w]) for w in range(R.shape[0])])
    er = e.real
    return float(er.mean()), float(er.std() / np.sqrt(er.shape[0])), e
