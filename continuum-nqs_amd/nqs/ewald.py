"""2D Ewald Coulomb on a general cell (call surface of src/nqs/ewald.py:26-73).

make_ewald_lattice(A, alpha, ewald_alpha, n_real, n_recip) -> (ewald_pair, zeta_M, coulomb_energy)

The lattice tables (images A n over the square index box |n_x|,|n_y| <= n_real, reciprocal
vectors G = B m over |m| <= n_recip without G = 0, R_G = (2 pi/area) erfc(|G|/2 eta)/|G|,
V_G0 = -2 sqrt(pi)/(area eta), the Madelung term zeta_M) are built once on the host at set-up,
as the reference does; every energy is evaluated by the nqs_ewald kernel (float64,
compensated accumulation, reciprocal part through rho_G = sum_i exp(i G.r_i)).
coulomb_energy(R): R (N, 2) or (W, N, 2) -> energy (scalar or (W,)).  ewald_pair(r): the pair
potential V(r) for r (2,) or (..., 2), evaluated as the two-electron energy minus zeta_M.
"""
import ctypes
import math

import numpy as np
import torch
from scipy.special import erfc

from . import _lib
from .cell import square_cell


class Ewald:
    def __init__(self, A, alpha=1.0, ewald_alpha=None, n_real=5, n_recip=10):
        A = np.asarray(A, dtype=float)
        self.A = A
        self.area = float(abs(np.linalg.det(A)))
        B = 2.0 * np.pi * np.linalg.inv(A).T
        self.eta = math.sqrt(math.pi) / math.sqrt(self.area) if ewald_alpha is None else float(ewald_alpha)
        self.alpha = float(alpha)
        idx = np.arange(-n_real, n_real + 1)
        nx, ny = np.meshgrid(idx, idx, indexing="ij")
        self.img = np.stack([nx.ravel(), ny.ravel()], -1).astype(float) @ A.T
        jdx = np.arange(-n_recip, n_recip + 1)
        mx, my = np.meshgrid(jdx, jdx, indexing="ij")
        mc = np.stack([mx.ravel(), my.ravel()], -1).astype(float)
        mc = mc[(mc[:, 0] ** 2 + mc[:, 1] ** 2) > 0]
        self.G = mc @ B.T
        gm = np.linalg.norm(self.G, axis=-1)
        self.RG = (2.0 * np.pi / self.area) * erfc(gm / (2.0 * self.eta)) / gm
        self.VG0 = -2.0 * math.sqrt(math.pi) / (self.area * self.eta)
        rs = np.linalg.norm(self.img, axis=-1)
        rs = rs[rs > 1e-10]
        self.zeta_M = float(self.alpha * (float(np.sum(erfc(self.eta * rs) / rs)) - 2.0 * self.eta / math.sqrt(math.pi)
                                          + float(np.sum(self.RG)) + self.VG0))
        self._dev = None

    def _tables(self):
        dev = _lib.device()
        if self._dev is None or self._dev[0] != dev:
            t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
            self._dev = (dev, t(self.img), t(self.G), t(self.RG))
        _, img, G, RG = self._dev
        e = _lib.EwaldDesc()
        e.alpha, e.eta, e.v_g0, e.zeta_m = self.alpha, self.eta, self.VG0, self.zeta_M
        e.n_img, e.n_g = img.shape[0], G.shape[0]
        e.img_dev, e.g_dev, e.rg_dev = img.data_ptr(), G.data_ptr(), RG.data_ptr()
        return e

    def coulomb_energy(self, R):
        lib = _lib.load()
        dev = _lib.device()
        R = torch.as_tensor(np.asarray(R) if not torch.is_tensor(R) else R)
        single = R.dim() == 2
        R = (R[None] if single else R).to(device=dev, dtype=torch.float64).contiguous()
        out = torch.empty(R.shape[0], dtype=torch.float64, device=dev)
        e = self._tables()
        _lib.check(lib.nqs_ewald(ctypes.byref(e), _lib.ptr(R), R.shape[0], R.shape[1], _lib.ptr(out),
                                 _lib.stream()), "nqs_ewald")
        return out[0] if single else out

    def ewald_pair(self, r):
        r = torch.as_tensor(np.asarray(r) if not torch.is_tensor(r) else r, dtype=torch.float64)
        shp = r.shape[:-1]
        r2 = r.reshape(-1, 2)
        R = torch.stack([r2, torch.zeros_like(r2)], 1)          # (M, 2, 2): electrons at r and 0
        v = self.coulomb_energy(R) - self.zeta_M
        return v.reshape(shp)


def make_ewald_lattice(A, alpha=1.0, ewald_alpha=None, n_real=5, n_recip=10):
    e = Ewald(A, alpha, ewald_alpha, n_real, n_recip)
    return e.ewald_pair, e.zeta_M, e.coulomb_energy


def make_ewald(L, alpha=1.0, ewald_alpha=None, n_real=4, n_recip=8):
    return make_ewald_lattice(square_cell(L), alpha=alpha, ewald_alpha=ewald_alpha, n_real=n_real, n_recip=n_recip)
