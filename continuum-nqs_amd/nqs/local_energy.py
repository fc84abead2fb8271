"""Local energy on the HIP path (call surface of src/nqs/local_energy.py:40-117).

E_loc(R, S) = Psi^-1 H Psi with
    H = sum_i [ -1/(2 m*) grad_i^2 + lambda_R (sigma_x (-i d_{y_i}) + sigma_y (-i d_{x_i}))
                - (h_z/2) sigma_z - ext_field(r_i) . sigma ]
computed by nqs_eloc: the kinetic term from a forward-Laplacian jet pass (value, the 2N
position tangents and the Laplacian carried through the network in one sweep), the SOC term
from the N single-spin-flipped configurations with the two tangents of the flipped electron,
Zeeman and the external field in the assembly kernel.  `use_folx` is accepted for drop-in
compatibility; the forward-Laplacian pass is the only kinetic path.

ext_field must be an ExtField (a Fourier series m(r) = sum_t c_t cos(q_t.r) + s_t sin(q_t.r),
optionally normalised to scale m/(|m| + 1e-12)), e.g. from spiral_reference.spiral_field or the
texture factories of nqs.pinning.  `pinning` is an optional list of nqs.pinning.GaussianPin
centres (scalar potential and/or spin field with a periodic Gaussian profile; no reference
counterpart).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .psiformer import LogPsi, as_walkers


def pin_table(pins):
    """(n_pin, 8) rows (cx, cy, width, v0, hx, hy, hz, 0) of a list of GaussianPin."""
    from .pinning import GaussianPin
    rows = []
    for p in pins:
        if not isinstance(p, GaussianPin):
            raise TypeError("pinning must be a list of nqs.pinning.GaussianPin")
        rows.append(p.row())
    return np.asarray(rows, float).reshape(-1, 8)


class ExtField:
    """m(r) = sum_t c_t cos(q_t . r) + s_t sin(q_t . r), each c_t, s_t a 3-vector (x, y, z);
    the field is f = m, or f = scale m / (|m| + 1e-12) when `scale` is given (a normalised
    texture, as the pinning targets of src/nqs/pinning.py:20-100 are)."""

    def __init__(self, q, c, s, scale=None):
        self.q = np.asarray(q, float).reshape(-1, 2)
        self.c = np.asarray(c, float).reshape(-1, 3)
        self.s = np.asarray(s, float).reshape(-1, 3)
        if not (len(self.q) == len(self.c) == len(self.s)):
            raise ValueError("ExtField: q, c, s need one row per term")
        if scale is not None and not (np.isfinite(scale) and scale > 0):
            raise ValueError("ExtField: scale must be a positive number (or None)")
        self.scale = None if scale is None else float(scale)

    def table(self):
        return np.concatenate([self.q, self.c, self.s], axis=1)          # (n, 8)

    def __call__(self, r):
        """Host evaluation for inspection: r (..., 2) -> (..., 3)."""
        r = np.asarray(r, float)
        ph = r @ self.q.T
        m = np.cos(ph) @ self.c + np.sin(ph) @ self.s
        if self.scale is None:
            return m
        return self.scale * m / (np.linalg.norm(m, axis=-1, keepdims=True) + 1e-12)


class LocalEnergy:
    def __init__(self, logpsi, m_star=1.0, lambda_r=1.0, h_z=0.0, ext_field=None, pinning=None):
        if not isinstance(logpsi, LogPsi):
            raise TypeError("make_local_energy needs a logpsi from nqs.psiformer.make_logpsi / slater_reference")
        if ext_field is not None and not isinstance(ext_field, ExtField):
            raise TypeError("ext_field must be an nqs.local_energy.ExtField (Fourier series)")
        self.logpsi, self.m_star, self.lambda_r, self.h_z, self.ext_field = logpsi, m_star, lambda_r, h_z, ext_field
        self._field = None
        # Gaussian pinning centres (nqs.pinning.GaussianPin): no counterpart in the reference
        self.pinning = None if pinning is None else pin_table(pinning)
        self._pin = None
        # device workspace name: calls that may run concurrently (on different streams) need their own
        self.ws_tag = "default"

    def _ham(self):
        h = _lib.HamDesc()
        h.m_star, h.lambda_r, h.h_z = float(self.m_star), float(self.lambda_r), float(self.h_z)
        if self.ext_field is not None:
            if self._field is None:
                self._field = torch.tensor(self.ext_field.table(), dtype=torch.float64, device=_lib.device())
            h.n_field = self._field.shape[0]
            h.field_dev = self._field.data_ptr()
            h.field_scale = 0.0 if self.ext_field.scale is None else self.ext_field.scale
        else:
            h.n_field = 0
            h.field_dev = None
        if self.pinning is not None and len(self.pinning):
            if self._pin is None:
                self._pin = torch.tensor(self.pinning, dtype=torch.float64, device=_lib.device())
            h.n_pin = self._pin.shape[0]
            h.pin_dev = self._pin.data_ptr()
            for i, v in enumerate(self.logpsi.model.A.reshape(-1)):
                h.cell[i] = float(v)
        else:
            h.n_pin = 0
            h.pin_dev = None
        return h

    def parts(self, R, S):
        """(E_loc, log psi, grad log psi (W, 2N), Laplacian of log psi), complex device tensors."""
        lib = _lib.load()
        R, S, single = as_walkers(R, S)
        W, N = R.shape[0], R.shape[1]
        m = self.logpsi.model
        d = m.desc()
        h = self._ham()
        e = torch.empty(W, 2, dtype=torch.float64, device=R.device)
        lp = torch.empty(W, 2, dtype=torch.float64, device=R.device)
        g = torch.empty(W, 2 * N, 2, dtype=torch.float64, device=R.device)
        lap = torch.empty(W, 2, dtype=torch.float64, device=R.device)
        nb = lib.nqs_workspace_bytes(ctypes.byref(d), _lib.OP_ELOC, W)
        ws = _lib.workspace(nb, self.ws_tag)
        _lib.check(lib.nqs_eloc(ctypes.byref(d), ctypes.byref(h), _lib.ptr(self.logpsi.packed), _lib.ptr(R),
                                _lib.ptr(S), W, _lib.ptr(e), _lib.ptr(lp), _lib.ptr(g), _lib.ptr(lap), _lib.ptr(ws),
                                ws.numel(), _lib.stream()), "nqs_eloc")
        c = torch.view_as_complex
        out = (c(e), c(lp), c(g), c(lap))
        return tuple(o[0] for o in out) if single else out

    def __call__(self, R, S):
        return self.parts(R, S)[0]


def make_local_energy(logpsi, m_star=1.0, lambda_r=1.0, h_z=0.0, ext_field=None, use_folx=False, pinning=None):
    return LocalEnergy(logpsi, m_star, lambda_r, h_z, ext_field, pinning)
