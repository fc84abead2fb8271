"""Pinning fields (call surface of src/nqs/pinning.py:20-119) and Gaussian pinning centres.

A pinning field adds -h_pin sum_i n_target(r_i).sigma_i to the Hamiltonian through the local
energy's ext_field.  Every target of the reference is a short Fourier series, or one normalised to
unit length, so each factory here returns an ExtField (the table the E_loc assembly kernel
evaluates per electron): the triple-Q Neel / Bloch skyrmion (normalised), the ferromagnet, the
cycloidal / flat spiral, the Overhauser z-SDW, and the texture of a converged Hartree-Fock run
(normalised).  The arithmetic of each term follows the reference line by line, so the fields agree
with it up to rounding.

GaussianPin (no reference counterpart: the reference has no localised pin) is a periodic Gaussian
centre that adds a scalar potential v0 g(r) and/or a spin field h g(r), coupled like ext_field;
pass a list of them as make_local_energy(..., pinning=[...]).
"""
from typing import NamedTuple, Tuple

import numpy as np

from .local_energy import ExtField


class GaussianPin(NamedTuple):
    """Centre c, width w: g(r) = sum over the 3 x 3 images around the minimum image of r - c of
    exp(-|.|^2 / (2 w^2)); the electron sees v0 g(r) and the field h g(r)."""
    center: Tuple[float, float]
    width: float
    v0: float = 0.0
    h: Tuple[float, float, float] = (0.0, 0.0, 0.0)

    def row(self):
        cx, cy = (float(x) for x in self.center)
        hx, hy, hz = (float(x) for x in self.h)
        if not (self.width > 0):
            raise ValueError("GaussianPin: width must be > 0")
        return [cx, cy, float(self.width), float(self.v0), hx, hy, hz, 0.0]


def make_gaussian_pins(centers, amplitude, width, kind="spin", axis=(0.0, 0.0, 1.0)):
    """One GaussianPin per centre: kind "spin" gives the field amplitude * axis_hat (a localised
    version of the reference's pinning fields), kind "scalar" the potential amplitude."""
    ax = np.asarray(axis, float)
    ax = ax / np.linalg.norm(ax)
    if kind == "spin":
        return [GaussianPin(tuple(c), width, 0.0, tuple(amplitude * ax)) for c in centers]
    if kind == "scalar":
        return [GaussianPin(tuple(c), width, float(amplitude)) for c in centers]
    raise ValueError(kind)


def make_skx_pin(q1, q2, h_pin, m0=0.3, chirality=1.0, neel=True):
    """Triple-Q skyrmion (pinning.py:20-37): arms q1, q2, q3 = -(q1 + q2);
    m = (chirality sum_j sin(q_j.r) e_j, m0 + sum_j cos(q_j.r)) with e_j = q_hat_j (Neel) or
    z_hat x q_hat_j (Bloch); f = h_pin m / (|m| + 1e-12)."""
    qs = np.stack([np.asarray(q1, float), np.asarray(q2, float), -(np.asarray(q1, float) + np.asarray(q2, float))])
    qhat = qs / np.linalg.norm(qs, axis=1, keepdims=True)
    ehat = qhat if neel else np.stack([-qhat[:, 1], qhat[:, 0]], 1)
    q = np.concatenate([np.zeros((1, 2)), qs])
    c = np.zeros((4, 3))
    s = np.zeros((4, 3))
    c[0, 2] = m0
    c[1:, 2] = 1.0
    s[1:, 0] = chirality * ehat[:, 0]
    s[1:, 1] = chirality * ehat[:, 1]
    return ExtField(q, c, s, scale=h_pin)


def make_fm_pin(h_pin, axis=(0.0, 0.0, 1.0)):
    """Uniform h_pin axis_hat (pinning.py:40-46)."""
    ax = np.asarray(axis, float)
    ax = ax / np.linalg.norm(ax)
    return ExtField(np.zeros((1, 2)), (h_pin * ax)[None], np.zeros((1, 3)))


def make_spiral_pin(Q, h_pin, cycloidal=True):
    """Single-Q spiral (pinning.py:49-61): cycloid h (sin(Q.r) q_hat, cos(Q.r)) or flat helix
    h (cos(Q.r), sin(Q.r), 0)."""
    Q = np.asarray(Q, float)
    qh = Q / (np.linalg.norm(Q) + 1e-12)
    if cycloidal:
        return ExtField(Q[None], [[0.0, 0.0, h_pin]], [[h_pin * qh[0], h_pin * qh[1], 0.0]])
    return ExtField(Q[None], [[h_pin, 0.0, 0.0]], [[0.0, h_pin, 0.0]])


def make_overhauser_pin(Q, h_pin):
    """Collinear z-SDW h (0, 0, cos(Q.r)) (pinning.py:64-71)."""
    return ExtField(np.asarray(Q, float)[None], [[0.0, 0.0, h_pin]], [[0.0, 0.0, 0.0]])


def load_hf_texture(path):
    """The converged Hartree-Fock texture file (.npz, pinning.py:74-84), read without pickles."""
    d = np.load(path, allow_pickle=False)
    out = dict(q1=d["q1"], q2=d["q2"], s_keys=d["s_keys"], s_vals=d["s_vals"], BL=float(d["BL"]),
               E_var=float(d["E_var"]), q_star=float(d["q_star"]), e_per_cell=float(d["e_per_cell"]))
    if "n_keys" in d.files:
        out["n_keys"] = d["n_keys"]
        out["n_vals"] = d["n_vals"]
    return out


def make_hf_skx_pin(s_keys, s_vals, q1, q2, h_pin):
    """Texture of a converged HF run (pinning.py:87-100): m(r) = Re sum_G S_G exp(i G.r) with
    G = m q1 + n q2, i.e. cos terms Re S_G and sin terms -Im S_G; f = h_pin m / (|m| + 1e-12)."""
    q1, q2 = np.asarray(q1, float), np.asarray(q2, float)
    G = np.array([m * q1 + n * q2 for (m, n) in np.asarray(s_keys)]).reshape(-1, 2)
    S = np.asarray(s_vals, dtype=complex).reshape(-1, 3)
    return ExtField(G, S.real, -S.imag, scale=h_pin)


def make_pin(kind, q1, q2, h_pin, **kw):
    """Dispatch by competitor name (pinning.py:103-119); "fluid" is no field (None)."""
    if kind == "fluid":
        return None
    if kind == "fm":
        return make_fm_pin(h_pin, **kw)
    if kind == "spiral":
        return make_spiral_pin(q1, h_pin)
    if kind == "spiral_xy":
        return make_spiral_pin(q1, h_pin, cycloidal=False)
    if kind == "overhauser":
        return make_overhauser_pin(q1, h_pin)
    if kind == "skx":
        return make_skx_pin(q1, q2, h_pin, **kw)
    if kind == "skx_bloch":
        return make_skx_pin(q1, q2, h_pin, neel=False, **kw)
    raise ValueError(kind)
