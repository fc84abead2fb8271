"""Oracle pinning textures (test infrastructure): src/nqs/pinning.py:20-119 restated in torch.

Each factory returns ext_field(r) -> (..., 3) for positions r (..., 2), the field coupled as
-sum_i f(r_i).sigma_i by the local energy.  Pins (tests/test_oracle.py): the triple-Q Neel and
Bloch skyrmions carry Berg-Luscher charge -1 per magnetic cell and flip it with the chirality, the
normalised textures have |f| = h_pin, the spiral, overhauser and ferromagnet reduce to their closed
forms, and a converged-HF texture built from the skyrmion's own harmonics reproduces it.
"""
import numpy as np
import torch


def _t(x):
    return torch.as_tensor(np.asarray(x, float), dtype=torch.float64)


def make_skx_pin(q1, q2, h_pin, m0=0.3, chirality=1.0, neel=True):
    """pinning.py:20-37: arms q1, q2, q3 = -(q1 + q2); in-plane part along q_hat (Neel) or
    z_hat x q_hat (Bloch); f = h_pin m / (|m| + 1e-12)."""
    qs = _t(np.stack([q1, q2, -(np.asarray(q1, float) + np.asarray(q2, float))]))
    qhat = qs / torch.linalg.norm(qs, dim=1, keepdim=True)
    ehat = qhat if neel else torch.stack([-qhat[:, 1], qhat[:, 0]], 1)

    def ext_field(r):
        r = torch.as_tensor(r, dtype=torch.float64)
        qr = r @ qs.T                                               # (..., 3)
        s = torch.sin(qr)
        m = torch.stack([chirality * (s * ehat[:, 0]).sum(-1), chirality * (s * ehat[:, 1]).sum(-1),
                         m0 + torch.cos(qr).sum(-1)], -1)
        return h_pin * m / (torch.linalg.norm(m, dim=-1, keepdim=True) + 1e-12)
    return ext_field


def make_fm_pin(h_pin, axis=(0.0, 0.0, 1.0)):
    """pinning.py:40-46: uniform h_pin axis_hat."""
    ax = _t(axis)
    ax = ax / torch.linalg.norm(ax)

    def ext_field(r):
        r = torch.as_tensor(r, dtype=torch.float64)
        return (h_pin * ax).expand(*r.shape[:-1], 3)
    return ext_field


def make_spiral_pin(Q, h_pin, cycloidal=True):
    """pinning.py:49-61: cycloid h (sin(Q.r) q_hat, cos(Q.r)), or the flat helix h (cos, sin, 0)."""
    Q = _t(Q)
    qh = Q / (torch.linalg.norm(Q) + 1e-12)

    def ext_field(r):
        ph = torch.as_tensor(r, dtype=torch.float64) @ Q
        if cycloidal:
            return h_pin * torch.stack([torch.sin(ph) * qh[0], torch.sin(ph) * qh[1], torch.cos(ph)], -1)
        return h_pin * torch.stack([torch.cos(ph), torch.sin(ph), torch.zeros_like(ph)], -1)
    return ext_field


def make_overhauser_pin(Q, h_pin):
    """pinning.py:64-71: h (0, 0, cos(Q.r))."""
    Q = _t(Q)

    def ext_field(r):
        ph = torch.as_tensor(r, dtype=torch.float64) @ Q
        return h_pin * torch.stack([torch.zeros_like(ph), torch.zeros_like(ph), torch.cos(ph)], -1)
    return ext_field


def make_hf_skx_pin(s_keys, s_vals, q1, q2, h_pin):
    """pinning.py:87-100: m(r) = Re sum_G S_G exp(i G.r), G = m q1 + n q2; f = h_pin m / (|m| + 1e-12)."""
    q1, q2 = np.asarray(q1, float), np.asarray(q2, float)
    Gs = _t(np.array([m * q1 + n * q2 for (m, n) in np.asarray(s_keys)]))
    S = torch.as_tensor(np.asarray(s_vals, dtype=complex))

    def ext_field(r):
        ph = torch.as_tensor(r, dtype=torch.float64) @ Gs.T            # (..., nG)
        m = torch.real(torch.exp(1j * ph) @ S)                           # (..., 3)
        return h_pin * m / (torch.linalg.norm(m, dim=-1, keepdim=True) + 1e-12)
    return ext_field


def make_pin(kind, q1, q2, h_pin, **kw):
    """pinning.py:103-119: dispatch by competitor name."""
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


def skyrmion_charge(field_fn, A, ngrid=48):
    """Lattice topological charge of n = f/|f| over the cell A (columns = lattice vectors): the
    sum over grid triangles of the signed solid angle 2 atan2(n1.(n2 x n3), 1 + n1.n2 + n2.n3 +
    n3.n1) / (4 pi) (Berg and Luscher, Nucl. Phys. B 190 (1981) 412), an integer for a smooth
    texture on a fine enough grid."""
    A = np.asarray(A, float)
    fr = np.arange(ngrid) / ngrid
    u, v = np.meshgrid(fr, fr, indexing="ij")
    pts = np.stack([u, v], -1) @ A.T
    n = field_fn(torch.tensor(pts)).numpy()
    n = n / np.linalg.norm(n, axis=-1, keepdims=True)

    def omega(a, b, c):
        num = np.einsum("...i,...i->...", a, np.cross(b, c))
        den = 1.0 + (a * b).sum(-1) + (b * c).sum(-1) + (c * a).sum(-1)
        return 2.0 * np.arctan2(num, den)

    n00 = n
    n10 = np.roll(n, -1, axis=0)
    n01 = np.roll(n, -1, axis=1)
    n11 = np.roll(n10, -1, axis=1)
    q = omega(n00, n10, n11) + omega(n00, n11, n01)
    return float(q.sum() / (4.0 * np.pi))
