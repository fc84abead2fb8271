"""CPU checks of the C ABI boundary (no compute calls): the library loads, every symbol that
include/nqs_hip.h declares is exported, the parameter layout matches the oracle's real_ravel
order, descriptor validation returns errors instead of crashing."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import bands, psiformer as opf
from nqs import _lib
from nqs.psiformer import ContinuumSpinorPsiformer, ravel_params

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), os.pardir))


def _declared():
    txt = open(os.path.join(ROOT, "include", "nqs_hip.h")).read()
    return sorted(set(re.findall(r"\b(nqs_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert {n for n, _, _ in _lib.SIGNATURES} <= set(names)
    assert lib.nqs_version() >= 1


def test_descriptor_layouts_match_the_bindings():
    lib = _lib.load()
    for which, cls in enumerate([_lib.ModelDesc, _lib.HamDesc, _lib.EwaldDesc, _lib.MinsrDesc]):
        assert lib.nqs_desc_size(which) == ctypes.sizeof(cls), cls.__name__
    assert lib.nqs_desc_size(4) == -1


ARCHS = [dict(embed_dim=16, depth=2, n_heads=4, n_freq=4),
         dict(embed_dim=16, depth=2, n_heads=2, n_freq=3, n_dets=3, use_jastrow=True, learnable_env=True, n_knots=6),
         dict(embed_dim=8, depth=12, n_heads=2, n_freq=2, n_dets=2, use_jastrow=True, jastrow_kind="cusp"),
         dict(embed_dim=128, depth=4, n_heads=8, n_freq=8, n_dets=8)]


@pytest.mark.parametrize("kw", ARCHS)
def test_param_layout_matches_oracle(kw):
    N = 16 if kw["embed_dim"] == 128 else 5
    g = bands.rashba_gs(7.0, N, 1.0, 1.0, 0.5, n_cut=10, twist=(0.1, 0.17))
    m = ContinuumSpinorPsiformer(n_elec=N, Gs=g["Gs"], W=g["W"], L=7.0, **kw)
    a = opf.Arch(N, g["Gs"], g["W"], 7.0, **kw)
    lo, P = opf.ravel_layout(a)
    lk = m.layout()
    assert [(x[0], tuple(x[1]), x[2], x[3]) for x in lo] == [(x[0], tuple(x[1]), x[2], x[3]) for x in lk]
    assert m.n_params() == P


def test_param_count_formula_c3():
    """P = [(4F+2)d + d] + depth (8d^2 + 11d) + 2d + 2(d+1)KN + 2K for N=16, d=128, 4 layers, K=8, F=8."""
    N, d, depth, K, F = 16, 128, 4, 8, 8
    g = bands.rashba_gs(21.27, N, 1.0, 0.8, 0.05, n_cut=10, twist=(0.1, 0.17))
    m = ContinuumSpinorPsiformer(n_elec=N, Gs=g["Gs"], W=g["W"], L=21.27, embed_dim=d, depth=depth, n_heads=8,
                                 n_freq=F, n_dets=K)
    assert m.n_params() == (4 * F + 2) * d + d + depth * (8 * d * d + 11 * d) + 2 * d + 2 * (d + 1) * K * N + 2 * K
    assert m.n_params() == 567696


def test_init_tree_ravels_to_layout():
    g = bands.rashba_gs(7.0, 5, 1.0, 1.0, 0.5, n_cut=10, twist=(0.1, 0.17))
    m = ContinuumSpinorPsiformer(n_elec=5, Gs=g["Gs"], W=g["W"], L=7.0, embed_dim=16, depth=2, n_heads=4, n_dets=2)
    p = m.init(0)
    th = ravel_params(p)
    assert th.shape[0] == m.n_params()
    lay = {n: (s, c, o) for n, s, c, o in m.layout()}
    s, c, o = lay["params/LayerNorm_0/scale"]
    assert np.all(th[o:o + s[0]].numpy() == 1.0)


def test_descriptor_validation_errors():
    lib = _lib.load()
    d = _lib.ModelDesc()
    d.n_elec, d.n_g, d.embed_dim, d.depth, d.n_heads, d.ff_mult, d.n_freq, d.n_dets = 5, 5, 16, 2, 3, 2, 4, 1
    d.cell[0] = d.cell[3] = 7.0
    d.cell_inv[0] = d.cell_inv[3] = 1 / 7.0
    assert lib.nqs_param_count(ctypes.byref(d)) == -1          # 16 not divisible by 3 heads
    assert b"multiple" in lib.nqs_last_error()
    d.n_heads = 4
    assert lib.nqs_param_count(ctypes.byref(d)) > 0
    d.cell_inv[0] = 1.0
    assert lib.nqs_param_count(ctypes.byref(d)) == -1
    assert lib.nqs_workspace_bytes(ctypes.byref(d), 99, 10) == -1
