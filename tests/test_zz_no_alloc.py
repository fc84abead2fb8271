"""The allocation-free log psi apply (nqs_psiformer_fwd_noalloc) against nqs_psiformer_fwd, both
called through the C ABI: the headline architecture in mixed precision (its activations need more
than the 64 MiB scratch at 2048 walkers, so the call runs in several chunks) and in float64, two
calls in a row with the second on fewer walkers, and a model too large for the scratch rejected
with NQS_ERR_WORKSPACE before any launch."""
import ctypes
import math

import numpy as np
import pytest
import torch

from nqs import _lib
from nqs.rashba_reference import noninteracting_gs, occupied_orbital_matrix, planewave_envelope
from nqs.psiformer import ContinuumSpinorPsiformer, LogPsi

pytestmark = pytest.mark.gpu

N_E, RS = 16, 3.0
ARCH = dict(embed_dim=128, depth=4, n_heads=8, n_freq=8, n_dets=8)


def _model(precision):
    L = RS * math.sqrt(math.pi * N_E)
    g = noninteracting_gs(N_E, L, 1.0, 0.8, 0.05, n_cut=10, twist=(0.1, 0.17))
    Gs, Wenv = planewave_envelope(*occupied_orbital_matrix(g))
    m = ContinuumSpinorPsiformer(n_elec=N_E, Gs=Gs, W=Wenv, L=L, precision=precision, **ARCH)
    params = m.init(3)
    rng = np.random.default_rng(4)
    kr = params["params"]["backflow"]["kernel"]
    z = torch.tensor(rng.standard_normal(tuple(kr.shape)) * 0.05)
    params["params"]["backflow"]["kernel"] = (z + 1j * z).to(torch.complex128)
    from nqs.psiformer import ravel_params
    return m, LogPsi(m, ravel_params(params).cuda()), L


def _walkers(W, L, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    R = (torch.rand(W, N_E, 2, generator=g, dtype=torch.float64, device="cuda") * L).contiguous()
    S = torch.randint(0, 2, (W, N_E), generator=g, device="cuda").to(torch.int32).contiguous()
    return R, S


def _fwd(lib, d, lp, R, S):
    W = R.shape[0]
    out = torch.empty(W, 2, dtype=torch.float64, device="cuda")
    nb = lib.nqs_workspace_bytes(ctypes.byref(d), _lib.OP_LOGPSI, W)
    ws = torch.empty(nb, dtype=torch.uint8, device="cuda")
    _lib.check(lib.nqs_psiformer_fwd(ctypes.byref(d), _lib.ptr(lp.packed), _lib.ptr(R), _lib.ptr(S), W, _lib.ptr(out),
                                     _lib.ptr(ws), nb, _lib.stream()), "nqs_psiformer_fwd")
    return out


def _fwd_noalloc(lib, d, lp, R, S):
    W = R.shape[0]
    out = torch.full((W, 2), float("nan"), dtype=torch.float64, device="cuda")
    _lib.check(lib.nqs_psiformer_fwd_noalloc(ctypes.byref(d), _lib.ptr(lp.packed), _lib.ptr(R), _lib.ptr(S), W,
                                             _lib.ptr(out), _lib.stream()), "nqs_psiformer_fwd_noalloc")
    return out


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_noalloc_logpsi_equals_the_workspace_entry_point(precision):
    lib = _lib.load()
    m, lp, L = _model(precision)
    d = m.desc()
    assert lib.nqs_workspace_bytes(ctypes.byref(d), _lib.OP_LOGPSI, 2048) > (64 << 20)
    for W, seed in ((2048, 11), (300, 12)):
        R, S = _walkers(W, L, seed)
        ref = _fwd(lib, d, lp, R, S)
        got = _fwd_noalloc(lib, d, lp, R, S)
        torch.cuda.synchronize()
        diff = float((got - ref).abs().max())
        print(f"{precision} W={W}: max |noalloc - fwd| = {diff:.3e}", flush=True)
        assert torch.isfinite(got).all()
        assert diff <= 1e-12 * max(1.0, float(ref.abs().max()))


def test_noalloc_rejects_a_model_beyond_the_scratch():
    """K = 8192 determinants of 32 orbitals: the backflow projection alone needs 128 MiB for one walker,
    so the call returns NQS_ERR_WORKSPACE before any launch (the other arguments are valid device
    buffers of the right sizes, never read)."""
    lib = _lib.load()
    L, N = 10.0, 32
    g = noninteracting_gs(N, L, 1.0, 0.5, 0.0, n_cut=12, twist=(0.1, 0.17))
    Gs, Wenv = planewave_envelope(*occupied_orbital_matrix(g))
    m = ContinuumSpinorPsiformer(n_elec=N, Gs=Gs, W=Wenv, L=L, precision="fp64", embed_dim=32, depth=1,
                                 n_heads=4, n_freq=4, n_dets=8192)
    d = m.desc()
    assert lib.nqs_workspace_bytes(ctypes.byref(d), _lib.OP_LOGPSI, 1) > (64 << 20)
    packed = torch.zeros(lib.nqs_packed_bytes(ctypes.byref(d)), dtype=torch.uint8, device="cuda")
    out = torch.empty(1, 2, dtype=torch.float64, device="cuda")
    R = torch.zeros(1, N, 2, dtype=torch.float64, device="cuda")
    S = torch.zeros(1, N, dtype=torch.int32, device="cuda")
    rc = lib.nqs_psiformer_fwd_noalloc(ctypes.byref(d), _lib.ptr(packed), _lib.ptr(R), _lib.ptr(S), 1, _lib.ptr(out),
                                       _lib.stream())
    assert rc == _lib.NQS_ERR_WORKSPACE
    assert b"scratch" in lib.nqs_last_error()
