"""The vector-Jacobian product of log psi (nqs_psiformer_vjp, the backward of nqs.autograd's
differentiable log psi) against the per-walker rows O of nqs_psiformer_grad contracted with the
same cotangent, and its memory: it never forms O.

  g = sum_w [cot_w.re d Re log psi_w / d theta + cot_w.im d Im log psi_w / d theta]
    = cot.re^T O_re + cot.im^T O_im                        (src/nqs/minsr.py:56-65, summed)

Architectures with every kind of leaf: K > 1 (det_logw), learnable envelope (dW), RBF and cusp
Jastrows, in float64 (bound 1e-11 relative to max |g|) and in mixed precision (float32 backbone:
within three times the error of the float32 rows against float64); the configuration with N = 8,
d = 64, 3 layers, K = 4, RBF Jastrow and learnable envelope at 4096 walkers, where O would take
6.7 GB in float64 and the product's peak memory must stay below a quarter of it (0.56 GB
measured); and two identical calls bit for bit (slices summed in a fixed order).
"""
import math

import numpy as np
import pytest
import torch

from oracle import bands, psiformer as opf
from nqs.psiformer import ContinuumSpinorPsiformer, LogPsi
from nqs.autograd import vjp_theta
from nqs import minsr as nms

pytestmark = pytest.mark.gpu

ARCHS = {
    "base": dict(embed_dim=16, depth=2, n_heads=4, n_freq=4),
    "kdet_jastrow_env": dict(embed_dim=16, depth=2, n_heads=2, n_freq=3, n_dets=3, use_jastrow=True,
                             learnable_env=True, n_knots=6),
    "cusp": dict(embed_dim=8, depth=1, n_heads=2, n_freq=2, n_dets=2, use_jastrow=True, jastrow_kind="cusp"),
}


def _case(kw, precision, N=5, L=7.0, W=96, seed=3):
    g = bands.rashba_gs(L, N, 1.0, 1.0, 0.5, n_cut=10, twist=(0.1, 0.17))
    arch = opf.Arch(N, g["Gs"], g["W"], L, **kw)
    theta = opf.init_theta(arch, seed, backflow_scale=0.05, scale_all=0.02)
    model = ContinuumSpinorPsiformer(n_elec=N, Gs=g["Gs"], W=g["W"], L=L, precision=precision, **kw)
    rng = np.random.default_rng(seed)
    R = torch.tensor(rng.uniform(0, L, (W, N, 2)), device="cuda")
    S = torch.tensor(rng.integers(0, 2, (W, N)), device="cuda", dtype=torch.int32)
    cot = torch.tensor(rng.normal(size=(W, 2)), device="cuda")
    return LogPsi(model, theta.cuda()), R, S, cot


def _by_rows(lp, R, S, cot):
    O, _ = nms.grad_params(lp, R, S)
    W, P = O.W, O.P
    Od = O.data[:, :P].double()
    return cot[:, 0] @ Od[:W] + cot[:, 1] @ Od[W:]


@pytest.mark.parametrize("name", list(ARCHS))
def test_vjp_equals_contracted_O_rows(name):
    """float64: the product equals the contracted rows to 1e-11 of max |g|.  Mixed precision (float32
    backbone): both float32 evaluations, the product and the contracted float32 rows, are compared
    with the float64 result; the product's error stays within three times the rows' own error."""
    lp, R, S, cot = _case(ARCHS[name], "fp64")
    g = vjp_theta(lp, R, S, cot)
    ref = _by_rows(lp, R, S, cot)
    den = float(ref.abs().max())
    err = float((g - ref).abs().max()) / den
    lpm, _, _, _ = _case(ARCHS[name], "mixed")
    e_vjp = float((vjp_theta(lpm, R, S, cot) - ref).abs().max()) / den
    e_rows = float((_by_rows(lpm, R, S, cot) - ref).abs().max()) / den
    print(f"{name}: float64 |vjp - rows| {err:.2e}; mixed vs float64: vjp {e_vjp:.2e}, rows {e_rows:.2e}", flush=True)
    assert err < 1e-11
    assert e_vjp < max(3.0 * e_rows, 1e-6)


def test_vjp_c2_memory_and_determinism():
    N, RS = 8, 2.0
    L = RS * math.sqrt(math.pi * N)
    kw = dict(embed_dim=64, depth=3, n_heads=4, n_freq=8, n_dets=4, use_jastrow=True, n_knots=16, learnable_env=True)
    lp, R, S, cot = _case(kw, "fp64", N=N, L=L, W=4096, seed=5)
    P = lp.model.n_params()
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats()
    base = torch.cuda.memory_allocated()
    g1 = vjp_theta(lp, R, S, cot, ws_tag="vjp_test")
    torch.cuda.synchronize()
    peak = torch.cuda.max_memory_allocated() - base
    o_bytes = 2 * 4096 * P * 8
    print(f"P = {P}: vjp peak {peak / 2**20:.0f} MiB, O rows would take {o_bytes / 2**20:.0f} MiB", flush=True)
    assert peak < o_bytes / 4
    g2 = vjp_theta(lp, R, S, cot, ws_tag="vjp_test")
    assert torch.equal(g1, g2)
    ref = _by_rows(lp, R[:512], S[:512], cot[:512])
    g3 = vjp_theta(lp, R[:512], S[:512], cot[:512], ws_tag="vjp_test")
    assert float((g3 - ref).abs().max() / ref.abs().max()) < 1e-11
