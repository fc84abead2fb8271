"""The fused value-only forward (one launch from the Fourier features to the backflow activations,
csrc/fused.hip) against the layer-by-layer forward of the same model and against the oracle.

Headline architecture: N = 16, d = 128, 4 layers, 8 heads (head width 16), K = 8, n_freq = 8, mixed
precision with bf16x3 dense layers, a non-trivial network (backflow kernel (1+i) N(0, 0.05^2), every
leaf perturbed by N(0, 0.02^2)).  NQS_LAYERWISE_FORWARD=1 switches the library back to the
layer-by-layer kernels for the comparison (psiformer.py:118-189).

  * log psi of 4096 walkers and of 33 walkers (the last workgroup holds a single configuration):
    bit for bit the layer-by-layer forward's (the same float32 operations in the same order: the
    dense blocks accumulate as k_gemm_nt_bf does, and the LayerNorm, GELU and softmax rules are
    written without FMA contraction, so they round alike in every kernel that inlines them);
  * log psi of 8 walkers within 5e-4 of the float64 oracle (bf16x3 rounding of the dense layers);
  * the Metropolis sweeps over the fused forward make the same accept/reject decisions as over the
    layer-by-layer forward, given the same supplied streams;
  * the spin-flip pass of E_loc (value and the two position tangents of every single-spin-flipped
    configuration, src/nqs/local_energy.py:63-115) on the fused first-order kernel: E_loc bit for bit
    the layer-by-layer jets' (the same dense blocks, and the LayerNorm and attention tangent rules
    written without FMA contraction in both paths).
"""
import math
import os
import time

import numpy as np
import pytest
import torch

from oracle import psiformer as opf
from nqs.rashba_reference import noninteracting_gs, occupied_orbital_matrix, planewave_envelope
from nqs.psiformer import ContinuumSpinorPsiformer, LogPsi
from nqs.sampler import Chain
from nqs.local_energy import make_local_energy

pytestmark = pytest.mark.gpu

N_E, RS, LAM, HZ = 16, 3.0, 0.8, 0.05
ARCH = dict(embed_dim=128, depth=4, n_heads=8, n_freq=8, n_dets=8)
ENV = "NQS_LAYERWISE_FORWARD"


@pytest.fixture(scope="module")
def c3m():
    L = RS * math.sqrt(math.pi * N_E)
    g = noninteracting_gs(N_E, L, 1.0, LAM, HZ, n_cut=10, twist=(0.1, 0.17))
    Gs, Wenv = planewave_envelope(*occupied_orbital_matrix(g))
    arch = opf.Arch(N_E, Gs, Wenv, L, **ARCH)
    theta = opf.init_theta(arch, 0, backflow_scale=0.05, scale_all=0.02)
    m = ContinuumSpinorPsiformer(n_elec=N_E, Gs=Gs, W=Wenv, L=L, precision="mixed", **ARCH)
    lp = LogPsi(m, theta.cuda())
    rng = np.random.default_rng(11)
    R = torch.tensor(rng.uniform(0, L, (4096, N_E, 2)), device="cuda")
    S = torch.tensor(rng.integers(0, 2, (4096, N_E)), device="cuda", dtype=torch.int32)
    return dict(L=L, arch=arch, theta=theta, lp=lp, R=R, S=S)


def _layerwise(fn):
    os.environ[ENV] = "1"
    try:
        return fn()
    finally:
        os.environ.pop(ENV, None)


def _wrap(d):
    return torch.remainder(d + math.pi, 2 * math.pi) - math.pi


@pytest.mark.parametrize("W", [4096, 33])
def test_fused_forward_equals_layerwise(c3m, W):
    lp, R, S = c3m["lp"], c3m["R"][:W], c3m["S"][:W]
    v = lp(R, S)
    v_lw = _layerwise(lambda: lp(R, S))
    dre = float((v.real - v_lw.real).abs().max())
    dim = float(_wrap(v.imag - v_lw.imag).abs().max())
    exact = bool(torch.equal(v, v_lw))
    print(f"fused vs layer-by-layer, {W} walkers: bit-identical {exact}, max |d Re| {dre:.3e}, |d Im| {dim:.3e}",
          flush=True)
    assert exact, (dre, dim)


def test_fused_forward_against_oracle(c3m):
    v = c3m["lp"](c3m["R"][:8], c3m["S"][:8]).cpu().numpy()
    for w in range(8):
        ref = complex(opf.logpsi(c3m["arch"], c3m["theta"], torch.tensor(c3m["R"][w].cpu().numpy()),
                                 torch.tensor(c3m["S"][w].cpu().numpy())))
        assert abs(v[w].real - ref.real) < 5e-4
        assert abs((v[w].imag - ref.imag + np.pi) % (2 * np.pi) - np.pi) < 5e-4


def test_fused_sweeps_match_layerwise_decisions(c3m):
    lp, L = c3m["lp"], c3m["L"]
    R0, S0 = c3m["R"][:512].clone(), c3m["S"][:512].clone()

    def run():
        ch = Chain(lp, R0.clone(), S0.clone())
        acc = ch.run(5, 0.3 * RS, 2, key=77)
        return ch.R.clone(), ch.S.clone(), [a.clone() if torch.is_tensor(a) else a for a in acc]
    Rf, Sf, _ = run()
    Rl, Sl, _ = _layerwise(run)
    assert torch.equal(Sf, Sl)
    assert float((Rf - Rl).abs().max()) == 0.0


def test_fused_forward_timing(c3m):
    """Informational: the value-only forward of 8192 walkers, fused and layer by layer."""
    lp = c3m["lp"]
    R = torch.cat([c3m["R"], c3m["R"]]).contiguous()
    S = torch.cat([c3m["S"], c3m["S"]]).contiguous()

    def t(n=5):
        lp(R, S)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            lp(R, S)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n * 1e3
    tf = t()
    tl = _layerwise(t)
    print(f"value-only forward, 8192 walkers: fused {tf:.3f} ms, layer by layer {tl:.3f} ms", flush=True)


@pytest.mark.parametrize("W", [1024, 7])
def test_fused_spin_flip_eloc_equals_layerwise(c3m, W):
    """E_loc with the Rashba term (spin-flip pass fused) against the layer-by-layer jets."""
    lp = c3m["lp"]
    R, S = c3m["R"][:W], c3m["S"][:W]
    E = make_local_energy(lp, 1.0, LAM, HZ)
    e = E(R, S)
    e_lw = _layerwise(lambda: E(R, S))
    d = (e - e_lw).abs()
    scale = float(e_lw.abs().median())
    print(f"spin-flip pass fused vs layer-by-layer, {W} walkers: max |dE_loc| {float(d.max()):.3e}, "
          f"median |dE_loc| {float(d.median()):.3e}, median |E_loc| {scale:.3e}", flush=True)
    assert bool(torch.isfinite(e).all())
    assert torch.equal(e, e_lw)


def test_fused_spin_flip_timing(c3m):
    """Informational: E_loc of 4096 walkers with the spin-flip pass fused and layer by layer."""
    lp = c3m["lp"]
    E = make_local_energy(lp, 1.0, LAM, HZ)
    R, S = c3m["R"], c3m["S"]

    def t(n=3):
        E(R, S)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            E(R, S)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n * 1e3
    tf = t()
    tl = _layerwise(t)
    print(f"E_loc, 4096 walkers: spin-flip pass fused {tf:.2f} ms, layer by layer {tl:.2f} ms", flush=True)
