"""Mixed-precision accuracy at the headline architecture (N = 16, rs = 3, lambda = 0.8,
h_z = 0.05, d = 128, 4 layers, 8 heads, K = 8, F = 8) against the float64 GPU path, which the
other GPU tests pin to the oracle (log psi 1e-10, E_loc 1e-8).

Walkers are drawn from |psi|^2 by the float64 chain (40 sweeps from uniform positions); then,
on the same walkers with the same parameters, each mixed dense-layer mode (bf16x3, fp32, bf16):
  per-step energy  |mean Re E_loc(mixed) - mean Re E_loc(fp64)|  <= 1e-4 Ha (0.1 mHa) for
  bf16x3 and fp32, log psi (real part) within 2e-3.  Plain bf16 (operands keep 8 significant
  bits) misses the 0.1 mHa bar (measured at 512 walkers: about 2.4e-3 Ha per step, E_loc rms
  0.28 Ha, log psi up to 6e-2); its test only bounds it loosely (1e-2 Ha).
Also checks the dense layers of every mode on small shapes (K not a multiple of the k tile,
rows and columns not multiples of the block tile) through the logpsi parity of a tiny model.
"""
import math

import numpy as np
import pytest
import torch

from nqs.rashba_reference import noninteracting_gs, occupied_orbital_matrix, planewave_envelope
from nqs.psiformer import ContinuumSpinorPsiformer, ravel_params, LogPsi
from nqs.local_energy import make_local_energy
from nqs.sampler import Chain

pytestmark = pytest.mark.gpu

N_E, RS, LAM, HZ = 16, 3.0, 0.8, 0.05
ARCH = dict(embed_dim=128, depth=4, n_heads=8, n_freq=8, n_dets=8)


def _system(W, seed=0):
    L = RS * math.sqrt(math.pi * N_E)
    g = noninteracting_gs(N_E, L, 1.0, LAM, HZ, n_cut=10, twist=(0.1, 0.17))
    Gs, Wenv = planewave_envelope(*occupied_orbital_matrix(g))
    m = ContinuumSpinorPsiformer(n_elec=N_E, Gs=Gs, W=Wenv, L=L, precision="fp64", **ARCH)
    params = m.init(seed)
    rng = np.random.default_rng(seed + 1)
    kr = params["params"]["backflow"]["kernel"]
    z = torch.tensor(rng.standard_normal(tuple(kr.shape)) * 0.05)
    params["params"]["backflow"]["kernel"] = (z + 1j * z).to(torch.complex128)
    theta = ravel_params(params).cuda()
    R = torch.tensor(rng.uniform(0, L, (W, N_E, 2)), device="cuda")
    S = torch.tensor(rng.integers(0, 2, (W, N_E)), device="cuda", dtype=torch.int32)
    return dict(L=L, Gs=Gs, Wenv=Wenv, theta=theta, R=R, S=S, m64=m)


@pytest.fixture(scope="module")
def c3():
    sysd = _system(512)
    lp = LogPsi(sysd["m64"], sysd["theta"])
    ch = Chain(lp, sysd["R"], sysd["S"])
    ch.run(40, 0.3 * RS, 2, key=11)
    sysd["R"], sysd["S"] = ch.R, ch.S
    sysd["lp64"] = lp(sysd["R"], sysd["S"])
    sysd["e64"] = make_local_energy(lp, 1.0, LAM, HZ)(sysd["R"], sysd["S"])
    return sysd


@pytest.mark.parametrize("gemm,tol_e,tol_lp", [("bf16x3", 1e-4, 2e-3), ("fp32", 1e-4, 2e-3), ("bf16", 1e-2, None)])
def test_step_energy_mixed_vs_fp64(c3, gemm, tol_e, tol_lp):
    m = ContinuumSpinorPsiformer(n_elec=N_E, Gs=c3["Gs"], W=c3["Wenv"], L=c3["L"], precision="mixed", gemm=gemm,
                                 **ARCH)
    lp = LogPsi(m, c3["theta"])
    lv = lp(c3["R"], c3["S"])
    e = make_local_energy(lp, 1.0, LAM, HZ)(c3["R"], c3["S"])
    assert torch.isfinite(e).all()
    dE = abs(float(e.real.mean() - c3["e64"].real.mean()))
    de = (e - c3["e64"]).abs()
    print(f"gemm={gemm}: dE_step={dE:.3e} Ha, max|dlogpsi.re|={float((lv - c3['lp64']).real.abs().max()):.3e}, "
          f"dE_loc rms={float(de.pow(2).mean().sqrt()):.3e} median={float(de.median()):.3e} "
          f"max={float(de.max()):.3e}, |E_loc| median={float(c3['e64'].abs().median()):.3e}", flush=True)
    assert dE < tol_e
    if tol_lp is not None:
        assert float((lv - c3["lp64"]).real.abs().max()) < tol_lp


@pytest.mark.parametrize("gemm", ["bf16x3", "fp32", "bf16"])
def test_small_shapes_each_gemm_mode(gemm):
    """Odd sizes: d = 24 (not a multiple of 32), 3 heads, F = 3 (34 -> 16 padded inputs), N = 7."""
    N, L = 7, 8.0
    g = noninteracting_gs(N, L, 1.0, 1.0, 0.5, n_cut=10, twist=(0.1, 0.17))
    Gs, Wenv = planewave_envelope(*occupied_orbital_matrix(g))
    kw = dict(embed_dim=24, depth=2, n_heads=3, n_freq=3, n_dets=2)
    m64 = ContinuumSpinorPsiformer(n_elec=N, Gs=Gs, W=Wenv, L=L, precision="fp64", **kw)
    params = m64.init(5)
    kr = params["params"]["backflow"]["kernel"]
    z = torch.tensor(np.random.default_rng(2).standard_normal(tuple(kr.shape)) * 0.05)
    params["params"]["backflow"]["kernel"] = (z + 1j * z).to(torch.complex128)
    theta = ravel_params(params).cuda()
    rng = np.random.default_rng(9)
    R = torch.tensor(rng.uniform(0, L, (37, N, 2)), device="cuda")
    S = torch.tensor(rng.integers(0, 2, (37, N)), device="cuda", dtype=torch.int32)
    ref = LogPsi(m64, theta)(R, S)
    m = ContinuumSpinorPsiformer(n_elec=N, Gs=Gs, W=Wenv, L=L, precision="mixed", gemm=gemm, **kw)
    v = LogPsi(m, theta)(R, S)
    tol = 2e-1 if gemm == "bf16" else 2e-4
    err = float((v - ref).real.abs().max())
    print(f"gemm={gemm}: small-shape max|dlogpsi.re|={err:.3e}", flush=True)
    assert err < tol


def test_batch_size_does_not_change_the_bits():
    """Every walker's log psi is a function of that walker alone: a large batch (one launch per
    layer over 131k rows) equals, bit for bit, the same walkers evaluated in batches of 1024."""
    sysd = _system(8192, seed=4)
    m = ContinuumSpinorPsiformer(n_elec=N_E, Gs=sysd["Gs"], W=sysd["Wenv"], L=sysd["L"], precision="mixed", **ARCH)
    lp = LogPsi(m, sysd["theta"])
    big = lp(sysd["R"], sysd["S"])
    small = torch.cat([lp(sysd["R"][i:i + 1024], sysd["S"][i:i + 1024]) for i in range(0, 8192, 1024)])
    assert torch.equal(big, small)


def test_value_forward_equals_laplacian_pass_value_channel():
    """The value-only forward (mixed precision: GELU applied inside the Dense_1 GEMM's A staging,
    value-only determinant head) against the value channel of the local energy's Laplacian pass
    (separate GELU jet kernel, Laplacian determinant head) on the same walkers: the same float32
    arithmetic in a different kernel arrangement, so log psi agrees to float32 rounding."""
    import math
    from nqs.rashba_reference import noninteracting_gs, occupied_orbital_matrix, planewave_envelope
    from nqs.psiformer import ContinuumSpinorPsiformer, LogPsi, ravel_params
    from nqs.local_energy import make_local_energy
    N, RS = 16, 3.0
    L = RS * math.sqrt(math.pi * N)
    g = noninteracting_gs(N, L, 1.0, 0.8, 0.05, n_cut=10, twist=(0.1, 0.17))
    Gs, Wenv = planewave_envelope(*occupied_orbital_matrix(g))
    m = ContinuumSpinorPsiformer(n_elec=N, Gs=Gs, W=Wenv, L=L, precision="mixed", embed_dim=128, depth=4, n_heads=8,
                                 n_freq=8, n_dets=8)
    params = m.init(1)
    kr = params["params"]["backflow"]["kernel"]
    z = torch.tensor(np.random.default_rng(3).standard_normal(tuple(kr.shape)) * 0.05)
    params["params"]["backflow"]["kernel"] = (z + 1j * z).to(torch.complex128)
    lp = LogPsi(m, ravel_params(params).cuda())
    rng = np.random.default_rng(6)
    R = torch.tensor(rng.uniform(0, L, (256, N, 2)), device="cuda")
    S = torch.tensor(rng.integers(0, 2, (256, N)), device="cuda", dtype=torch.int32)
    v = lp(R, S)
    _, lv, _, _ = make_local_energy(lp, 1.0, 0.8, 0.05).parts(R, S)
    d = (v - lv)
    dre = float(d.real.abs().max())
    dim = float((torch.remainder(d.imag + math.pi, 2 * math.pi) - math.pi).abs().max())
    print(f"value forward vs Laplacian-pass value channel: max |d Re| {dre:.2e}, |d Im| {dim:.2e}", flush=True)
    assert dre < 1e-4 and dim < 1e-4
