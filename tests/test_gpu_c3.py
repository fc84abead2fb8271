"""The headline configuration's architecture on the float64 GPU path against the oracle.

N = 16, rs = 3 (L = rs sqrt(pi N)), lambda_R = 0.8, h_z = 0.05, d = 128, 4 layers, 8 heads (head
width 16), K = 8, n_freq = 8 (P = 567,696), plane-wave envelope of the non-interacting ground state
(twist (0.1, 0.17)), a non-trivial network (backflow kernel (1+i) N(0, 0.05^2), every leaf perturbed
by N(0, 0.02^2)), walkers after 20 sweeps.  This runs the exact kernels of the headline workload:
the 16 x 16 register determinant head with K = 8, the register attention jet at head width 16 and the
d = 128 GEMM tiles.

  * log psi of 8 walkers: 1e-10 (real part; imaginary part mod 2 pi)     psiformer.py:118-189
  * E_loc (kinetic + Rashba SOC + Zeeman) of 2 walkers: 1e-8 relative      local_energy.py:63-115
  * Ewald energy of 64 walkers (n_real 4, n_recip 8): 1e-12 relative        ewald.py:47-65
  * O rows of 2 walkers: 1e-9 of max |O|                                     minsr.py:56-65
  * the mixed path (bf16x3 dense layers): E_loc within 1e-2 Ha of the oracle per walker, and the
    per-step energy of 256 walkers within 1e-4 Ha of the float64 path.
"""
import math

import numpy as np
import pytest
import torch

from oracle import psiformer as opf, local_energy as ole, minsr as oms, ewald as oew
from nqs.rashba_reference import noninteracting_gs, occupied_orbital_matrix, planewave_envelope
from nqs.psiformer import ContinuumSpinorPsiformer, LogPsi
from nqs.local_energy import make_local_energy
from nqs.sampler import Chain
from nqs.ewald import make_ewald
from nqs import minsr as nms

pytestmark = pytest.mark.gpu

N_E, RS, LAM, HZ = 16, 3.0, 0.8, 0.05
ARCH = dict(embed_dim=128, depth=4, n_heads=8, n_freq=8, n_dets=8)


@pytest.fixture(scope="module")
def c3():
    L = RS * math.sqrt(math.pi * N_E)
    g = noninteracting_gs(N_E, L, 1.0, LAM, HZ, n_cut=10, twist=(0.1, 0.17))
    Gs, Wenv = planewave_envelope(*occupied_orbital_matrix(g))
    arch = opf.Arch(N_E, Gs, Wenv, L, **ARCH)
    theta = opf.init_theta(arch, 0, backflow_scale=0.05, scale_all=0.02)
    m64 = ContinuumSpinorPsiformer(n_elec=N_E, Gs=Gs, W=Wenv, L=L, precision="fp64", **ARCH)
    W = 256
    rng = np.random.default_rng(5)
    R = torch.tensor(rng.uniform(0, L, (W, N_E, 2)), device="cuda")
    S = torch.tensor(rng.integers(0, 2, (W, N_E)), device="cuda", dtype=torch.int32)
    lp = LogPsi(m64, theta.cuda())
    ch = Chain(lp, R, S)
    ch.run(20, 0.3 * RS, 2, key=31)
    return dict(L=L, Gs=Gs, Wenv=Wenv, arch=arch, theta=theta, m64=m64, lp=lp, R=ch.R.clone(), S=ch.S.clone())


def _cpu(c3, w):
    return torch.tensor(c3["R"][w].cpu().numpy()), torch.tensor(c3["S"][w].cpu().numpy())


def test_c3_parameter_count(c3):
    assert c3["m64"].n_params() == 567696 == c3["theta"].shape[0]


def test_c3_logpsi_matches_oracle(c3):
    v = c3["lp"](c3["R"][:8], c3["S"][:8]).cpu().numpy()
    for w in range(8):
        ref = complex(opf.logpsi(c3["arch"], c3["theta"], *_cpu(c3, w)))
        assert abs(v[w].real - ref.real) < 1e-10
        assert abs((v[w].imag - ref.imag + np.pi) % (2 * np.pi) - np.pi) < 1e-10


@pytest.fixture(scope="module")
def oracle_eloc(c3):
    Eo = ole.make_local_energy(opf.make_logpsi(c3["arch"], c3["theta"]), 1.0, LAM, HZ)
    return np.array([Eo(*_cpu(c3, w)) for w in range(2)])


def test_c3_eloc_matches_oracle(c3, oracle_eloc):
    E = make_local_energy(c3["lp"], 1.0, LAM, HZ)(c3["R"][:2], c3["S"][:2]).cpu().numpy()
    for w in range(2):
        print(f"walker {w}: E_loc gpu {E[w]:.10f} oracle {oracle_eloc[w]:.10f}", flush=True)
        assert abs(E[w] - oracle_eloc[w]) < 1e-8 * max(1.0, abs(oracle_eloc[w]))


def test_c3_ewald_matches_oracle(c3):
    _, _, ce = make_ewald(c3["L"], 1.0, n_real=4, n_recip=8)
    _, _, ceo = oew.make_ewald(c3["L"], 1.0, n_real=4, n_recip=8)
    e = ce(c3["R"][:64]).cpu().numpy()
    ref = np.array([ceo(r) for r in c3["R"][:64].cpu().numpy()])
    assert np.max(np.abs(e - ref)) < 1e-12 * np.max(np.abs(ref))


def test_c3_O_rows_match_oracle_autodiff(c3):
    R, S = c3["R"][:2], c3["S"][:2]
    O = nms.per_sample_O(c3["m64"], c3["theta"].cuda(), None, R, S).cpu().numpy()
    Oo = oms.per_sample_O(c3["arch"], c3["theta"], R.cpu().numpy(), S.cpu().numpy()).numpy()
    assert np.max(np.abs(O - Oo)) < 1e-9 * np.max(np.abs(Oo))


def test_c3_mixed_path_against_oracle(c3, oracle_eloc):
    """Per walker the mixed E_loc is within 1e-2 Ha of the oracle (bf16x3 rounding of the dense
    layers); the per-step energy over all 256 walkers is within 1e-4 Ha of the float64 GPU path,
    which itself matches the oracle to 1e-8 relative per walker (test above)."""
    m = ContinuumSpinorPsiformer(n_elec=N_E, Gs=c3["Gs"], W=c3["Wenv"], L=c3["L"], precision="mixed", **ARCH)
    lp = LogPsi(m, c3["theta"].cuda())
    e = make_local_energy(lp, 1.0, LAM, HZ)(c3["R"], c3["S"])
    e64 = make_local_energy(c3["lp"], 1.0, LAM, HZ)(c3["R"], c3["S"])
    dw = np.abs(e[:2].cpu().numpy() - oracle_eloc)
    dE = abs(float(e.real.mean() - e64.real.mean()))
    print(f"c3 mixed: |E_loc - oracle| on 2 walkers {dw.max():.3e} Ha; per-step energy vs float64 over 256 "
          f"walkers {dE:.3e} Ha", flush=True)
    assert dw.max() < 1e-2
    assert dE < 1e-4
