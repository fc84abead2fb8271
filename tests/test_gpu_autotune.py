"""Launch-shape autotuner: tuned launches give bit-identical results to the default shapes (each
(config, head) pair and each determinant is computed by one wave whatever the workgroup size, and
every tile shape of the bf16 dense-layer GEMM accumulates each output in the same order), the
choices are recorded per shape class (kernel 1 attention, 2 determinant head, 4 GEMM tiles), and
disabling the tuner falls back to the defaults."""
import numpy as np
import pytest
import torch

from oracle import bands
from nqs import _lib
from nqs.psiformer import ContinuumSpinorPsiformer, LogPsi, ravel_params
from nqs.local_energy import make_local_energy

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autotuned_launches_match_default_shapes(precision):
    N, L = 7, 8.0
    g = bands.rashba_gs(L, N, 1.0, 1.0, 0.5, n_cut=10, twist=(0.1, 0.17))
    m = ContinuumSpinorPsiformer(n_elec=N, Gs=g["Gs"], W=g["W"], L=L, embed_dim=32, depth=2, n_heads=4, n_freq=4,
                                 n_dets=3, precision=precision)
    params = m.init(3)
    kr = params["params"]["backflow"]["kernel"]
    z = torch.tensor(np.random.default_rng(2).standard_normal(tuple(kr.shape)) * 0.05)
    params["params"]["backflow"]["kernel"] = (z + 1j * z).to(torch.complex128)
    lp = LogPsi(m, ravel_params(params).cuda())
    rng = np.random.default_rng(4)
    R = torch.tensor(rng.uniform(0, L, (300, N, 2)), device="cuda")
    S = torch.tensor(rng.integers(0, 2, (300, N)), device="cuda", dtype=torch.int32)
    E = make_local_energy(lp, 1.0, 1.0, 0.5)
    lib = _lib.load()
    lib.nqs_autotune_enable(0)
    ref = [t.clone() for t in E.parts(R, S)]
    lib.nqs_autotune_reset()
    lib.nqs_autotune_enable(1)
    first = E.parts(R, S)              # tunes the new shape classes
    second = E.parts(R, S)             # uses the recorded choices
    rep = _lib.autotune_report()
    print(rep)
    assert "kernel 1" in rep and "kernel 2" in rep
    if precision == "mixed":
        assert "kernel 4" in rep and "128x128" in rep
    for a, b, c in zip(ref, first, second):
        assert torch.equal(a, b) and torch.equal(a, c)
