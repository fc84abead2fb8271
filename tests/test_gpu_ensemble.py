"""The batched uneven-N ensemble (no reference counterpart; each system follows the reference's
single-system step).

* nqs_minsr_direction_batched gives, for every system, exactly what nqs_minsr_direction gives
  for that system alone (bf16 Gram with a float32 O, and the float64 path).
* The ensemble's one-graph sweeps (side-stream branches) make exactly the decisions of eager
  per-system sweeps with the same keys.
* The 24-system configuration (N in {4 .. 32}, rs in [1, 6], lambda_R in [0, 1], h_z in [0, 0.2],
  d = 64, 3 layers, 4 heads, K = 4, 1024 walkers each): one full step, every solve succeeds; the
  throughput of a second step is printed.
"""
import time

import numpy as np
import pytest
import torch

from nqs import ensemble as nen
from nqs import minsr as nms

pytestmark = pytest.mark.gpu

SMALL = dict(embed_dim=16, depth=1, n_heads=2, n_freq=2, n_dets=2)
SYS3 = [dict(N=4, rs=1.5, lambda_r=0.6, h_z=0.1), dict(N=6, rs=2.5, lambda_r=0.3, h_z=0.0),
        dict(N=8, rs=2.0, lambda_r=0.9, h_z=0.05)]


@pytest.mark.parametrize("precision,gram", [("mixed", "bf16"), ("fp64", "fp64")])
def test_batched_minsr_equals_single_system(precision, gram):
    ens = nen.Ensemble(SYS3, arch=SMALL, n_walkers=128, precision=precision, graph=False, seed=3)
    ens.burn(4)
    Os, eps = [], []
    for s in ens.systems:
        e = s.E(s.chain.R, s.chain.S) + s.coulomb(s.chain.R)
        ec, _ = nms.clip_energies(e, 5.0)
        O, _ = nms.grad_params(s.lp, s.chain.R, s.chain.S)
        Os.append(O)
        eps.append(ec)
    single = []
    for O, e in zip(Os, eps):
        Oc = nms.OMatrix(O.data.clone(), O.W, O.P)
        single.append(nms.minsr_direction(Oc, e, 1e-3, gram, 1).clone())
        assert int(nms.minsr_direction.last_info.item()) == 0
    dth, info = nen.minsr_direction_batched(Os, eps, [1e-3] * 3, gram, 1)
    assert torch.equal(info, torch.zeros_like(info))
    for a, b in zip(single, dth):
        assert torch.equal(a, b)


def test_ensemble_graph_sweeps_equal_eager():
    a = nen.Ensemble(SYS3, arch=SMALL, n_walkers=256, graph=True, seed=5)
    b = nen.Ensemble(SYS3, arch=SMALL, n_walkers=256, graph=False, seed=5)
    for k in range(3):
        ra = a.sweep(key=k)
        rb = b.sweep(key=k)
        torch.cuda.synchronize()
        assert ra == rb
        for sa, sb in zip(a.systems, b.systems):
            assert torch.equal(sa.chain.R, sb.chain.R) and torch.equal(sa.chain.S, sb.chain.S)
            assert torch.equal(sa.chain.logp, sb.chain.logp)


def test_c5_ensemble_step():
    systems = nen.make_c5_systems(24, seed=0)
    ens = nen.Ensemble(systems, n_walkers=1024, precision="mixed", graph=True, seed=0)
    print("systems:", [(s["N"], round(s["rs"], 2)) for s in systems], flush=True)
    ens.burn(5)
    st = ens.step(key=1)
    assert np.all(np.isfinite(st))
    assert int(ens.last_info.abs().sum()) == 0 and ens.failed.sum() == 0
    torch.cuda.synchronize()
    t = time.perf_counter()
    tm = []
    st = ens.step(key=2, timers=tm)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    print(f"c5 step: {1e3 * dt:.1f} ms, {24 * 1024 / dt:.0f} walker-steps/s over 24 systems; stage ms "
          f"{ {k: round(v, 1) for k, v in tm[0].items()} }", flush=True)
    assert np.all(np.isfinite(st))
