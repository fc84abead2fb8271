"""Graph-captured sweeps (nqs.graphs.GraphSweeps): replays make exactly the decisions of the eager
nqs_mcmc_sweeps call given the same random streams (bit-identical R, S, log|psi|^2), across several
replays with other GPU work and synchronisations in between; a VMC step that replays the graph
equals the eager one."""
import numpy as np
import pytest
import torch

from oracle import bands
from nqs.psiformer import ContinuumSpinorPsiformer, LogPsi, ravel_params
from nqs.sampler import Chain, make_streams
from nqs.graphs import GraphSweeps
from nqs.local_energy import make_local_energy
from nqs import minsr as nms

pytestmark = pytest.mark.gpu


def _model(precision="fp64"):
    N, L = 5, 7.0
    g = bands.rashba_gs(L, N, 1.0, 1.0, 0.5, n_cut=10, twist=(0.1, 0.17))
    m = ContinuumSpinorPsiformer(n_elec=N, Gs=g["Gs"], W=g["W"], L=L, embed_dim=16, depth=2, n_heads=4, n_freq=4,
                                 n_dets=2, precision=precision)
    params = m.init(1)
    kr = params["params"]["backflow"]["kernel"]
    z = torch.tensor(np.random.default_rng(2).standard_normal(tuple(kr.shape)) * 0.05)
    params["params"]["backflow"]["kernel"] = (z + 1j * z).to(torch.complex128)
    return m, ravel_params(params).cuda(), N, L


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replays_match_eager_sweeps(precision):
    m, theta, N, L = _model(precision)
    W = 256
    rng = np.random.default_rng(5)
    R0 = torch.tensor(rng.uniform(0, L, (W, N, 2)), device="cuda")
    S0 = torch.tensor(rng.integers(0, 2, (W, N)), device="cuda", dtype=torch.int32)
    lp = LogPsi(m, theta)
    eager = Chain(lp, R0.clone(), S0.clone())
    graphed = Chain(lp, R0.clone(), S0.clone())
    gs = GraphSweeps(graphed, 4, 0.6, 2)
    E = make_local_energy(lp, 1.0, 1.0, 0.5)
    for rep in range(3):
        st = make_streams(4, W, N, 2, key=100 + rep, device=torch.device("cuda"))
        a_e, s_e = eager.run(4, 0.6, 2, streams=st)
        a_g, s_g = gs.run(streams=st)
        torch.cuda.synchronize()
        assert torch.equal(eager.R, graphed.R) and torch.equal(eager.S, graphed.S)
        assert torch.equal(eager.logp, graphed.logp)
        assert torch.equal(a_e, a_g) and torch.equal(s_e, s_g)
        # unrelated eager work between replays
        e = E(graphed.R, graphed.S)
        torch.cuda.synchronize()
        assert bool(torch.isfinite(e).all())


def test_vmc_step_with_graph_equals_eager():
    m, theta, N, L = _model("fp64")
    W = 128
    rng = np.random.default_rng(6)
    R0 = torch.tensor(rng.uniform(0, L, (W, N, 2)), device="cuda")
    S0 = torch.tensor(rng.integers(0, 2, (W, N)), device="cuda", dtype=torch.int32)
    kw = dict(m_star=1.0, lambda_r=1.0, h_z=0.5, lr=0.05, lam=1e-3, n_sweeps=3, step_size=0.6)
    a = nms.VMCStep(m, theta.clone(), R0.clone(), S0.clone(), **kw)
    b = nms.VMCStep(m, theta.clone(), R0.clone(), S0.clone(), graph=True, **kw)
    for k in range(3):
        sa = a.step(key=k).cpu().numpy()
        sb = b.step(key=k).cpu().numpy()
        assert np.array_equal(sa, sb)
        assert torch.equal(a.theta, b.theta)
