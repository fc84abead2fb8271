"""GPU parity of the hand-written backward (per-walker O) and of the MinSR solve.

O: against the oracle's reverse-mode autodiff (torch.func.grad of Re and Im log psi, the
restatement of minsr.py:56-65) to 1e-9 relative, and against central finite differences
(h = 1e-5) of the kernel's own float64 log psi.  MinSR: against the oracle's sample-space LU
solve and the explicit parameter-space solve (minsr.py:76-88, tests/test_nqs_minsr.py:46-61,
rel < 1e-6); bf16 / fp32 Gram variants against the float64 one with their stated tolerances.
"""
import numpy as np
import pytest
import torch

from oracle import bands, psiformer as opf, minsr as oms
from nqs.psiformer import ContinuumSpinorPsiformer, make_logpsi
from nqs import minsr as nms

pytestmark = pytest.mark.gpu

ARCHS = {
    "base": dict(embed_dim=16, depth=2, n_heads=4, n_freq=4),
    "kdet_jastrow_env": dict(embed_dim=16, depth=2, n_heads=2, n_freq=3, n_dets=3, use_jastrow=True,
                             learnable_env=True, n_knots=6),
    "cusp": dict(embed_dim=8, depth=1, n_heads=2, n_freq=2, n_dets=2, use_jastrow=True, jastrow_kind="cusp"),
    "cusp_periodic": dict(embed_dim=8, depth=1, n_heads=2, n_freq=2, n_dets=1, use_jastrow=True,
                          jastrow_kind="cusp_periodic"),
}


def _pair(name, precision="fp64", N=5, L=7.0, seed=3):
    g = bands.rashba_gs(L, N, 1.0, 1.0, 0.5, n_cut=10, twist=(0.1, 0.17))
    kw = ARCHS[name]
    arch = opf.Arch(N, g["Gs"], g["W"], L, **kw)
    theta = opf.init_theta(arch, seed, backflow_scale=0.05, scale_all=0.02)
    model = ContinuumSpinorPsiformer(n_elec=N, Gs=g["Gs"], W=g["W"], L=L, precision=precision, **kw)
    return g, arch, theta, model


def _configs(W, N, L, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0, L, (W, N, 2)), rng.integers(0, 2, (W, N))


@pytest.mark.parametrize("name", list(ARCHS))
def test_O_matches_oracle_autodiff(name):
    g, arch, theta, model = _pair(name)
    R, S = _configs(5, 5, 7.0, 21)
    O = nms.per_sample_O(model, theta.cuda(), None, R, S).cpu().numpy()
    Oo = oms.per_sample_O(arch, theta, R, S).numpy()
    assert O.shape == Oo.shape
    scale = np.max(np.abs(Oo))
    assert np.max(np.abs(O - Oo)) < 1e-9 * scale


def test_O_matches_central_finite_differences():
    g, arch, theta, model = _pair("kdet_jastrow_env")
    R, S = _configs(2, 5, 7.0, 22)
    O = nms.per_sample_O(model, theta.cuda(), None, R, S).cpu().numpy()
    P = theta.shape[0]
    h = 1e-5
    for k in list(range(0, P, max(1, P // 40))) + [P - 1]:
        e = torch.zeros_like(theta)
        e[k] = h
        fp = make_logpsi(model, (theta + e).cuda())(R, S).cpu().numpy()
        fm = make_logpsi(model, (theta - e).cuda())(R, S).cpu().numpy()
        fd = (fp - fm) / (2 * h)
        assert np.max(np.abs(fd - O[:, k])) < 2e-6 * max(1.0, np.max(np.abs(fd)))


def test_mixed_precision_O_close_to_fp64():
    g, arch, theta, m64 = _pair("base")
    _, _, _, m32 = _pair("base", precision="mixed")
    R, S = _configs(6, 5, 7.0, 23)
    O64 = nms.per_sample_O(m64, theta.cuda(), None, R, S).cpu().numpy()
    O32 = nms.per_sample_O(m32, theta.cuda(), None, R, S).cpu().numpy()
    assert np.max(np.abs(O32 - O64)) < 1e-4 * np.max(np.abs(O64))


def _minsr_case(W=32, seed=4):
    g, arch, theta, model = _pair("base", seed=seed)
    R, S = _configs(W, 5, 7.0, seed)
    O = nms.per_sample_O(model, theta.cuda(), None, R, S)
    rng = np.random.default_rng(seed)
    eps = torch.tensor(rng.normal(size=W) + 1j * rng.normal(size=W))
    return O, eps


@pytest.mark.parametrize("lam", [1e-2, 1e-4])
def test_minsr_matches_oracle_and_paramsr(lam):
    O, eps = _minsr_case()
    d = nms.minsr_direction(O, eps, lam, gram="fp64").cpu()
    assert int(nms.minsr_direction.last_info.item()) == 0
    d_lu = oms.minsr_direction(O.cpu(), eps, lam)
    d_ps = oms.paramsr_direction(O.cpu(), eps, lam)
    den = float(torch.max(torch.abs(d_ps)))
    assert float(torch.max(torch.abs(d - d_lu))) / den < 1e-6
    assert float(torch.max(torch.abs(d - d_ps))) / den < 1e-6
    d_gpu_ps = nms.paramsr_direction(O, eps, lam).cpu()
    assert float(torch.max(torch.abs(d_gpu_ps - d_ps))) / den < 1e-6


@pytest.mark.parametrize("W", [160, 224, 416])
def test_minsr_short_last_outer_panel(W):
    """2W = 320, 448, 832: the Cholesky's last 512-column outer panel is short (five, seven and
    five 64-column blocks past the last full panel), and the chained triangular solves run 5, 7
    and 13 blocks.  The float64 direction against the oracle's LU solve."""
    O, eps = _minsr_case(W=W, seed=6)
    d = nms.minsr_direction(O, eps, 1e-4, gram="fp64").cpu()
    assert int(nms.minsr_direction.last_info.item()) == 0
    d_lu = oms.minsr_direction(O.cpu(), eps, 1e-4)
    den = float(torch.max(torch.abs(d_lu)))
    assert float(torch.max(torch.abs(d - d_lu))) / den < 1e-6


@pytest.mark.parametrize("gram,tol,W", [("fp32", 5e-4, 64), ("bf16", 5e-3, 64), ("fp32", 5e-4, 384),
                                        ("bf16", 5e-3, 384)])
def test_low_precision_gram_close_to_fp64(gram, tol, W):
    """bf16 / f32 MFMA Gram (f32 accumulate) + float64 Cholesky + one refinement step from the
    float32 O: update vector within tol (relative, max norm) of the float64 MinSR; the bounds are
    about ten times the measured errors (fp32 2.2e-5 / 4.8e-5, bf16 4.1e-5 / 4.8e-4).  W = 384
    (2W = 768 = 3 x 256) takes the 256-tile bf16 Gram on the blocked operand layout."""
    g, arch, theta, m64 = _pair("base", seed=5)
    _, _, _, m32 = _pair("base", precision="mixed", seed=5)
    R, S = _configs(W, 5, 7.0, 5)
    eps = torch.tensor(np.random.default_rng(5).normal(size=W) + 0j)
    O64, _ = nms.grad_params(make_logpsi(m64, theta.cuda()), R, S)
    O32, _ = nms.grad_params(make_logpsi(m32, theta.cuda()), R, S)
    d64 = nms.minsr_direction(O64, eps, 1e-3, gram="fp64").cpu()
    d = nms.minsr_direction(O32, eps, 1e-3, gram=gram, refine=1).cpu()
    rel = float(torch.max(torch.abs(d - d64)) / torch.max(torch.abs(d64)))
    print(f"{gram} Gram, W={W}: dtheta rel {rel:.3e} (bound {tol:.0e})", flush=True)
    assert rel < tol, rel


@pytest.mark.parametrize("gram", ["fp64", "fp32", "bf16"])
def test_minsr_direction_is_deterministic(gram):
    """Two identical calls give bit-identical directions: every reduction of the solve (column sums,
    Gram tail, centring deflation, Cholesky, triangular solves, refinement) sums in a fixed order.
    W = 1024 gives the deflation's column part four 256-row slabs per column."""
    g, arch, theta, m = _pair("base", precision="fp64" if gram == "fp64" else "mixed", seed=7)
    R, S = _configs(1024, 5, 7.0, 7)
    eps = torch.tensor(np.random.default_rng(7).normal(size=1024) + 1j * np.random.default_rng(8).normal(size=1024))
    O, _ = nms.grad_params(make_logpsi(m, theta.cuda()), R, S)
    base = O.data.clone()
    d1 = nms.minsr_direction(O, eps, 1e-3, gram=gram)
    O.data.copy_(base)
    d2 = nms.minsr_direction(O, eps, 1e-3, gram=gram)
    assert int(nms.minsr_direction.last_info.item()) == 0
    assert torch.equal(d1, d2)


@pytest.mark.parametrize("gram", ["fp64", "fp32", "bf16"])
def test_split_minsr_calls_equal_direction(gram):
    """nqs_minsr_gram + nqs_minsr_factor + nqs_minsr_solve (the factorisation on another stream, ordered by
    stream waits, as VMCStep runs it beside the local energies) give nqs_minsr_direction's bits."""
    g, arch, theta, m = _pair("base", precision="fp64" if gram == "fp64" else "mixed", seed=9)
    R, S = _configs(512, 5, 7.0, 9)
    eps = torch.tensor(np.random.default_rng(9).normal(size=512) + 1j * np.random.default_rng(10).normal(size=512))
    O, _ = nms.grad_params(make_logpsi(m, theta.cuda()), R, S)
    base = O.data.clone()
    d1 = nms.minsr_direction(O, eps, 1e-3, gram=gram)
    O.data.copy_(base)
    ms = nms.MinSRSplit(O, 1e-3, gram=gram)
    ms.gram()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        ms.factor()
    torch.cuda.current_stream().wait_stream(side)
    d2 = ms.solve(eps)
    assert int(ms.info.item()) == 0
    assert torch.equal(d1, d2)


def test_overlapped_step_equals_serial_step():
    """A VMC step with the Gram factorised on a side stream beside the local energies equals the same
    step with every stage on one stream, bit for bit (theta, energies, statistics)."""
    N, L, hz = 5, 7.0, 0.5
    g = bands.rashba_gs(L, N, 1.0, 1.0, hz, n_cut=10, twist=(0.1, 0.17))
    model = ContinuumSpinorPsiformer(n_elec=N, Gs=g["Gs"], W=g["W"], L=L, embed_dim=16, depth=2, n_heads=4,
                                     precision="mixed")
    theta = opf.init_theta(opf.Arch(N, g["Gs"], g["W"], L, embed_dim=16, depth=2, n_heads=4), 5,
                           backflow_scale=0.05, scale_all=0.02)
    R, S = _configs(512, N, L, 11)
    out = []
    for overlap in (True, False):
        vs = nms.VMCStep(model, theta.cuda().clone(), torch.tensor(R).cuda(), torch.tensor(S, dtype=torch.int32).cuda(),
                         m_star=1.0, lambda_r=1.0, h_z=hz, lr=0.05, lam=1e-3, n_sweeps=2, step_size=0.5,
                         overlap=overlap)
        stats = [vs.step(key=k) for k in range(2)]
        torch.cuda.synchronize()
        out.append((vs.theta.clone(), vs.last["e"].clone(), torch.stack(stats)))
    assert vs.failed_steps == 0
    for a, b in zip(out[0], out[1]):
        assert torch.equal(a, b)


def test_bf16_gram_tail_split_matches_unsplit():
    """W = 4096: the 256-tile bf16 Gram has 528 tiles, so on a 256-CU device its last round holds
    16 tiles, which the single-system call splits 16 ways along K (partials summed in a fixed
    order); the batched call of one system keeps whole tiles.  Well-conditioned synthetic O
    (i.i.d. Gaussian, P = 32768 >> 2W), no refinement: the two directions agree to rounding
    (1e-4) and both lie within bf16 accuracy (2e-2) of the float64 Gram's."""
    from nqs.ensemble import minsr_direction_batched
    W, P = 4096, 32768
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    ldo = nms._ldo(P)
    data = torch.zeros(2 * W, ldo, dtype=torch.float32, device="cuda")
    data[:, :P] = torch.randn(2 * W, P, generator=g, device="cuda")
    O32 = nms.OMatrix(data, W, P)
    O64 = nms.OMatrix(data.double(), W, P)
    eps = torch.complex(torch.randn(W, generator=g, dtype=torch.float64, device="cuda"),
                        torch.randn(W, generator=g, dtype=torch.float64, device="cuda"))
    lam = 1e-3
    d_split = nms.minsr_direction(O32, eps, lam, gram="bf16", refine=0)
    (d_whole,), info = minsr_direction_batched([O32], [eps], [lam], "bf16", 0)
    d64 = nms.minsr_direction(O64, eps, lam, gram="fp64", refine=0)
    assert int(info[0]) == 0
    den = float(torch.max(torch.abs(d64)))
    rel_split = float(torch.max(torch.abs(d_split - d_whole))) / den
    rel64 = float(torch.max(torch.abs(d_split - d64))) / den
    print(f"split vs whole {rel_split:.2e}, bf16 vs fp64 {rel64:.2e}", flush=True)
    assert rel_split < 1e-4
    assert rel64 < 2e-2


def test_clip_matches_oracle():
    rng = np.random.default_rng(0)
    for W in (7, 256, 1000):
        e = rng.normal(size=W) + 1j * rng.normal(size=W)
        e[::17] += 50.0
        ec, st = nms.clip_energies(torch.tensor(e), 5.0)
        ref = oms.clip_energies(e, 5.0)
        assert np.array_equal(ec.cpu().numpy(), ref), float(np.max(np.abs(ec.cpu().numpy() - ref)))
        s = st.cpu().numpy()
        assert abs(s[0] - e.real.mean()) < 1e-12 and abs(s[1] - e.real.std() / np.sqrt(W)) < 1e-12


def test_minsr_descends_to_Egs():
    """tests/test_nqs_minsr.py:64-76 on the HIP path: perturbed backflow, 256 walkers, 80 steps."""
    N, L, hz = 5, 7.0, 0.5
    g = bands.rashba_gs(L, N, 1.0, 1.0, hz, n_cut=10, twist=(0.1, 0.17))
    model = ContinuumSpinorPsiformer(n_elec=N, Gs=g["Gs"], W=g["W"], L=L, embed_dim=16, depth=2, n_heads=4)
    params = model.init(0)
    rng = np.random.default_rng(7)
    for leaf in ("kernel", "bias"):
        x = params["params"]["backflow"][leaf]
        params["params"]["backflow"][leaf] = x + 0.1 * torch.tensor(rng.standard_normal(tuple(x.shape))
                                                                    + 1j * rng.standard_normal(tuple(x.shape)))
    params, hist = nms.train_minsr(model, params, L, N, 3, m_star=1.0, lambda_r=1.0, h_z=hz, n_walkers=256,
                                   n_steps=80, lr=0.1, lam=1e-3, n_sweeps=8, step_size=0.5, burn=40, log_every=20)
    E_final = float(np.mean(hist["E"][-5:]))
    assert hist["E"][0] > g["E_gs"] + 1e-2
    assert g["E_gs"] - 2e-2 < E_final < g["E_gs"] + 1e-2
