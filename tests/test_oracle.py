"""CPU pins of the oracle (no GPU): closed forms, identities and invariants from the reference's
own tests, restated against oracle/ (tests/test_nqs_local_energy.py, test_nqs_psiformer.py,
test_nqs_spiral.py, test_nqs_ewald.py, test_nqs_cell.py, test_nqs_minsr.py,
test_nqs_rashba_reference.py)."""
import math

import numpy as np
import pytest
import torch

from oracle import bands, ewald as oew, local_energy as ole, minsr as oms, psiformer as opf, sampler as osa
from oracle import observables as oob


def _rashba(N=5, L=7.0, lr=1.0, hz=0.5):
    return bands.rashba_gs(L, N, 1.0, lr, hz, n_cut=10, twist=(0.1, 0.17))


def test_bands_closed_form_and_brute_force():
    g = _rashba()
    assert abs(g["E_gs"] - g["E_closed"]) < 1e-12
    # brute force: lowest N eigenvalues of the block-diagonal plane-wave Hamiltonian
    L, N = 7.0, 5
    n = np.arange(-10, 11)
    ev = []
    for a in n:
        for b in n:
            kx, ky = (2 * np.pi * a + 0.1) / L, (2 * np.pi * b + 0.17) / L
            ev += list(np.linalg.eigvalsh(bands.bloch_h(kx, ky, 1.0, 1.0, 0.5)))
    assert abs(sum(sorted(ev)[:N]) - g["E_gs"]) < 1e-12
    # Rashba ring minimum E = -lambda^2 m*/2 at |k| = lambda m* (h_z = 0)
    lo, _ = bands.band_energies(1.0, 0.0, 1.0, 1.0, 0.0)
    assert abs(lo + 0.5) < 1e-15


def _slater_logpsi(g):
    return lambda R, S: opf.slater_logpsi(g["Gs"], g["W"], R, S)


@pytest.mark.parametrize("N,L,lr,hz", [(5, 7.0, 1.0, 0.5), (7, 8.0, 1.0, 0.5)])
def test_oracle_zero_variance(N, L, lr, hz):
    g = bands.rashba_gs(L, N, 1.0, lr, hz, n_cut=10, twist=(0.1, 0.17))
    E = ole.make_local_energy(_slater_logpsi(g), 1.0, lr, hz)
    rng = np.random.default_rng(0)
    vals = np.array([E(torch.tensor(rng.uniform(0, L, (N, 2))), torch.tensor(rng.integers(0, 2, N))) for _ in range(6)])
    assert np.std(vals.real) < 1e-7 and abs(vals.real.mean() - g["E_gs"]) < 1e-7 and np.max(np.abs(vals.imag)) < 1e-7


def test_oracle_zero_variance_fails_for_a_planted_sign_error():
    """A flipped SOC sign (sigma_y -> -sigma_y) must break the certificate: the pin is sensitive."""
    g = _rashba()
    E = ole.make_local_energy(_slater_logpsi(g), 1.0, 1.0, 0.5)
    saved = ole._SY
    try:
        ole._SY = ((0.0, 1.0j), (-1.0j, 0.0))
        rng = np.random.default_rng(0)
        vals = np.array([E(torch.tensor(rng.uniform(0, 7, (5, 2))), torch.tensor(rng.integers(0, 2, 5)))
                         for _ in range(6)])
    finally:
        ole._SY = saved
    assert np.std(vals.real) > 1e-3


def test_oracle_polarized_limit():
    N, L, hz = 5, 9.0, 4.0
    g = bands.rashba_gs(L, N, 1.0, 0.0, hz, n_cut=10, twist=(0.1, 0.17))
    E = ole.make_local_energy(_slater_logpsi(g), 1.0, 0.0, hz)
    R = torch.tensor(np.random.default_rng(1).uniform(0, L, (N, 2)))
    e = E(R, torch.zeros(N, dtype=torch.long))
    ekin = float(np.sum(g["Gs"] ** 2) / 2.0)
    assert abs(e.real - (ekin - 0.5 * hz * N)) < 1e-7 and abs(e.real - g["E_gs"]) < 1e-7


def test_oracle_spiral_zero_variance():
    res = bands.spiral_gs(7, 8.0, (1, 0), 0.3, lambda_r=1.0, h_z=0.0, n_cut=10, twist=(0.1, 0.17))
    assert res["boundary_weight"] < 1e-13 and res["gap"] > 1e-6
    lp = lambda R, S: opf.slater_logpsi(res["Gs"], res["W"], R, S)
    E = ole.make_local_energy(lp, 1.0, 1.0, 0.0, ext_field=bands.spiral_field(0.3, res["Q"]))
    rng = np.random.default_rng(0)
    vals = np.array([E(torch.tensor(rng.uniform(0, 8, (7, 2))), torch.tensor(rng.integers(0, 2, 7))) for _ in range(4)])
    assert np.std(vals.real) < 1e-7 and abs(vals.real.mean() - res["E_gs"]) < 1e-7


def _arch(**kw):
    g = _rashba()
    base = dict(embed_dim=16, depth=2, n_heads=4, n_freq=4)
    base.update(kw)
    return g, opf.Arch(5, g["Gs"], g["W"], 7.0, **base)


def test_oracle_psiformer_zero_backflow_is_slater_and_antisymmetric():
    g, a = _arch(n_dets=1)
    th = opf.init_theta(a, 0)
    rng = np.random.default_rng(3)
    R, S = torch.tensor(rng.uniform(0, 7, (5, 2))), torch.tensor(rng.integers(0, 2, 5))
    d = complex(opf.logpsi(a, th, R, S)) - complex(opf.slater_logpsi(g["Gs"], g["W"], R, S))
    assert abs(d.real) < 1e-12 and abs((d.imag + np.pi) % (2 * np.pi) - np.pi) < 1e-12
    g, a = _arch(n_dets=3, use_jastrow=True, learnable_env=True)
    th = opf.init_theta(a, 1, backflow_scale=0.1, scale_all=0.05)
    perm = torch.tensor([1, 0, 2, 3, 4])
    r = np.exp(complex(opf.logpsi(a, th, R[perm], S[perm])) - complex(opf.logpsi(a, th, R, S)))
    assert abs(r + 1.0) < 1e-10


def test_oracle_library_ops_match_torch_definitions():
    """LayerNorm / tanh-GELU restatements against torch.nn.functional's implementations of the same
    definitions (LayerNorm: biased variance; eps 1e-6)."""
    x = torch.randn(7, 16, dtype=torch.float64)
    p = {"n/scale": torch.randn(16, dtype=torch.float64), "n/bias": torch.randn(16, dtype=torch.float64)}
    ref = torch.nn.functional.layer_norm(x, (16,), p["n/scale"], p["n/bias"], eps=1e-6)
    assert torch.max(torch.abs(opf.layer_norm(x, p, "n") - ref)) < 1e-12
    ref = torch.nn.functional.gelu(x, approximate="tanh")
    assert torch.max(torch.abs(opf.gelu_tanh(x) - ref)) < 1e-14


def _torch_block(p, b, d, h):
    """torch.nn.TransformerEncoderLayer with the reference block's weights copied in: pre-norm
    (norm_first), eps 1e-6, feed-forward d -> 2d -> d with tanh-GELU, no dropout.  torch's
    layer splits heads as contiguous slices of width d/h and scales QK^T by (d/h)^-1/2."""
    gelu = lambda t: torch.nn.functional.gelu(t, approximate="tanh")
    layer = torch.nn.TransformerEncoderLayer(d, h, dim_feedforward=2 * d, dropout=0.0, activation=gelu,
                                             layer_norm_eps=1e-6, batch_first=True, norm_first=True,
                                             dtype=torch.float64)
    a = b + "/attn"
    with torch.no_grad():
        layer.self_attn.in_proj_weight.copy_(torch.cat([p[a + "/q/kernel"].T, p[a + "/k/kernel"].T,
                                                        p[a + "/v/kernel"].T]))
        layer.self_attn.in_proj_bias.copy_(torch.cat([p[a + "/q/bias"], p[a + "/k/bias"], p[a + "/v/bias"]]))
        layer.self_attn.out_proj.weight.copy_(p[a + "/o/kernel"].T)
        layer.self_attn.out_proj.bias.copy_(p[a + "/o/bias"])
        layer.linear1.weight.copy_(p[b + "/Dense_0/kernel"].T)
        layer.linear1.bias.copy_(p[b + "/Dense_0/bias"])
        layer.linear2.weight.copy_(p[b + "/Dense_1/kernel"].T)
        layer.linear2.bias.copy_(p[b + "/Dense_1/bias"])
        layer.norm1.weight.copy_(p[b + "/LayerNorm_0/scale"])
        layer.norm1.bias.copy_(p[b + "/LayerNorm_0/bias"])
        layer.norm2.weight.copy_(p[b + "/LayerNorm_1/scale"])
        layer.norm2.bias.copy_(p[b + "/LayerNorm_1/bias"])
    return layer.eval()


@pytest.mark.parametrize("d,h,depth", [(128, 8, 4), (32, 4, 2), (16, 4, 0)])
def test_oracle_backbone_matches_torch_transformer_layers(d, h, depth):
    """The backbone (src/nqs/psiformer.py:67-81,118-136) against torch's own library layers with the
    same weights: embed Dense, then per block torch.nn.TransformerEncoderLayer(norm_first=True), then
    torch.nn.LayerNorm.  Post-norm placement, a missing final LayerNorm, interleaved heads, erf-GELU
    or a different attention scale each move the result by far more than the tolerance.  Every leaf
    is perturbed so that LayerNorm scales and biases are not the identity."""
    g = _rashba(N=16, L=21.2694, lr=0.8, hz=0.05)
    a = opf.Arch(16, g["Gs"], g["W"], 21.2694, embed_dim=d, depth=depth, n_heads=h, n_freq=8, n_dets=8)
    th = opf.init_theta(a, 11, backflow_scale=0.05, scale_all=0.1)
    p = opf.unravel(a, th)
    rng = np.random.default_rng(12)
    R, S = torch.tensor(rng.uniform(0, 21.2694, (16, 2))), torch.tensor(rng.integers(0, 2, 16))
    x = opf.backbone(a, p, R, S)
    # library path: features by an explicit loop over (axis, sin/cos, harmonic)
    s = R / 21.2694
    cols = [f(2 * math.pi * m * s[:, ax]) for ax in range(2) for f in (torch.sin, torch.cos) for m in range(1, 9)]
    tok = torch.cat([torch.stack(cols, -1), torch.nn.functional.one_hot(S, 2).double()], -1)
    emb = torch.nn.Linear(34, d, dtype=torch.float64)
    with torch.no_grad():
        emb.weight.copy_(p["params/embed/kernel"].T)
        emb.bias.copy_(p["params/embed/bias"])
        y = emb(tok)[None]
        for l in range(depth):
            y = _torch_block(p, f"params/TransformerBlock_{l}", d, h)(y)
        y = torch.nn.functional.layer_norm(y[0], (d,), p["params/LayerNorm_0/scale"], p["params/LayerNorm_0/bias"],
                                           eps=1e-6)
    assert torch.max(torch.abs(x - y)) < 1e-12 * max(1.0, float(torch.max(torch.abs(y))))
    # the planted mistakes above are visible: post-norm at the same weights differs at O(1)
    if depth:
        with torch.no_grad():
            yb = emb(tok)[None]
            blk = _torch_block(p, "params/TransformerBlock_0", d, h)
            blk.norm_first = False
            yb = torch.nn.functional.layer_norm(blk(yb)[0], (d,), eps=1e-6)
            x1 = opf.layer_norm(x, {"n/scale": torch.ones(d, dtype=torch.float64),
                                    "n/bias": torch.zeros(d, dtype=torch.float64)}, "n")
        assert torch.max(torch.abs(yb - x1)) > 1e-2


def test_oracle_mha_and_softmax_match_torch_functional():
    """mha (psiformer.py:50-64) against torch.nn.functional.multi_head_attention_forward with the
    packed q/k/v projection, and the softmax step against torch.softmax."""
    d, h, N = 128, 8, 16
    rng = np.random.default_rng(5)
    t = lambda *s: torch.tensor(rng.standard_normal(s) / math.sqrt(s[0]), dtype=torch.float64)
    p = {f"m/{n}/kernel": t(d, d) for n in "qkvo"}
    p.update({f"m/{n}/bias": t(d) for n in "qkvo"})
    x = torch.tensor(rng.standard_normal((N, d)), dtype=torch.float64)
    ours = opf.mha(x, p, "m", h)
    ref, _ = torch.nn.functional.multi_head_attention_forward(
        x[:, None], x[:, None], x[:, None], d, h,
        torch.cat([p["m/q/kernel"].T, p["m/k/kernel"].T, p["m/v/kernel"].T]),
        torch.cat([p["m/q/bias"], p["m/k/bias"], p["m/v/bias"]]), None, None, False, 0.0,
        p["m/o/kernel"].T, p["m/o/bias"], training=False, need_weights=False)
    assert torch.max(torch.abs(ours - ref[:, 0])) < 1e-12
    # interleaved heads (head j = columns j, j+h, ...) would be a different function
    perm = torch.arange(d).reshape(d // h, h).T.reshape(-1)
    p2 = dict(p)
    for n in "qkv":
        p2[f"m/{n}/kernel"], p2[f"m/{n}/bias"] = p[f"m/{n}/kernel"][:, perm], p[f"m/{n}/bias"][perm]
    p2["m/o/kernel"] = p["m/o/kernel"][perm]
    assert torch.max(torch.abs(opf.mha(x, p2, "m", h) - ref[:, 0])) > 1e-3
    s = torch.tensor(rng.standard_normal((h, N, N)) * 4, dtype=torch.float64)
    z = s - s.max(-1, keepdim=True).values
    assert torch.max(torch.abs(torch.exp(z) / torch.exp(z).sum(-1, keepdim=True) - torch.softmax(s, -1))) < 1e-15


def test_oracle_backflow_layout():
    """bf[k, a, i] = out[i, k n_orb + a] (psiformer.py:139-143: reshape (N, K, n_orb), transpose
    (1, 2, 0)), checked by perturbing one backflow bias entry at a time with K != N: only the
    elements of row a of determinant k change, M_k[a, i] = env[a, i] (1 + out[i, k n_orb + a]); the
    finite difference of the oracle's log psi is compared with tr(M_k^-1 dM_k) weighted by the
    determinant's share of the sum, built here independently of the oracle's reshape."""
    g = _rashba()
    a = opf.Arch(5, g["Gs"], g["W"], 7.0, embed_dim=16, depth=1, n_heads=4, n_freq=4, n_dets=3)
    th = opf.init_theta(a, 4, backflow_scale=0.05)
    p = opf.unravel(a, th)
    rng = np.random.default_rng(6)
    R, S = torch.tensor(rng.uniform(0, 7, (5, 2))), torch.tensor(rng.integers(0, 2, 5))
    base = opf.logpsi_from_params(a, p, R, S)
    # a bias on output column c = k*N + a multiplies row a of determinant k by (1 + bf + delta):
    # with det_logw = 0, d logpsi / d bias_c = w_k * sum_i (M_k^-1)[i, a] env[a, i]
    x = opf.backbone(a, p, R, S).to(torch.complex128)
    bf = (x @ p["params/backflow/kernel"] + p["params/backflow/bias"])
    Gs, W = torch.tensor(a.Gs), torch.tensor(a.W)
    env = torch.einsum("gi,agi->ai", torch.exp(1j * (Gs @ R.T).to(torch.complex128)), W[:, :, S.long()])
    Ms = [torch.stack([env[aa] * (1 + bf[:, k * 5 + aa]) for aa in range(5)]) for k in range(3)]
    logd = torch.stack([torch.linalg.slogdet(M)[1] + torch.log(torch.linalg.slogdet(M)[0]) for M in Ms])
    z = torch.exp(logd - logd.real.max())
    wk = z / z.sum()
    h = 1e-7
    for c in (0, 4, 6, 14):
        k, aa = divmod(c, 5)
        p2 = dict(p)
        b2 = p["params/backflow/bias"].clone()
        b2[c] += h
        p2["params/backflow/bias"] = b2
        fd = (opf.logpsi_from_params(a, p2, R, S) - base) / h
        Minv = torch.linalg.inv(Ms[k])
        an = wk[k] * (Minv[:, aa] * env[aa]).sum()
        assert abs(complex(fd - an)) < 1e-5 * max(1.0, abs(complex(an))), (c, fd, an)


def test_oracle_backbone_permutation_equivariant():
    g, a = _arch()
    th = opf.init_theta(a, 2, scale_all=0.05)
    p = opf.unravel(a, th)
    rng = np.random.default_rng(4)
    R, S = torch.tensor(rng.uniform(0, 7, (5, 2))), torch.tensor(rng.integers(0, 2, 5))
    perm = torch.tensor([3, 0, 4, 1, 2])
    x = opf.backbone(a, p, R, S)
    xp = opf.backbone(a, p, R[perm], S[perm])
    assert torch.max(torch.abs(x[perm] - xp)) < 1e-12


def test_oracle_ewald_invariants():
    L = 7.0
    vals, zs = [], []
    for c in (0.8, 1.0, 1.25):
        ep, z, _ = oew.make_ewald(L, 1.0, ewald_alpha=c * math.sqrt(math.pi) / L, n_real=8, n_recip=14)
        vals.append(ep([1.3, -2.1]))
        zs.append(z)
    assert np.ptp(vals) < 1e-7 and np.ptp(zs) < 1e-7
    ep, zM, ce = oew.make_ewald(8.0, 1.0, n_real=6, n_recip=12)
    R = np.random.default_rng(0).uniform(0, 8, (6, 2))
    E = ce(R)
    assert abs(E - ce(R + np.array([2.3, -1.1]))) < 1e-8
    assert abs(ep([1.1, 2.7]) - ep([-1.1, -2.7])) < 1e-10
    assert abs(ep([1e-4, 0.0]) * 1e-4 - 1.0) < 2e-2


def test_oracle_minsr_equals_paramsr_and_O_matches_fd():
    g, a = _arch(embed_dim=12, n_heads=3)
    th = opf.init_theta(a, 0, backflow_scale=0.05)
    rng = np.random.default_rng(4)
    R = rng.uniform(0, 7, (6, 5, 2))
    S = rng.integers(0, 2, (6, 5))
    O = oms.per_sample_O(a, th, R, S)
    eps = torch.tensor(rng.normal(size=6) + 1j * rng.normal(size=6))
    for lam in (1e-2, 1e-4):
        d1, d2 = oms.minsr_direction(O, eps, lam), oms.paramsr_direction(O, eps, lam)
        assert float(torch.max(torch.abs(d1 - d2)) / torch.max(torch.abs(d2))) < 1e-6
    # central finite differences on a few parameters
    h = 1e-5
    for k in (3, 500, th.shape[0] - 1):
        e = torch.zeros_like(th)
        e[k] = h
        fp = complex(opf.logpsi(a, th + e, torch.tensor(R[0]), torch.tensor(S[0])))
        fm = complex(opf.logpsi(a, th - e, torch.tensor(R[0]), torch.tensor(S[0])))
        fd = (fp - fm) / (2 * h)
        assert abs(fd - complex(O[0, k])) < 1e-6 * max(1.0, abs(fd))


def test_oracle_accept_rule():
    """sampler.py:41-43: accept iff u < exp(min(0, d)); d = 2 Re L' - 2 Re L; NaN rejects."""
    W, N = 4, 2
    L0 = np.array([0.0, 0.0, 0.0, 0.0])
    Lnew = np.array([0.1, -0.5, np.nan, -np.inf])
    logpsi_b = lambda R, S: Lnew.astype(complex)
    R = np.full((W, N, 2), 0.5)
    S = np.zeros((W, N), int)
    u = np.array([0.99, math.exp(-1.0) - 1e-12, 0.0, 0.0])
    st = osa.make_streams(1, W, N, 0, 0)
    st["xi"][:] = 0.0
    _, _, acc, _, _ = osa.sweep(logpsi_b, R, S, np.eye(2), 0.1, st["xi"][0], u, st["flip_i"][0], st["u_flip"][0],
                                lp_cur=2 * L0)
    assert acc.tolist() == [True, True, False, False]


# ------------------------------------------------------------------ two-electron exact diagonalisation
def _relative_problem(L, n_cut, m_star, alpha):
    """The lambda = h_z = 0 singlet reduced to one particle in the relative coordinate: plane waves
    k = 2 pi n / L (n in the same disc), H[n, n'] = k^2/m* delta + alpha 2 pi/(L^2 |k - k'|)."""
    from oracle import ed
    v = ed.disc_vectors(n_cut)
    k = 2.0 * np.pi * v / L
    H = np.diag((k * k).sum(1) / m_star).astype(float)
    q = np.linalg.norm(k[:, None, :] - k[None, :, :], axis=-1)
    with np.errstate(divide="ignore"):
        H += np.where(q > 0, alpha * 2.0 * np.pi / (L * L * np.where(q > 0, q, 1.0)), 0.0)
    return float(np.linalg.eigvalsh(H)[0])


@pytest.mark.parametrize("lr,hz,twist", [(0.5, 0.0, (0.0, 0.0)), (1.0, 0.5, (0.1, 0.17)), (0.3, 0.2, (0.4, -0.2))])
def test_ed_noninteracting_limit_is_the_band_sum(lr, hz, twist):
    """alpha = 0: the two-electron ground energy over all momentum sectors equals the sum of the two
    lowest closed-form band energies (rashba_reference.py:49-54)."""
    from oracle import ed
    L, cut = 5.0, 4
    g = bands.rashba_gs(L, 2, 1.0, lr, hz, n_cut=cut, twist=twist)
    sectors = [(i, j) for i in range(-2, 3) for j in range(-2, 3)]
    e = ed.ground_energy(L, cut, sectors=sectors, lambda_r=lr, h_z=hz, alpha=0.0, twist=twist)
    assert abs(e - g["E_closed"]) < 1e-10


@pytest.mark.parametrize("alpha", [0.3, 1.0])
def test_ed_spin_free_singlet_equals_relative_coordinate_problem(alpha):
    """lambda = h_z = 0: the antisymmetrised spinful ground state in the K = 0 sector is the singlet,
    whose spatial part depends on r1 - r2 only; its energy must equal the lowest eigenvalue of the
    one-body relative-coordinate problem in the same plane-wave set (independent construction:
    no spins, no antisymmetrisation, reduced mass m*/2)."""
    from oracle import ed
    L = 2.0 * math.sqrt(2.0 * math.pi)
    for cut in (4, 7):
        HA, _ = ed.sector_hamiltonian(L, cut, (0, 0), 1.0, 0.0, 0.0, alpha)
        assert abs(float(np.linalg.eigvalsh(HA)[0]) - _relative_problem(L, cut, 1.0, alpha)) < 1e-11


def test_ed_rashba_sign_convention_invariance_and_variational_convergence():
    """The spectrum cannot depend on the sign convention of the Rashba term (a global spin rotation);
    the ground energy decreases monotonically with the cutoff (variational in the basis)."""
    from oracle import ed
    L = 2.0 * math.sqrt(2.0 * math.pi)
    e_p = ed.ground_energy(L, 5, lambda_r=0.5, alpha=1.0)
    e_m = ed.ground_energy(L, 5, lambda_r=-0.5, alpha=1.0)
    assert abs(e_p - e_m) < 1e-11
    es = [ed.ground_energy(L, c, lambda_r=0.5, alpha=1.0) for c in (3, 5, 7, 9)]
    assert all(b < a for a, b in zip(es, es[1:]))
    assert es[-2] - es[-1] < 1e-4


def test_ed_madelung_matches_the_ewald_self_energy():
    """zeta_M of the ED (converged sums) equals the Ewald restatement's (default sums) to 1e-12 and
    is independent of the splitting parameter."""
    from oracle import ed
    L = 5.0
    A = np.diag([L, L])
    z = ed.madelung(A)
    _, z0, _ = oew.make_ewald_lattice(A, 1.0)
    _, z1, _ = oew.make_ewald_lattice(A, 1.0, ewald_alpha=0.7 * math.sqrt(math.pi) / L, n_real=9, n_recip=12)
    assert abs(z - z0) < 1e-12 and abs(z - z1) < 1e-12


def test_periodic_cusp_jastrow_is_smooth_periodic_and_cusped():
    """The chord-distance cusp Jastrow (jastrow_kind "cusp_periodic"): periodic, |r| at short range,
    and (unlike the minimum-image form) with a gradient that is continuous across the cell edge."""
    L = 6.0
    a = opf.Arch(2, np.zeros((2, 2)), np.zeros((2, 2, 2), complex), L, use_jastrow=True, jastrow_kind="cusp_periodic")
    am = opf.Arch(2, np.zeros((2, 2)), np.zeros((2, 2, 2), complex), L, use_jastrow=True, jastrow_kind="cusp")
    p = {"params/jastrow_loga": torch.tensor(0.3, dtype=torch.float64)}
    S = torch.tensor([0, 1])

    def J(arch, x, y):
        R = torch.tensor([[0.0, 0.0], [x, y]], dtype=torch.float64)
        return float(opf.jastrow(arch, p, R, S))

    # periodicity and short-range agreement with the minimum-image form
    assert abs(J(a, 1.3, 0.4) - J(a, 1.3 + L, 0.4 - L)) < 1e-12
    assert abs(J(a, 0.05, 0.02) - J(am, 0.05, 0.02)) < 1e-5
    # gradient across the edge x = L/2 (pair displacement crosses the Wigner-Seitz boundary)
    def gx(arch, x):
        h = 1e-6
        return (J(arch, x + h, 0.7) - J(arch, x - h, 0.7)) / (2 * h)
    jump_p = abs(gx(a, L / 2 + 1e-4) - gx(a, L / 2 - 1e-4))
    jump_m = abs(gx(am, L / 2 + 1e-4) - gx(am, L / 2 - 1e-4))
    assert jump_p < 1e-3 and jump_m > 1e-2


def test_pinning_potential_closed_forms():
    """Gaussian pinning centres (no reference counterpart): periodicity, continuity across the
    cell edge, and the image sum against its Poisson-summed Fourier form
    sum_n exp(-|u - A n|^2 / 2w^2) = (2 pi w^2 / area) sum_G exp(-w^2 |G|^2 / 2) cos(G.u)."""
    L, w = 10.0, 1.5
    A = np.array([[L, 2.0], [0.0, 0.9 * L]])
    area = abs(np.linalg.det(A))
    B = 2 * np.pi * np.linalg.inv(A).T
    pin = [(2.0, 3.0, w, 0.05, 0.0, 0.0, 0.0), (6.0, 1.0, w, 0.0, 0.01, -0.02, 0.03)]
    rng = np.random.default_rng(0)
    r = rng.uniform(0, 1, (7, 2)) @ A.T
    g = ole.pin_profiles(r, pin, A)
    m = np.array([(i, j) for i in range(-9, 10) for j in range(-9, 10)], float)
    G = m @ B.T
    for p, (cx, cy, *_) in enumerate(pin):
        u = r - np.array([cx, cy])
        four = (2 * np.pi * w * w / area) * (np.exp(-0.5 * w * w * (G * G).sum(1))[None] * np.cos(u @ G.T)).sum(1)
        assert np.max(np.abs(g[p] - four)) < 1e-13
    # potential and field are the profiles weighted by v0 and h
    assert abs(ole.pinning_potential(r, pin, A) - 0.05 * g[0].sum()) < 1e-15
    assert np.max(np.abs(ole.pinning_field(r, pin, A) - g[1][:, None] * np.array([0.01, -0.02, 0.03]))) < 1e-17
    # periodic under any lattice shift, continuous across the cell edge
    shift = A @ np.array([1.0, -2.0])
    assert np.max(np.abs(ole.pin_profiles(r + shift, pin, A) - g)) < 1e-12
    for y in (0.5, 3.0, 8.7):
        a_ = ole.pin_profiles(np.array([[1e-9 + 2.0 * y / (0.9 * L), y]]), pin, A)
        b_ = ole.pin_profiles(np.array([[1e-9 + 2.0 * y / (0.9 * L) + L, y]]), pin, A)
        assert np.max(np.abs(a_ - b_)) < 1e-12


def _hex_cell(a=7.0):
    A = np.array([[a, a / 2], [0.0, a * np.sqrt(3) / 2]])
    B = 2 * np.pi * np.linalg.inv(A).T
    return A, B[:, 0], B[:, 1]


def test_oracle_skyrmion_textures():
    """pinning.py:20-37 restated: the triple-Q texture on its own (hexagonal) magnetic cell has
    skyrmion number 1 for |m0| small (Neel and Bloch, either chirality) and 0 when m0 > 3 keeps
    m_z > 0; near r = 0 the in-plane part is radial (Neel) or azimuthal (Bloch), which needs the
    three arms at 120 degrees; |f| = h_pin."""
    from oracle import pinning as opn
    A, q1, q2 = _hex_cell()
    for neel in (True, False):
        for ch in (1.0, -1.0):
            f = opn.make_skx_pin(q1, q2, 0.1, m0=0.3, chirality=ch, neel=neel)
            assert abs(opn.skyrmion_charge(f, A) - 1.0) < 1e-9
            assert abs(opn.skyrmion_charge(opn.make_skx_pin(q1, q2, 0.1, m0=3.5, chirality=ch, neel=neel), A)) < 1e-9
            r = torch.tensor([[1e-4, 3e-5], [-2e-5, 7e-5]])
            v = f(r)
            cross = r[:, 0] * v[:, 1] - r[:, 1] * v[:, 0]          # (r x f)_z
            dot = v[:, 0] * r[:, 0] + v[:, 1] * r[:, 1]
            inplane = torch.linalg.norm(v[:, :2], dim=1) * torch.linalg.norm(r, dim=1)
            if neel:
                assert float((cross.abs() / inplane).max()) < 1e-6 and float((ch * dot).min()) > 0
            else:
                assert float((dot.abs() / inplane).max()) < 1e-6 and float((ch * cross).min()) > 0
            pts = torch.tensor(np.random.default_rng(1).uniform(0, 7, (50, 2)))
            assert float((torch.linalg.norm(f(pts), dim=1) - 0.1).abs().max()) < 1e-10
    # a planted sin <-> cos swap flips the skyrmion number
    qs = torch.tensor(np.stack([q1, q2, -(q1 + q2)]))
    qh = qs / torch.linalg.norm(qs, dim=1, keepdim=True)

    def planted(r):
        ph = r @ qs.T
        m = torch.stack([(torch.cos(ph) * qh[:, 0]).sum(-1), (torch.cos(ph) * qh[:, 1]).sum(-1),
                         0.3 + torch.sin(ph).sum(-1)], -1)
        return m / torch.linalg.norm(m, dim=-1, keepdim=True)
    assert abs(opn.skyrmion_charge(planted, A) + 1.0) < 1e-9


def test_oracle_simple_pins_and_hf_texture():
    """pinning.py:40-119 restated: closed forms of the ferromagnet, spiral, flat helix and
    Overhauser targets; a converged-HF texture whose harmonics are the skyrmion's own
    (S_0 = (0, 0, m0), S_qj = (-i chi e_j, 1)) reproduces make_skx_pin; make_pin dispatch."""
    from oracle import pinning as opn
    A, q1, q2 = _hex_cell()
    Q = np.array([0.4, -0.3])
    qh = Q / np.linalg.norm(Q)
    r = torch.tensor(np.random.default_rng(2).uniform(0, 7, (20, 2)))
    ph = (r @ torch.tensor(Q)).numpy()
    fm = opn.make_fm_pin(0.2, axis=(1.0, 2.0, 2.0))(r).numpy()
    assert np.max(np.abs(fm - 0.2 * np.array([1, 2, 2]) / 3.0)) < 1e-15
    sp = opn.make_spiral_pin(Q, 0.2)(r).numpy()
    # (q_hat carries the reference's +1e-12 in its norm)
    assert np.max(np.abs(sp - 0.2 * np.stack([np.sin(ph) * qh[0], np.sin(ph) * qh[1], np.cos(ph)], 1))) < 1e-12
    xy = opn.make_pin("spiral_xy", Q, None, 0.2)(r).numpy()
    assert np.max(np.abs(xy - 0.2 * np.stack([np.cos(ph), np.sin(ph), 0 * ph], 1))) < 1e-15
    oh = opn.make_pin("overhauser", Q, None, 0.2)(r).numpy()
    assert np.max(np.abs(oh - 0.2 * np.stack([0 * ph, 0 * ph, np.cos(ph)], 1))) < 1e-15
    assert opn.make_pin("fluid", Q, None, 0.2) is None
    for neel, ch in ((True, 1.0), (False, -1.0)):
        qs = np.stack([q1, q2, -(q1 + q2)])
        qhat = qs / np.linalg.norm(qs, axis=1, keepdims=True)
        e = qhat if neel else np.stack([-qhat[:, 1], qhat[:, 0]], 1)
        keys = [(0, 0), (1, 0), (0, 1), (-1, -1)]
        vals = [(0.0, 0.0, 0.3)] + [(-1j * ch * e[j, 0], -1j * ch * e[j, 1], 1.0) for j in range(3)]
        hf = opn.make_hf_skx_pin(keys, vals, q1, q2, 0.1)(r).numpy()
        sk = opn.make_pin("skx" if neel else "skx_bloch", q1, q2, 0.1, chirality=ch)(r).numpy()
        assert np.max(np.abs(hf - sk)) < 1e-14


def test_oracle_spin_estimators_spin_independent_trial():
    """observables.py:31-47 restated; test_nqs_observables.py:17-39's analytic case: a
    spin-independent trial is the +x product state, so sigma_x = 1 and Re sigma_y = 0 exactly for
    every electron, and the texture sums give m_x = 1 per electron."""
    from oracle import observables as oob

    def logpsi(R, S):
        return (-0.1 * (R ** 2).sum()).to(torch.complex128)
    rng = np.random.default_rng(0)
    N, L = 6, 5.0
    R = rng.uniform(0, L, (20, N, 2))
    S = rng.integers(0, 2, (20, N))
    sp = []
    for w in range(20):
        sx, sy, sz = oob.spin_estimators(logpsi, R[w], S[w])
        assert np.max(np.abs(sx - 1.0)) == 0.0 and np.max(np.abs(sy.real)) == 0.0
        assert np.array_equal(sz.real, 1.0 - 2.0 * S[w])
        sp.append(np.stack([sx.real, sy.real, sz.real], -1))
    sp = np.array(sp)
    m, rho = oob.texture_sums(R, sp, np.diag([L, L]), 7)
    assert rho.sum() == 20 * N and abs(m[..., 0].sum() / (20 * N) - 1.0) < 1e-15
    assert np.array_equal(m[..., 0], rho)


def test_oracle_structure_factor_and_order_parameter():
    """observables.py:94-121 restated: S_zz(0) = <(sum_i sz_i)^2>/N exactly; uncorrelated spins give
    a flat S_zz (test_nqs_observables.py:88-101); +x spins give m_x(0) = 1 and S_perp(0) = N."""
    from oracle import observables as oob
    rng = np.random.default_rng(1)
    W, N, L = 1200, 8, 5.0
    R = rng.uniform(0, L, (W, N, 2))
    sz = 1.0 - 2.0 * rng.integers(0, 2, (W, N))
    sp = np.stack([np.ones((W, N)), np.zeros((W, N)), sz], -1)
    B = 2 * np.pi / L
    qs = np.array([(i * B, j * B) for i in range(-3, 4) for j in range(-3, 4)])
    m, szz, sperp = oob.spin_fourier(R, sp, qs)
    k0 = np.where(np.linalg.norm(qs, axis=1) < 1e-12)[0][0]
    assert abs(szz[k0] - np.mean(sz.sum(1) ** 2) / N) < 1e-12
    assert abs(m[0, k0] - 1.0) < 1e-15 and abs(sperp[k0] - N) < 1e-12
    fin = np.linalg.norm(qs, axis=1) > 1e-12
    assert szz[fin].max() < 2.5 and 0.5 < szz[fin].mean() < 1.6


def test_oracle_berg_luscher_and_ratio_stats():
    """observables.py:79-91 restated: charge 1 on the triple-Q skyrmion texture (equal to the
    oracle pinning module's solid-angle sum) and 0 on a uniform one; ratio statistics of a constant
    ratio are exact (ess 1, std 0) and the shift cancels in mean r / sqrt(mean |r|^2)."""
    from oracle import observables as oob, pinning as opn
    A, q1, q2 = _hex_cell()
    f = opn.make_skx_pin(q1, q2, 1.0, m0=0.5)
    ng = 24
    fr = np.arange(ng) / ng
    u, v = np.meshgrid(fr, fr, indexing="ij")
    m = f(torch.tensor(np.stack([u, v], -1) @ A.T)).numpy()
    assert abs(oob.berg_luscher(m) - 1.0) < 1e-9
    assert abs(oob.berg_luscher(m) - opn.skyrmion_charge(f, A, ng)) < 1e-9
    assert abs(oob.berg_luscher(np.tile([0.0, 0.0, 1.0], (ng, ng, 1)))) < 1e-12
    la = np.random.default_rng(2).normal(size=50) + 1j * np.random.default_rng(3).normal(size=50)
    st = oob.ratio_stats(la, la + (0.3 + 0.4j))
    assert abs(st[3] - 1.0) < 1e-12 and st[5] < 1e-12
    assert abs(complex(st[0], st[1]) / np.sqrt(st[2]) - np.exp(0.4j)) < 1e-12


def _small_arch():
    g = bands.rashba_gs(7.0, 4, 1.0, 1.0, 0.5, n_cut=8, twist=(0.1, 0.17))
    arch = opf.Arch(4, g["Gs"], g["W"], 7.0, embed_dim=8, depth=1, n_heads=2, n_freq=2)
    return arch, opf.init_theta(arch, 1, backflow_scale=0.05, scale_all=0.02)


def test_oracle_surrogate_gradient_pins():
    """train.py:48-58 restated: reverse mode through the surrogate equals the per-sample form
    (2/W) sum_w Re[conj(e_w - <e>) O_w] with the oracle's O, and central differences."""
    from oracle import train as otr
    arch, theta = _small_arch()
    rng = np.random.default_rng(3)
    R, S = rng.uniform(0, 7.0, (5, 4, 2)), rng.integers(0, 2, (5, 4))
    e = rng.normal(size=5) + 1j * rng.normal(size=5)
    g = otr.surrogate_gradient(arch, theta, R, S, e).numpy()
    O = oms.per_sample_O(arch, theta, R, S).numpy()
    ec = e - e.mean()
    g2 = (2.0 / 5) * (ec.real[:, None] * O.real + ec.imag[:, None] * O.imag).sum(0)
    assert np.max(np.abs(g - g2)) < 1e-12 * max(1.0, np.max(np.abs(g)))
    for k in range(2):
        v = torch.tensor(rng.normal(size=theta.shape[0]))
        h = 1e-6
        fd = (otr.surrogate(arch, theta + h * v, R, S, e) - otr.surrogate(arch, theta - h * v, R, S, e)) / (2 * h)
        assert abs(float(fd) - float(g @ v.numpy())) < 1e-6 * max(1.0, abs(float(fd)))


def test_oracle_adam_pins():
    """optax.adam restated: real leaves follow torch.optim.Adam; a constant gradient gives the
    step -lr g/(|g| + eps) at every count (complex elements with g_re -/+ i g_im); a complex
    element with g_im = 0 moves like a real one."""
    from oracle import train as otr
    arch, theta = _small_arch()
    layout, P = opf.ravel_layout(arch)
    real_idx = np.concatenate([np.arange(off, off + (int(np.prod(s)) if len(s) else 1))
                               for _, s, c, off in layout if not c])
    rng = np.random.default_rng(4)
    lr = 0.01
    th_o, st = theta.numpy(), None
    th_t = theta.clone().requires_grad_(True)
    opt = torch.optim.Adam([th_t], lr=lr, betas=(0.9, 0.999), eps=1e-8)
    for it in range(5):
        g = rng.normal(size=P)
        th_o, st = otr.adam_step(arch, th_o, g, st, lr=lr)
        th_t.grad = torch.tensor(g)
        opt.step()
        ref = th_t.detach().numpy()
        assert np.max(np.abs(th_o[real_idx] - ref[real_idx])) < 1e-14
    for mode, sgn in (("jax", -1.0), ("descent", 1.0)):
        g = rng.normal(size=P)
        th, st = theta.numpy(), None
        for it in range(4):
            prev = th.copy()
            th, st = otr.adam_step(arch, th, g, st, lr=lr, complex_grads=mode)
            for _, s, c, off in layout:
                n = int(np.prod(s)) if len(s) else 1
                if c:
                    gc = g[off:off + n] + sgn * 1j * g[off + n:off + 2 * n]
                    dz = (th[off:off + n] - prev[off:off + n]) + 1j * (th[off + n:off + 2 * n] - prev[off + n:off + 2 * n])
                    assert np.max(np.abs(dz + lr * gc / (np.abs(gc) + 1e-8))) < 1e-14
                else:
                    d = th[off:off + n] - prev[off:off + n]
                    gg = g[off:off + n]
                    assert np.max(np.abs(d + lr * gg / (np.abs(gg) + 1e-8))) < 1e-14
    g = rng.normal(size=P)
    for _, s, c, off in layout:
        if c:
            n = int(np.prod(s)) if len(s) else 1
            g[off + n:off + 2 * n] = 0.0
    th, _ = otr.adam_step(arch, theta.numpy(), g, None, lr=lr)
    for _, s, c, off in layout:
        if c:
            n = int(np.prod(s)) if len(s) else 1
            gg = g[off:off + n]
            assert np.max(np.abs(th[off:off + n] - (theta.numpy()[off:off + n] - lr * gg / (np.abs(gg) + 1e-8)))) < 1e-15
            assert np.array_equal(th[off + n:off + 2 * n], theta.numpy()[off + n:off + 2 * n])


def test_oracle_rotation_overlap_is_the_zero_variance_ratio():
    """For a C4-closed single-plane-wave state whose spinors follow the rotation (an eigenstate of
    R_4 and of R_2), psi(R_n^-1 X) U(S) / psi(X), evaluated with the oracle's Slater log psi at
    random configurations, equals rotation_overlap_planewaves at every configuration; a state that
    is not an eigenstate gives a ratio that varies."""
    L = 7.0
    g = 2 * np.pi / L
    Gs = g * np.array([[1, 0], [0, 1], [-1, 0], [0, -1]], float)
    U4 = np.diag([np.exp(-1j * np.pi / 4), np.exp(1j * np.pi / 4)])
    c0 = np.array([1.0, 0.4 + 0.2j]) / np.linalg.norm([1.0, 0.4 + 0.2j])
    rng = np.random.default_rng(9)
    for sym in (True, False):
        if sym:
            Chi = np.stack([np.linalg.matrix_power(U4, k) @ c0 for k in range(4)])
        else:
            Chi = rng.normal(size=(4, 2)) + 1j * rng.normal(size=(4, 2))
            Chi /= np.linalg.norm(Chi, axis=1, keepdims=True)
        W = np.zeros((4, 4, 2), complex)
        W[np.arange(4), np.arange(4)] = Chi
        for n in (4, 2):
            t = 2 * np.pi / n
            rinv = np.array([[np.cos(t), np.sin(t)], [-np.sin(t), np.cos(t)]])      # turn by -t
            vals = []
            for _ in range(5):
                R, S = rng.uniform(0, L, (4, 2)), rng.integers(0, 2, 4)
                Rr = np.mod(R @ rinv.T, L)
                lr = complex(opf.slater_logpsi(Gs, W, torch.tensor(Rr), torch.tensor(S)))
                l0 = complex(opf.slater_logpsi(Gs, W, torch.tensor(R), torch.tensor(S)))
                vals.append(np.exp(lr - l0 - 1j * (np.pi / n) * (4 - 2 * S.sum())))
            exact = oob.rotation_overlap_planewaves(Gs, Chi, n)
            if sym:
                assert abs(abs(exact) - 1) < 1e-12
                assert np.max(np.abs(np.array(vals) - exact)) < 1e-10, (n, vals, exact)
            else:
                assert np.ptp(np.abs(vals)) > 1e-3
