"""GPU parity of the Ewald kernel and the Metropolis kernel against the oracle.

Ewald: per-walker energy vs the oracle's direct pair sum (ewald.py restated), 1e-12 relative;
the reference's invariants (tests/test_nqs_ewald.py, tests/test_nqs_cell.py) through the kernel.
Metropolis: given identical supplied streams, every accept/reject decision and the final
positions and spins equal the oracle's bit for bit (float64 path).
"""
import numpy as np
import pytest
import torch

from oracle import bands, ewald as oew, sampler as osa, psiformer as opf
from nqs.ewald import make_ewald, make_ewald_lattice
from nqs.cell import square_cell, rect_cell, hex_cell, cell_area
from nqs.psiformer import ContinuumSpinorPsiformer, make_logpsi
from nqs.sampler import Chain, run_chain
from nqs.slater_reference import make_logpsi_general

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("A,n_real,n_recip,N", [(square_cell(7.0), 4, 8, 6), (square_cell(21.27), 1, 6, 16),
                                                 (hex_cell(8.0), 3, 6, 5), (rect_cell(7.0, 12.1), 2, 5, 9)])
def test_ewald_matches_oracle(A, n_real, n_recip, N):
    _, zk, ce = make_ewald_lattice(A, alpha=1.0, n_real=n_real, n_recip=n_recip)
    _, zo, ceo = oew.make_ewald_lattice(A, alpha=1.0, n_real=n_real, n_recip=n_recip)
    assert abs(zk - zo) < 1e-12 * max(1.0, abs(zo))
    rng = np.random.default_rng(1)
    R = rng.uniform(0, 1, (16, N, 2)) @ np.asarray(A).T
    e = ce(R).cpu().numpy()
    ref = np.array([ceo(R[w]) for w in range(len(R))])
    assert np.max(np.abs(e - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_ewald_splitting_independence():
    L = 7.0
    r = np.array([1.3, -2.1])
    vals, zs = [], []
    for c in (0.8, 1.0, 1.25):
        ep, zM, _ = make_ewald(L, alpha=1.0, ewald_alpha=c * np.sqrt(np.pi) / L, n_real=8, n_recip=14)
        vals.append(float(ep(r)))
        zs.append(zM)
    assert np.ptp(vals) < 1e-7 and np.ptp(zs) < 1e-7


def test_ewald_periodicity_evenness_shortrange_alpha():
    L = 7.0
    ep, _, _ = make_ewald(L, alpha=1.0, n_real=8, n_recip=12)
    r = np.array([1.1, 2.7])
    v = float(ep(r))
    assert abs(v - float(ep(r + np.array([L, 0.0])))) < 1e-8
    assert abs(v - float(ep(r + np.array([0.0, L])))) < 1e-8
    assert abs(v - float(ep(-r))) < 1e-10
    for alpha in (1.0, 0.6):
        ep2, _, _ = make_ewald(L, alpha=alpha, n_real=8, n_recip=12)
        for d in (1e-3, 1e-4):
            assert abs(float(ep2(np.array([d, 0.0]))) * d - alpha) < 1e-2 * alpha + 5e-3
    ep1, z1, _ = make_ewald(L, alpha=1.0, n_real=6, n_recip=12)
    ep3, z3, _ = make_ewald(L, alpha=2.5, n_real=6, n_recip=12)
    r = np.array([1.3, -0.7])
    assert abs(2.5 * float(ep1(r)) - float(ep3(r))) < 1e-9 and abs(2.5 * z1 - z3) < 1e-9


def test_ewald_pair_sum_madelung_translation():
    L = 8.0
    ep, zM, ce = make_ewald(L, alpha=1.0, n_real=6, n_recip=12)
    R = np.random.default_rng(0).uniform(0, L, (6, 2))
    E = float(ce(R))
    pairs = sum(float(ep(R[i] - R[j])) for i in range(6) for j in range(i + 1, 6))
    assert abs(E - (pairs + 6 * zM / 2.0)) < 1e-9
    assert abs(E - float(ce(R + np.array([2.3, -1.1])))) < 1e-8


@pytest.mark.parametrize("A", [rect_cell(7.0, 7.0 * np.sqrt(3)), hex_cell(8.0)])
def test_ewald_splitting_independence_general_cell(A):
    R = np.random.default_rng(0).uniform(0, 1, (6, 2)) @ np.asarray(A).T
    Es = []
    for c in (0.7, 1.0, 1.4):
        eta = c * np.sqrt(np.pi) / np.sqrt(cell_area(A))
        _, _, ce = make_ewald_lattice(A, alpha=1.0, ewald_alpha=eta, n_real=9, n_recip=16)
        Es.append(float(ce(R)))
    assert np.ptp(Es) < 1e-6


def _mcmc_case(name):
    N, L = 5, 7.0
    g = bands.rashba_gs(L, N, 1.0, 1.0, 0.5, n_cut=10, twist=(0.1, 0.17))
    kw = dict(embed_dim=16, depth=2, n_heads=4, n_freq=4)
    arch = opf.Arch(N, g["Gs"], g["W"], L, **kw)
    theta = opf.init_theta(arch, 4, backflow_scale=0.1, scale_all=0.02)
    model = ContinuumSpinorPsiformer(n_elec=N, Gs=g["Gs"], W=g["W"], L=L, **kw)
    return N, L, arch, theta, make_logpsi(model, theta.cuda())


def test_metropolis_decisions_bit_exact():
    N, L, arch, theta, lp = _mcmc_case("base")
    Wk, ns, nf = 32, 3, 2
    rng = np.random.default_rng(7)
    R0 = rng.uniform(0, L, (Wk, N, 2))
    S0 = rng.integers(0, 2, (Wk, N))
    st = osa.make_streams(ns, Wk, N, nf, seed=8)
    ch = Chain(lp, R0, S0)
    # oracle: the same sweeps with the oracle's log psi
    logpsi_b = lambda R, S: opf.batched_logpsi(arch, theta, torch.tensor(R), torch.tensor(S)).numpy()
    R, S, lpc = R0.copy(), S0.copy(), None
    acc_o = np.zeros(Wk, int)
    accs_o = np.zeros(Wk, int)
    for s in range(ns):
        R, S, a, asp, lpc = osa.sweep(logpsi_b, R, S, np.eye(2) * L, 0.5, st["xi"][s], st["u_pos"][s],
                                      st["flip_i"][s], st["u_flip"][s], lp_cur=lpc)
        acc_o += a
        accs_o += asp.sum(0)
    ap, asp = ch.run(ns, 0.5, nf, streams=st)
    assert np.array_equal(np.rint(ap.cpu().numpy() * ns).astype(int), acc_o)
    assert np.array_equal(np.rint(asp.cpu().numpy() * ns * nf).astype(int), accs_o)
    assert np.array_equal(ch.S.cpu().numpy(), S)
    assert np.array_equal(ch.R.cpu().numpy(), R)          # bit for bit


def test_metropolis_out_of_range_flip_indices_are_rejected():
    """Flip indices outside [0, N) in a caller's stream are rejected proposals inside the kernels (no
    host-side check runs): the result equals the oracle's with those proposals forced to reject."""
    N, L, arch, theta, lp = _mcmc_case("base")
    Wk, ns, nf = 32, 2, 2
    rng = np.random.default_rng(17)
    R0 = rng.uniform(0, L, (Wk, N, 2))
    S0 = rng.integers(0, 2, (Wk, N))
    st = osa.make_streams(ns, Wk, N, nf, seed=18)
    bad = rng.random(st["flip_i"].shape) < 0.4
    st["flip_i"] = np.where(bad, rng.choice([-1, N, N + 7, -(2 ** 31)], st["flip_i"].shape), st["flip_i"])
    ref = dict(st, flip_i=np.where(bad, 0, st["flip_i"]), u_flip=np.where(bad, 2.0, st["u_flip"]))
    logpsi_b = lambda R, S: opf.batched_logpsi(arch, theta, torch.tensor(R), torch.tensor(S)).numpy()
    R, S, lpc = R0.copy(), S0.copy(), None
    accs_o = np.zeros(Wk, int)
    for s in range(ns):
        R, S, _, asp, lpc = osa.sweep(logpsi_b, R, S, np.eye(2) * L, 0.5, ref["xi"][s], ref["u_pos"][s],
                                      ref["flip_i"][s], ref["u_flip"][s], lp_cur=lpc)
        accs_o += asp.sum(0)
    ch = Chain(lp, R0, S0)
    _, asp = ch.run(ns, 0.5, nf, streams=st)
    assert np.array_equal(np.rint(asp.cpu().numpy() * ns * nf).astype(int), accs_o)
    assert np.array_equal(ch.S.cpu().numpy(), S)
    assert np.array_equal(ch.R.cpu().numpy(), R)


def test_sampler_density_uniform_and_energy():
    """test_nqs_sampler.py:31-56 on the HIP path (plane-wave determinant: uniform density)."""
    from nqs.local_energy import make_local_energy
    from nqs.sampler import vmc_energy
    g = bands.rashba_gs(9.0, 9, 1.0, 1.0, 0.5, n_cut=10, twist=(0.1, 0.17))
    lp = make_logpsi_general(g["Gs"], g["W"], 9.0)
    R, S, ap, asp = run_chain(lp, 9.0, 9, 1, n_walkers=2048, n_sweeps=150, burn=100, step=0.5)
    assert 0.1 < ap < 0.95
    pos = R.cpu().numpy().reshape(-1, 2) % 9.0
    H, _, _ = np.histogram2d(pos[:, 0], pos[:, 1], bins=6, range=[[0, 9.0], [0, 9.0]])
    assert np.std(H) / H.mean() < 1.8 / np.sqrt(H.mean())
    Em, Ee, _ = vmc_energy(make_local_energy(lp, 1.0, 1.0, 0.5), R[:256], S[:256])
    assert abs(Em - g["E_gs"]) < 1e-6
