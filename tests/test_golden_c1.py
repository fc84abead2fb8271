"""The committed two-electron trajectory (tests/golden/c1_minsr_steps.npz) against the current oracle,
on the CPU: the burn sweeps and step 0's sweeps bit for bit and step 0's local energies of 4
walkers.  A change of the oracle that the fixture no longer reflects fails here."""
import importlib.util
import os

import numpy as np
import torch

from oracle import local_energy as ole, psiformer as opf, sampler as osa

_HERE = os.path.dirname(os.path.abspath(__file__))
_spec = importlib.util.spec_from_file_location("make_c1", os.path.join(_HERE, "golden", "make_c1_minsr_steps.py"))
c1 = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(c1)


def test_c1_fixture_reproduces_with_the_oracle():
    gold = np.load(os.path.join(_HERE, "golden", "c1_minsr_steps.npz"))
    a = c1.arch()
    theta = opf.init_theta(a, 0)
    p = opf.unravel(a, theta)
    lpb = torch.func.vmap(lambda r, s: opf.logpsi_from_params(a, p, r, s))
    logpsi_b = lambda R_, S_: lpb(torch.as_tensor(R_), torch.as_tensor(S_)).numpy()
    R, S = c1.initial_walkers()
    st = osa.make_streams(c1.BURN, c1.W, c1.N, 2, c1.BURN_SEED)
    with np.errstate(invalid="ignore"):
        for s in range(c1.BURN):
            R, S, _, _, _ = osa.sweep(logpsi_b, R, S, a.A, c1.STEP, st["xi"][s], st["u_pos"][s], st["flip_i"][s],
                                      st["u_flip"][s])
    assert np.array_equal(R, gold["burn_R"]) and np.array_equal(S, gold["burn_S"])
    st = osa.make_streams(c1.N_SWEEPS, c1.W, c1.N, 2, c1.STEP_SEED0)
    with np.errstate(invalid="ignore"):
        for s in range(c1.N_SWEEPS):
            R, S, _, _, _ = osa.sweep(logpsi_b, R, S, a.A, c1.STEP, st["xi"][s], st["u_pos"][s], st["flip_i"][s],
                                      st["u_flip"][s])
    assert np.array_equal(R, gold["step_R"][0]) and np.array_equal(S, gold["step_S"][0])
    E = ole.make_local_energy(opf.make_logpsi(a, theta), 1.0, c1.LAM_R, c1.H_Z)
    ce = c1.coulomb()
    e = np.array([E(torch.tensor(R[w]), torch.tensor(S[w])) + ce(R[w]) for w in range(4)])
    assert np.max(np.abs(e - gold["step_e"][0][:4])) < 1e-13
    # every walker is in the singlet sector after the burn-in (parallel spins have zero amplitude in
    # the k = 0 closed-shell envelope, so flips into them are always rejected)
    assert np.all(gold["step_S"].sum(-1) == 1) and np.all(gold["burn_S"].sum(-1) == 1)
