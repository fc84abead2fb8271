"""Exact spinor Slater determinants on the HIP path (call surface of src/nqs/slater_reference.py:21-45).

A multi-G spinor Slater determinant is the Psiformer's determinant head with the backflow
switched off, so it runs through the same kernels: a depth-0 backbone whose backflow weights are
zero multiplies nothing into M = env, leaving log psi = log det env exactly.
"""
import numpy as np

from .psiformer import ContinuumSpinorPsiformer, LogPsi, ravel_params
from .rashba_reference import planewave_envelope


def make_logpsi_general(Gs, W, L=1.0, precision="fp64"):
    """logpsi(R, S) of det[phi_a(r_i, s_i)], phi_a(r, s) = sum_G W[a, G, s] exp(i G.r).
    The cell L only enters the (zero-weighted) features; any value gives the same log psi."""
    Gs = np.asarray(Gs, float)
    W = np.asarray(W, complex)
    model = ContinuumSpinorPsiformer(n_elec=W.shape[0], Gs=Gs, W=W, L=L, embed_dim=8, depth=0, n_heads=1,
                                     n_freq=1, precision=precision)
    return LogPsi(model, ravel_params(model.init(0)))


def make_logpsi_slater(Kvecs, Chi, L=1.0):
    Gs, W = planewave_envelope(Kvecs, Chi)
    return make_logpsi_general(Gs, W, L)
