"""A batched ensemble of independent systems of uneven electron count (no reference counterpart:
the reference runs one jit-compiled system of fixed N at a time).

Each system has its own cell (L = rs sqrt(pi N)), couplings (lambda_R, h_z), plane-wave spinor
envelope, parameters and walkers; all share one architecture.  A step advances every system:

  sweeps   one HIP graph replay with every system's Metropolis sweeps (nqs.graphs.EnsembleSweeps,
           side-stream branches so that small systems overlap);
  E_loc    per system (forward-Laplacian kinetic + SOC + Zeeman, Ewald), median/MAD clip, O rows,
           spread over the same number of side streams (each system with its own workspaces) so that
           small systems run beside large ones;
  MinSR    one nqs_minsr_direction_batched call: the bf16 Grams of all systems in one launch, the
           Cholesky factorizations and triangular solves of all systems sharing each launch;
  update   theta -= lr dtheta per system, repacked in place so the captured graph stays valid.

make_c5_systems draws the configuration of the 24-system benchmark: N from {4, 6, 8, 12, 16, 24,
32}, rs in [1, 6], lambda_R in [0, 1], h_z in [0, 0.2].
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from .psiformer import ContinuumSpinorPsiformer, LogPsi, ravel_params
from .rashba_reference import noninteracting_gs, occupied_orbital_matrix, planewave_envelope
from .sampler import Chain
from .local_energy import make_local_energy
from .ewald import make_ewald
from .minsr import grad_params, clip_energies, lib_minsr_desc
from .graphs import EnsembleSweeps

C5_N = (4, 6, 8, 12, 16, 24, 32)
C5_ARCH = dict(embed_dim=64, depth=3, n_heads=4, n_freq=8, n_dets=4)


def make_c5_systems(n_systems=24, seed=0):
    """[{N, rs, lambda_r, h_z}] of the uneven ensemble benchmark, seeded."""
    rng = np.random.default_rng(seed)
    return [dict(N=int(rng.choice(C5_N)), rs=float(rng.uniform(1.0, 6.0)), lambda_r=float(rng.uniform(0.0, 1.0)),
                 h_z=float(rng.uniform(0.0, 0.2))) for _ in range(n_systems)]


def minsr_direction_batched(Os, eps, lams, gram="auto", refine=1):
    """MinSR directions of several systems with the same walker count in one library call.
    Os: OMatrix per system, eps: (W,) complex per system, lams: shift per system.  Returns
    (list of dtheta, info (n_sys,) int32 device tensor)."""
    lib = _lib.load()
    dev = _lib.device()
    ns = len(Os)
    fp64 = Os[0].data.dtype == torch.float64
    if gram == "auto":
        gram = "fp64" if fp64 else "bf16"
    descs = (_lib.MinsrDesc * ns)()
    for i, (O, lam) in enumerate(zip(Os, lams)):
        d = lib_minsr_desc(O.W, O.P, O.ldo, fp64, gram, refine, lam)
        descs[i] = d
    nb = lib.nqs_minsr_batched_workspace_bytes(descs, ns)
    if nb < 0:
        _lib.check(-1, "nqs_minsr_batched_workspace_bytes")
    ws = _lib.workspace(nb, "minsr_batched")
    eps = [torch.as_tensor(e).to(dev, torch.complex128).contiguous() for e in eps]
    dth = [torch.empty(O.P, dtype=torch.float64, device=dev) for O in Os]
    info = torch.zeros(ns, dtype=torch.int32, device=dev)
    arr = lambda ts: (ctypes.c_void_p * ns)(*[t.data_ptr() for t in ts])
    _lib.check(lib.nqs_minsr_direction_batched(descs, ns, arr([O.data for O in Os]),
                                               arr([torch.view_as_real(e) for e in eps]), arr(dth), _lib.ptr(info),
                                               _lib.ptr(ws), ws.numel(), _lib.stream()),
               "nqs_minsr_direction_batched")
    return dth, info


class System:
    """One member of the ensemble: model, parameters, walkers, Hamiltonian, Ewald."""

    def __init__(self, N, rs, lambda_r, h_z, arch, n_walkers, precision, gemm, seed, n_real, n_recip):
        self.N, self.rs, self.lambda_r, self.h_z = int(N), float(rs), float(lambda_r), float(h_z)
        self.L = rs * math.sqrt(math.pi * N)
        g = noninteracting_gs(self.N, self.L, 1.0, self.lambda_r, self.h_z, n_cut=10, twist=(0.1, 0.17))
        Gs, Wenv = planewave_envelope(*occupied_orbital_matrix(g))
        self.E_gs = float(g["E_gs"])
        self.model = ContinuumSpinorPsiformer(n_elec=self.N, Gs=Gs, W=Wenv, L=self.L, precision=precision,
                                              gemm=gemm, **arch)
        params = self.model.init(seed)
        rng = np.random.default_rng(seed + 1)
        kr = params["params"]["backflow"]["kernel"]
        z = torch.tensor(rng.standard_normal(tuple(kr.shape)) * 0.05)
        params["params"]["backflow"]["kernel"] = (z + 1j * z).to(torch.complex128)
        dev = _lib.device()
        self.theta = ravel_params(params).to(dev).contiguous()
        self.lp = LogPsi(self.model, self.theta)
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed + 2)
        R = torch.rand(n_walkers, self.N, 2, generator=gen, dtype=torch.float64, device=dev) * self.L
        S = torch.randint(0, 2, (n_walkers, self.N), generator=gen, device=dev).to(torch.int32)
        self.chain = Chain(self.lp, R, S)
        self.step_size = 0.3 * self.rs
        _, _, self.coulomb = make_ewald(self.L, 1.0, n_real=n_real, n_recip=n_recip)
        self.E = make_local_energy(self.lp, 1.0, self.lambda_r, self.h_z)
        self.O = None
        self.ws_tag = f"ensemble{seed}"  # own workspaces: systems' E_loc and O run concurrently
        self.E.ws_tag = self.ws_tag + "_eloc"


class Ensemble:
    """Independent systems advanced together; see the module docstring."""

    def __init__(self, systems, arch=None, n_walkers=1024, precision="mixed", gemm="bf16x3", lr=0.05, lam=1e-4,
                 n_sweeps=10, clip_k=5.0, n_spin_flips=2, gram="auto", refine=1, graph=True, seed=0, n_real=4,
                 n_recip=8, branches=4, concurrent=True):
        arch = dict(C5_ARCH if arch is None else arch)
        self.systems = [System(s["N"], s["rs"], s["lambda_r"], s["h_z"], arch, n_walkers, precision, gemm,
                               seed + 17 * i, n_real, n_recip) for i, s in enumerate(systems)]
        self.n_walkers = n_walkers
        self.lr, self.lam, self.n_sweeps, self.clip_k, self.nf = lr, lam, n_sweeps, clip_k, n_spin_flips
        self.gram, self.refine = gram, refine
        self.graph = graph
        self.branches = branches
        self.concurrent = concurrent
        self._side = None
        self._sweeps = None
        self.failed = np.zeros(len(self.systems), int)

    def burn(self, n, key=0):
        for i, s in enumerate(self.systems):
            s.chain.run(n, s.step_size, self.nf, key=key + 101 * i)

    def sweep(self, key=0):
        """Advance every system by n_sweeps sweeps; per-system mean acceptance (pos, spin)."""
        keys = [key * 1009 + i for i in range(len(self.systems))]
        if self.graph:
            if self._sweeps is None:
                self._sweeps = EnsembleSweeps([s.chain for s in self.systems], self.n_sweeps,
                                              [s.step_size for s in self.systems], self.nf, self.branches)
            rates = self._sweeps.run(keys)
        else:
            rates = [s.chain.run(self.n_sweeps, s.step_size, self.nf, key=k) for s, k in zip(self.systems, keys)]
        return [(float(a.mean()), float(b.mean())) for a, b in rates]

    def step(self, key=0, timers=None):
        """One VMC step of every system; returns (n_sys, 4) statistics [mean, err, median, MAD].
        timers: a list that receives a dict of stage times in ms (sweeps, eloc_clip_grad, minsr, update)."""
        ev = []

        def mark():
            if timers is not None:
                e = torch.cuda.Event(enable_timing=True)
                e.record()
                ev.append(e)
        mark()
        self.sweep(key)
        mark()
        n = len(self.systems)
        eps, stats = [None] * n, [None] * n
        main = torch.cuda.current_stream()
        if self.concurrent:
            nb = max(1, min(int(self.branches), n))
            if self._side is None:
                self._side = [torch.cuda.Stream() for _ in range(nb)]  # default priority
                # largest first onto the least loaded stream, cost ~ N^3 (determinants and jets)
                load = [0.0] * nb
                self._branch = [0] * n
                for i in sorted(range(n), key=lambda i: -self.systems[i].N):
                    b = min(range(nb), key=lambda b: load[b])
                    self._branch[i] = b
                    load[b] += float(self.systems[i].N) ** 3
            for st in self._side:
                st.wait_stream(main)
        for i, s in enumerate(self.systems):
            st = self._side[self._branch[i]] if self.concurrent else main
            with torch.cuda.stream(st):
                e = s.E(s.chain.R, s.chain.S) + s.coulomb(s.chain.R)
                eps[i], stats[i] = clip_energies(e, self.clip_k)
                s.O, _ = grad_params(s.lp, s.chain.R, s.chain.S, s.O, ws_tag=s.ws_tag + "_grad")
            if st is not main:
                eps[i].record_stream(main)
                stats[i].record_stream(main)
        if self.concurrent:
            for st in self._side:
                main.wait_stream(st)
        mark()
        dth, info = minsr_direction_batched([s.O for s in self.systems], eps, [self.lam] * n, self.gram, self.refine)
        mark()
        ok = (info == 0).cpu().numpy()
        for i, (s, d) in enumerate(zip(self.systems, dth)):
            if not ok[i] or not bool(torch.isfinite(d).all()):
                self.failed[i] += 1
                continue
            s.theta -= self.lr * d
            s.lp.update(s.theta)
            s.chain.logp.copy_(2.0 * s.lp(s.chain.R, s.chain.S).real)
        self.last_info = info
        mark()
        if timers is not None:
            torch.cuda.synchronize()
            ms = [a.elapsed_time(b) for a, b in zip(ev[:-1], ev[1:])]
            timers.append(dict(zip(("sweeps", "eloc_clip_grad", "minsr", "update"), ms)))
        return torch.stack(stats).cpu().numpy()
