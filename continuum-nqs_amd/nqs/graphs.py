"""Metropolis sweeps captured once in a HIP graph and replayed every step.

nqs_mcmc_sweeps launches, per sweep, one proposal kernel, a full log psi forward (about 40 kernels)
and one accept kernel for the joint move and for each spin flip; over 10 sweeps that is more than a
thousand launches.  They allocate nothing, synchronise nothing and write no memset, so the whole
call can be captured into a graph on torch's capture stream and replayed with one launch.

The random streams are inputs: they live in static device buffers that `run` fills (from a seeded
generator, or from caller-given streams) before each replay, so a replay makes exactly the
decisions an eager Chain.run makes with the same streams.  Acceptance counters are cleared with a
fill kernel outside the graph.

EnsembleSweeps captures the sweeps of many independent chains (systems of different size) into
one graph, spread over a few branches (side streams forked from and joined back to the capture
stream) so that small systems can run side by side: one graph launch advances every system.
"""
import ctypes

import torch

from . import _lib
from .sampler import Chain, make_streams, _dev_streams


class _Slot:
    """Static buffers and the sweep call of one chain."""

    def __init__(self, chain, n_sweeps, step, n_spin_flips):
        if not isinstance(chain, Chain):
            raise TypeError("graph sweeps need an nqs.sampler.Chain")
        self.chain, self.n_sweeps, self.step, self.nf = chain, int(n_sweeps), float(step), int(n_spin_flips)
        lib = _lib.load()
        W, N = chain.R.shape[0], chain.R.shape[1]
        dev = chain.R.device
        st = make_streams(self.n_sweeps, W, N, self.nf, 0, dev)
        self.xi, self.up, self.fi, self.uf = _dev_streams(st, dev)
        self.acc_p = torch.zeros(W, dtype=torch.int32, device=dev)
        self.acc_s = torch.zeros(W, dtype=torch.int32, device=dev)
        self.desc = chain.logpsi.model.desc()
        nb = lib.nqs_workspace_bytes(ctypes.byref(self.desc), _lib.OP_MCMC, W)
        self.ws = torch.empty(int(nb), dtype=torch.uint8, device=dev)

    def launch(self, R, S, logp, stream):
        lib = _lib.load()
        W = R.shape[0]
        _lib.check(lib.nqs_mcmc_sweeps(ctypes.byref(self.desc), _lib.ptr(self.chain.logpsi.packed), _lib.ptr(R),
                                       _lib.ptr(S), _lib.ptr(logp), W, self.n_sweeps, self.nf, self.step,
                                       _lib.ptr(self.xi), _lib.ptr(self.up), _lib.ptr(self.fi), _lib.ptr(self.uf),
                                       _lib.ptr(self.acc_p), _lib.ptr(self.acc_s), _lib.ptr(self.ws),
                                       self.ws.numel(), ctypes.c_void_p(stream.cuda_stream)),
                   "nqs_mcmc_sweeps")

    def warm(self):
        """One eager call on copies of the state (sets kernel attributes and tunes launch shapes)."""
        c = self.chain
        self.launch(c.R.clone(), c.S.clone(), c.logp.clone(), torch.cuda.current_stream())

    def fill(self, key=0, streams=None):
        W, N = self.chain.R.shape[0], self.chain.R.shape[1]
        if streams is None:
            streams = make_streams(self.n_sweeps, W, N, self.nf, key, self.chain.R.device)
        xi, up, fi, uf = _dev_streams(streams, self.chain.R.device)
        self.xi.copy_(xi)
        self.up.copy_(up)
        self.fi.copy_(fi)
        self.uf.copy_(uf)
        self.acc_p.zero_()
        self.acc_s.zero_()

    def rates(self):
        return (self.acc_p.double() / max(self.n_sweeps, 1),
                self.acc_s.double() / max(self.n_sweeps * self.nf, 1))


class GraphSweeps:
    """Graph of `n_sweeps` sweeps (1 joint move + `n_spin_flips` flips each) of `chain`'s walkers
    with step `step`.  The chain's R, S, log|psi|^2 and log psi parameters are bound by address:
    keep the same Chain and LogPsi objects (re-capture after replacing them)."""

    def __init__(self, chain, n_sweeps, step, n_spin_flips=2):
        self.slot = _Slot(chain, n_sweeps, step, n_spin_flips)
        self.chain = chain
        self.slot.warm()
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.slot.launch(chain.R, chain.S, chain.logp, torch.cuda.current_stream())
        torch.cuda.synchronize()

    def run(self, key=0, streams=None):
        """Advance the chain by n_sweeps sweeps; returns per-walker acceptance rates (pos, spin)."""
        self.slot.fill(key, streams)
        self.graph.replay()
        return self.slot.rates()


class EnsembleSweeps:
    """One graph with the sweeps of every chain in `chains` (steps[i] for chain i), spread
    round-robin over `branches` side streams (default priority) joined back before the end."""

    def __init__(self, chains, n_sweeps, steps, n_spin_flips=2, branches=4):
        if len(steps) != len(chains):
            raise ValueError("one step size per chain")
        self.chains = list(chains)
        self.slots = [_Slot(c, n_sweeps, s, n_spin_flips) for c, s in zip(chains, steps)]
        for sl in self.slots:
            sl.warm()
        torch.cuda.synchronize()
        nb = max(1, min(int(branches), len(self.slots)))
        self.graph = torch.cuda.CUDAGraph()
        side = [torch.cuda.Stream() for _ in range(nb)]
        with torch.cuda.graph(self.graph):
            main = torch.cuda.current_stream()
            for s in side:
                s.wait_stream(main)
            for i, sl in enumerate(self.slots):
                s = side[i % nb]
                with torch.cuda.stream(s):
                    c = sl.chain
                    sl.launch(c.R, c.S, c.logp, s)
            for s in side:
                main.wait_stream(s)
        torch.cuda.synchronize()

    def run(self, keys=None, streams=None):
        """Advance every chain; keys[i] seeds chain i (or streams[i] gives its streams).  Returns
        the per-walker acceptance rates of each chain."""
        for i, sl in enumerate(self.slots):
            sl.fill(0 if keys is None else keys[i], None if streams is None else streams[i])
        self.graph.replay()
        return [sl.rates() for sl in self.slots]
