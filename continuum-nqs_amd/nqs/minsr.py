"""MinSR natural gradient on the HIP path (call surface of src/nqs/minsr.py:30-159).

  real_ravel(params)                       -> (theta, unravel)
  per_sample_O(model, theta, unravel, R, S) -> O (W, P) complex   (hand-written backward kernel)
  minsr_direction(O, eps, lam)             -> dtheta (P,)          (MFMA Gram, float64 Cholesky)
  paramsr_direction(O, eps, lam)           -> the explicit P x P solve, as a check (torch on GPU)
  train_minsr(model, params, L, n_elec, key, ...)                  (the VMC loop, one step = VMCStep)

Hot-path layout: the O matrix is kept as the real (2W, ldo) array [Re O; Im O] that
nqs_psiformer_grad writes (float64 for precision="fp64" models, float32 for "mixed") and
nqs_minsr_direction consumes (read only with the bf16 Gram, centred in place otherwise).  Gram precision: "bf16" (MFMA bf16 in, f32
accumulate), "fp32" or "fp64"; the Cholesky, the solves and the refinement step are float64.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from .cell import as_cell
from .psiformer import LogPsi, as_walkers, ravel_params, _flatten
from .local_energy import make_local_energy
from .sampler import Chain


# ------------------------------------------------------------------ ravel
def real_ravel(pytree):
    """Flatten a nested dict to a real float64 vector (complex leaves -> [re, im]), key-sorted;
    returns (theta, unravel) as minsr.py:30-53."""
    leaves = _flatten(pytree)
    specs = [(name, tuple(torch.as_tensor(l).shape), torch.as_tensor(l).is_complex()) for name, l in leaves]
    theta = ravel_params(pytree)

    def unravel(v):
        v = torch.as_tensor(v)
        out, i = {}, 0
        for name, shape, cplx in specs:
            n = int(np.prod(shape)) if shape else 1
            if cplx:
                t = torch.complex(v[i:i + n], v[i + n:i + 2 * n]).reshape(shape)
                i += 2 * n
            else:
                t = v[i:i + n].reshape(shape)
                i += n
            node = out
            parts = name.split("/")
            for p in parts[:-1]:
                node = node.setdefault(p, {})
            node[parts[-1]] = t
        return out
    return theta, unravel


def _ldo(P):
    return int(math.ceil(P / 256.0) * 256)


class OMatrix:
    """The real O array [Re O; Im O] (2W, ldo) in the model's backbone precision."""

    def __init__(self, data, W, P):
        self.data, self.W, self.P = data, W, P

    @property
    def ldo(self):
        return self.data.shape[1]

    def complex(self):
        d = self.data[:, :self.P].double()
        return torch.complex(d[:self.W], d[self.W:])


def grad_params(logpsi, R, S, out=None, ws_tag="default"):
    """O rows of the walkers (R, S) by the backward kernel; returns (OMatrix, log psi).  ws_tag names
    the device workspace (calls running concurrently on different streams need their own)."""
    lib = _lib.load()
    R, S, _ = as_walkers(R, S)
    W = R.shape[0]
    m = logpsi.model
    d = m.desc()
    P = m.n_params()
    dt = torch.float64 if m.precision == "fp64" else torch.float32
    if out is None or out.data.shape != (2 * W, _ldo(P)) or out.data.dtype != dt:
        # The backward writes every column below P; only the pad columns need clearing.
        buf = torch.empty(2 * W, _ldo(P), dtype=dt, device=R.device)
        buf[:, P:].zero_()
        out = OMatrix(buf, W, P)
    lp = torch.empty(W, 2, dtype=torch.float64, device=R.device)
    nb = lib.nqs_workspace_bytes(ctypes.byref(d), _lib.OP_GRAD, W)
    ws = _lib.workspace(nb, ws_tag)
    _lib.check(lib.nqs_psiformer_grad(ctypes.byref(d), _lib.ptr(logpsi.packed), _lib.ptr(R), _lib.ptr(S), W,
                                      _lib.ptr(out.data), out.ldo, _lib.ptr(lp), _lib.ptr(ws), ws.numel(),
                                      _lib.stream()), "nqs_psiformer_grad")
    return out, torch.view_as_complex(lp)


def per_sample_O(model, theta, unravel, R, S):
    """O[s, k] = d log psi(x_s) / d theta_k (complex; real params) -> (W, P) complex128."""
    th = theta if torch.is_tensor(theta) and theta.dim() == 1 else ravel_params(theta)
    O, _ = grad_params(LogPsi(model, th), R, S)
    return O.complex()


# ------------------------------------------------------------------ the sample-space solve
_GRAM = {"fp64": 0, "fp32": 1, "bf16": 2}


def minsr_direction(O, eps, lam, gram="auto", refine=1):
    """dtheta = Otil^T (Otil Otil^T + lam I)^-1 etil (minsr.py:76-81).  O: an OMatrix or a (W, P)
    complex tensor; an OMatrix is left unchanged by the bf16 Gram and centred in place by the
    fp32 / fp64 Gram."""
    lib = _lib.load()
    dev = _lib.device()
    if not isinstance(O, OMatrix):
        Oc = torch.as_tensor(O).to(dev)
        W, P = Oc.shape
        data = torch.zeros(2 * W, _ldo(P), dtype=torch.float64, device=dev)
        data[:W, :P] = Oc.real
        data[W:, :P] = Oc.imag
        O = OMatrix(data, W, P)
    W, P = O.W, O.P
    fp64 = O.data.dtype == torch.float64
    if gram == "auto":
        gram = "fp64" if fp64 else "bf16"
    eps = torch.as_tensor(eps).to(device=dev, dtype=torch.complex128).contiguous()
    d = lib_minsr_desc(W, P, O.ldo, fp64, gram, refine, lam)
    nb = lib.nqs_minsr_workspace_bytes(ctypes.byref(d))
    if nb < 0:
        _lib.check(-1, "nqs_minsr_workspace_bytes")
    ws = _lib.workspace(nb, "minsr")
    dth = torch.empty(P, dtype=torch.float64, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    _lib.check(lib.nqs_minsr_direction(ctypes.byref(d), _lib.ptr(O.data), _lib.ptr(torch.view_as_real(eps)),
                                       _lib.ptr(dth), _lib.ptr(info), _lib.ptr(ws), ws.numel(), _lib.stream()),
               "nqs_minsr_direction")
    minsr_direction.last_info = info
    return dth


class MinSRSplit:
    """minsr_direction in three calls on one workspace (nqs_minsr_gram, nqs_minsr_factor,
    nqs_minsr_solve; together exactly nqs_minsr_direction): gram() centres O and forms the shifted
    Gram, factor() its Cholesky factor (info), solve(eps) the direction.  Each call runs on the
    current torch stream; the caller orders them (factor may run on another stream than the local
    energies, which the Gram does not need)."""

    def __init__(self, O, lam, gram="auto", refine=1):
        self.lib = _lib.load()
        dev = _lib.device()
        fp64 = O.data.dtype == torch.float64
        if gram == "auto":
            gram = "fp64" if fp64 else "bf16"
        self.O = O
        self.d = lib_minsr_desc(O.W, O.P, O.ldo, fp64, gram, refine, lam)
        nb = self.lib.nqs_minsr_workspace_bytes(ctypes.byref(self.d))
        if nb < 0:
            _lib.check(-1, "nqs_minsr_workspace_bytes")
        self.ws = _lib.workspace(nb, "minsr")
        self.info = torch.zeros(1, dtype=torch.int32, device=dev)

    def gram(self):
        _lib.check(self.lib.nqs_minsr_gram(ctypes.byref(self.d), _lib.ptr(self.O.data), _lib.ptr(self.ws),
                                           self.ws.numel(), _lib.stream()), "nqs_minsr_gram")

    def factor(self):
        _lib.check(self.lib.nqs_minsr_factor(ctypes.byref(self.d), _lib.ptr(self.info), _lib.ptr(self.ws),
                                             self.ws.numel(), _lib.stream()), "nqs_minsr_factor")

    def solve(self, eps):
        dev = _lib.device()
        eps = torch.as_tensor(eps).to(device=dev, dtype=torch.complex128).contiguous()
        dth = torch.empty(self.O.P, dtype=torch.float64, device=dev)
        _lib.check(self.lib.nqs_minsr_solve(ctypes.byref(self.d), _lib.ptr(self.O.data),
                                            _lib.ptr(torch.view_as_real(eps)), _lib.ptr(dth), _lib.ptr(self.ws),
                                            self.ws.numel(), _lib.stream()), "nqs_minsr_solve")
        return dth


def lib_minsr_desc(W, P, ldo, fp64, gram, refine, lam):
    d = _lib.MinsrDesc()
    d.n_walkers, d.n_params, d.ldo = W, P, ldo
    d.o_dtype = _lib.NQS_FP64 if fp64 else _lib.NQS_MIXED
    d.gram = _GRAM[gram]
    d.refine = int(refine)
    d.lam = float(lam)
    return d


def paramsr_direction(O, eps, lam):
    """The explicit parameter-space solve (minsr.py:84-88), torch linear algebra on the GPU; a
    check for small P, not part of the step."""
    dev = _lib.device()
    O = torch.as_tensor(O.complex() if isinstance(O, OMatrix) else O).to(dev)
    eps = torch.as_tensor(eps).to(dev, torch.complex128)
    dO = O - O.mean(0, keepdim=True)
    Ot = torch.cat([dO.real, dO.imag], 0)
    ec = eps - eps.mean()
    et = torch.cat([ec.real, ec.imag])
    P = Ot.shape[1]
    return torch.linalg.solve(Ot.T @ Ot + lam * torch.eye(P, dtype=torch.float64, device=dev), Ot.T @ et)


def clip_energies(e, clip_k=5.0):
    """Median/MAD clip (minsr.py:145-148) -> (e_clip complex (W,), stats [mean, err, median, MAD])."""
    lib = _lib.load()
    e = torch.as_tensor(e).to(_lib.device(), torch.complex128).contiguous()
    out = torch.empty_like(e)
    stats = torch.empty(4, dtype=torch.float64, device=e.device)
    _lib.check(lib.nqs_clip_energies(_lib.ptr(torch.view_as_real(e)), e.shape[0], float(clip_k),
                                     _lib.ptr(torch.view_as_real(out)), _lib.ptr(stats), _lib.stream()),
               "nqs_clip_energies")
    return out, stats


# ------------------------------------------------------------------ one VMC step
class VMCStep:
    """Persistent-walker MinSR VMC step (minsr.py:139-156): sweeps -> E_loc (+ Coulomb) ->
    clip -> O -> MinSR direction -> theta -= lr dtheta.  All device-resident."""

    def __init__(self, model, theta, R, S, m_star=1.0, lambda_r=1.0, h_z=0.0, ext_field=None, coulomb=None,
                 lr=0.05, lam=1e-3, n_sweeps=10, step_size=0.4, clip_k=5.0, n_spin_flips=2, gram="auto",
                 refine=1, graph=False, pinning=None, overlap=True):
        self.model = model
        self.theta = theta.to(_lib.device(), torch.float64).contiguous()
        self.lp = LogPsi(model, self.theta)
        self.chain = Chain(self.lp, R, S)
        self.ham = dict(m_star=m_star, lambda_r=lambda_r, h_z=h_z, ext_field=ext_field, pinning=pinning)
        self.coulomb = coulomb
        self.lr, self.lam, self.n_sweeps, self.step_size, self.clip_k = lr, lam, n_sweeps, step_size, clip_k
        self.n_spin_flips, self.gram, self.refine = n_spin_flips, gram, refine
        self.O = None
        self.last_info = None
        self.timers = None
        self.graph = graph      # replay the sweeps from a captured HIP graph (nqs.graphs.GraphSweeps)
        self._gs = None
        # factorise the Gram on a side stream (default priority) while the local energies run
        self.overlap = overlap
        self._side = None
        self.failed = torch.zeros((), dtype=torch.int64, device=_lib.device())   # skipped updates

    @property
    def failed_steps(self):
        """Number of steps whose update was skipped (non-positive-definite Gram or non-finite
        direction); reading it synchronises with the device."""
        return int(self.failed.item())

    def _ev(self):
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        return e

    def energies(self):
        E = make_local_energy(self.lp, **self.ham)
        e = E(self.chain.R, self.chain.S)
        if self.coulomb is not None:
            e = e + self.coulomb(self.chain.R)
        return e

    def step(self, key=0, streams=None):
        """One step; `streams` (sampler.make_streams layout) replaces the sweeps' random draws."""
        t = [self._ev()] if self.timers is not None else None
        if self.graph:
            if self._gs is None:
                from .graphs import GraphSweeps
                self._gs = GraphSweeps(self.chain, self.n_sweeps, self.step_size, self.n_spin_flips)
            ap, asp = self._gs.run(key=key, streams=streams)
        else:
            ap, asp = self.chain.run(self.n_sweeps, self.step_size, self.n_spin_flips, key=key, streams=streams)
        if t: t.append(self._ev())
        # O, then the centred Gram; its Cholesky factor does not need the local energies, so it runs
        # on a side stream beside them (overlap), and the solve waits for both
        self.O, _ = grad_params(self.lp, self.chain.R, self.chain.S, self.O)
        if t: t.append(self._ev())
        ms = MinSRSplit(self.O, self.lam, self.gram, self.refine)
        ms.gram()
        if t: t.append(self._ev())
        main = torch.cuda.current_stream()
        if self.overlap:
            if self._side is None:
                self._side = torch.cuda.Stream()
            self._side.wait_stream(main)
            with torch.cuda.stream(self._side):
                ms.factor()
        else:
            ms.factor()
        e = self.energies()
        e_clip, stats = clip_energies(e, self.clip_k)
        if t: t.append(self._ev())
        if self.overlap:
            main.wait_stream(self._side)
        d = ms.solve(e_clip)
        info = self.last_info = ms.info  # 0, or 1 + the first non-positive pivot (device int32)
        if t: t.append(self._ev())
        # A step whose shifted Gram was not positive definite in this precision (info != 0) or whose
        # direction is not finite leaves theta unchanged; the decision stays on the device (no host
        # sync) and is counted in self.failed (device int64), which train_minsr reports.
        ok = (info[0] == 0) & torch.isfinite(d).all()
        d = torch.where(ok, d, torch.zeros_like(d))
        self.failed += (~ok).to(torch.int64)
        self.theta -= self.lr * d
        if self.graph:
            # same buffers: the captured sweeps keep reading them
            self.lp.update(self.theta)
            self.chain.logp.copy_(2.0 * self.lp(self.chain.R, self.chain.S).real)
        else:
            self.lp = LogPsi(self.model, self.theta)
            self.chain.logpsi = self.lp
            self.chain.logp = (2.0 * self.lp(self.chain.R, self.chain.S).real).contiguous()
        if t:
            t.append(self._ev())
            self.timers.append(t)
        self.last = dict(e=e, stats=stats, acc=ap, acc_spin=asp, dtheta=d)
        return stats


def train_minsr(model, params, L, n_elec, key, *, m_star=1.0, lambda_r=1.0, h_z=0.0, ext_field=None, coulomb=None,
                use_folx=False, n_walkers=256, n_steps=100, lr=0.05, lam=1e-3, n_sweeps=10, step_size=0.4, burn=80,
                clip_k=5.0, log_every=20, history=None, R0=None, S0=None, return_walkers=False, gram="auto",
                pinning=None):
    """Optimise the Psiformer with MinSR (minsr.py:92-159).  Returns (params, history) or
    (params, history, R, S)."""
    dev = _lib.device()
    theta, unravel = real_ravel(params)
    A = as_cell(L)
    g = torch.Generator(device=dev)
    g.manual_seed(int(key))
    if R0 is None:
        R = torch.rand(n_walkers, n_elec, 2, generator=g, dtype=torch.float64, device=dev) @ torch.tensor(A.T, device=dev)
        S = torch.randint(0, 2, (n_walkers, n_elec), generator=g, device=dev).to(torch.int32)
    else:
        R = torch.as_tensor(np.asarray(R0) if not torch.is_tensor(R0) else R0).to(dev, torch.float64)
        S = torch.as_tensor(np.asarray(S0) if not torch.is_tensor(S0) else S0).to(dev, torch.int32)
    st = VMCStep(model, theta, R, S, m_star=m_star, lambda_r=lambda_r, h_z=h_z, ext_field=ext_field,
                 coulomb=coulomb, lr=lr, lam=lam, n_sweeps=n_sweeps, step_size=step_size, clip_k=clip_k, gram=gram,
                 pinning=pinning)
    seed = int(torch.randint(0, 2 ** 30, (1,), generator=g, device=dev).item())
    if R0 is None and burn > 0:
        st.chain.run(burn, step_size, 2, key=seed)
    hist = history if history is not None else {"E": [], "Eerr": []}
    hist.setdefault("failed_steps", [])
    for it in range(n_steps):
        stats = st.step(key=seed + 1000 + it)
        s = stats.cpu().numpy()
        hist["E"].append(float(s[0]))
        hist["Eerr"].append(float(s[1]))
        hist["failed_steps"].append(st.failed_steps)
        if log_every and (it % log_every == 0 or it == n_steps - 1):
            print(f"    minsr {it:4d}: E={s[0]:+.6f} +/- {s[1]:.1e} acc={float(st.last['acc'].mean()):.2f}",
                  flush=True)
    if st.failed_steps:
        import warnings
        warnings.warn(f"train_minsr: {st.failed_steps} of {n_steps} updates skipped (the shifted Gram was not "
                      f"positive definite in the {st.gram} precision, or the direction was not finite)")
    out = unravel(st.theta.cpu())
    if return_walkers:
        return out, hist, st.chain.R, st.chain.S
    return out, hist
