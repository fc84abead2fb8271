"""Continuum spinful Psiformer on the gfx950 HIP path (call surface of src/nqs/psiformer.py:100-196).

    model  = ContinuumSpinorPsiformer(n_elec=N, Gs=Gs, W=W, L=L, embed_dim=..., depth=..., ...)
    params = model.init(key)                # nested dict in the flax tree layout
    logpsi = make_logpsi(model, params)     # logpsi(R, S) -> complex128 (device tensor)

R is (N, 2) or (W, N, 2) float64, S is (N,) or (W, N) integer spins in {0, 1}; batches need no
vmap.  The log psi apply, its derivatives and everything downstream run in libnqs_hip.so; the
parameters live on the GPU as the flat real_ravel vector theta plus a packed kernel layout.
`precision="mixed"` runs the backbone (features, attention, MLP, LayerNorm, backflow) in
float32; the envelope, determinants, Jastrow and log-sum-exp stay float64 either way.  `gemm`
picks how a mixed model's dense layers use the matrix cores: "bf16x3" (default; each float32
operand split into a bf16 hi/lo pair, three bf16 products summed in float32), "fp32" (float32
MFMA) or "bf16" (operands rounded to bf16 once).
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from .cell import as_cell


class ContinuumSpinorPsiformer:
    def __init__(self, n_elec, Gs, W, L, embed_dim=32, depth=3, n_heads=4, ff_mult=2, n_freq=4, n_dets=1,
                 learnable_env=False, use_jastrow=False, jastrow_kind="rbf", n_knots=10, cusp=1.0,
                 precision="fp64", gemm="bf16x3"):
        self.n_elec = int(n_elec)
        self.Gs = np.asarray(Gs, dtype=float)
        self.W = np.asarray(W, dtype=complex)
        self.L = L
        self.A = as_cell(L)
        self.embed_dim, self.depth, self.n_heads, self.ff_mult = int(embed_dim), int(depth), int(n_heads), int(ff_mult)
        self.n_freq, self.n_dets = int(n_freq), int(n_dets)
        self.learnable_env, self.use_jastrow = bool(learnable_env), bool(use_jastrow)
        if jastrow_kind not in ("rbf", "cusp", "cusp_periodic"):
            raise ValueError("jastrow_kind must be 'rbf', 'cusp' or 'cusp_periodic'")
        self.jastrow_kind, self.n_knots, self.cusp = jastrow_kind, int(n_knots), float(cusp)
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        if gemm not in ("bf16x3", "fp32", "bf16"):
            raise ValueError("gemm must be 'bf16x3', 'fp32' or 'bf16'")
        self.gemm = gemm
        if self.W.shape != (self.n_elec, self.Gs.shape[0], 2):
            raise ValueError(f"W must have shape (n_elec, n_G, 2) = {(self.n_elec, self.Gs.shape[0], 2)}")
        self._dev = None

    # ------------------------------------------------------------------ descriptors / layout
    def desc(self):
        d = _lib.ModelDesc()
        d.n_elec, d.n_g, d.embed_dim, d.depth = self.n_elec, self.Gs.shape[0], self.embed_dim, self.depth
        d.n_heads, d.ff_mult, d.n_freq, d.n_dets = self.n_heads, self.ff_mult, self.n_freq, self.n_dets
        d.learnable_env, d.use_jastrow = int(self.learnable_env), int(self.use_jastrow)
        d.jastrow_kind = {"rbf": 0, "cusp": 1, "cusp_periodic": 2}[self.jastrow_kind]
        d.n_knots = self.n_knots
        d.precision = _lib.NQS_FP64 if self.precision == "fp64" else _lib.NQS_MIXED
        d.gemm = {"bf16x3": _lib.NQS_GEMM_BF16X3, "fp32": _lib.NQS_GEMM_FP32, "bf16": _lib.NQS_GEMM_BF16}[self.gemm]
        for i, v in enumerate(self.A.reshape(-1)):
            d.cell[i] = float(v)
        for i, v in enumerate(np.linalg.inv(self.A).reshape(-1)):
            d.cell_inv[i] = float(v)
        d.cusp = self.cusp
        return d

    def layout(self):
        """[(name, shape, is_complex, offset)] of theta, from the library (real_ravel order)."""
        lib = _lib.load()
        d = self.desc()
        out, i = [], 0
        name = ctypes.create_string_buffer(256)
        shape = (ctypes.c_int64 * 3)()
        nd, off, cp = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32()
        while lib.nqs_param_leaf(ctypes.byref(d), i, name, 256, shape, ctypes.byref(nd), ctypes.byref(off),
                                 ctypes.byref(cp)) == 0:
            out.append((name.value.decode(), tuple(shape[k] for k in range(nd.value)), bool(cp.value), off.value))
            i += 1
        return out

    def n_params(self):
        return int(_lib.load().nqs_param_count(ctypes.byref(self.desc())))

    # ------------------------------------------------------------------ parameters
    def init(self, key=0, R=None, S=None):
        """Parameters as a nested dict {'params': {...}} of float64 / complex128 CPU tensors.
        Dense kernels: truncated normal (+-2 sigma) with std sqrt(1/fan_in) (lecun normal, the
        flax Dense default); biases 0; LayerNorm scale 1; backflow, dW, det_logw, Jastrow 0
        (zero-initialised as in psiformer.py:124,139-141,157,180-181).  `key` seeds numpy."""
        seed = int(key) if np.ndim(key) == 0 else int(np.asarray(key).ravel()[-1])
        rng = np.random.default_rng(seed)
        tree = {}
        for name, shape, cplx, _ in self.layout():
            leaf = name.split("/")[-1]
            n = int(np.prod(shape)) if shape else 1
            if leaf == "kernel" and "backflow" not in name:
                std = math.sqrt(1.0 / shape[0]) / 0.87962566103423978
                z = rng.standard_normal(n)
                bad = np.abs(z) > 2.0
                while bad.any():
                    z[bad] = rng.standard_normal(int(bad.sum()))
                    bad = np.abs(z) > 2.0
                v = (std * z).reshape(shape)
            elif leaf == "scale":
                v = np.ones(shape)
            else:
                v = np.zeros(shape)
            t = torch.tensor(v, dtype=torch.complex128 if cplx else torch.float64)
            node = tree
            parts = name.split("/")
            for p in parts[:-1]:
                node = node.setdefault(p, {})
            node[parts[-1]] = t
        return tree

    # ------------------------------------------------------------------ device state
    def _device_consts(self):
        dev = _lib.device()
        if self._dev is None or self._dev[0] != dev:
            gs = torch.tensor(self.Gs, dtype=torch.float64, device=dev).contiguous()
            w = torch.view_as_real(torch.tensor(self.W, dtype=torch.complex128)).contiguous().to(dev)
            self._dev = (dev, gs, w)
        return self._dev

    def pack(self, theta, out=None):
        """theta (P,) float64 on the GPU -> packed kernel layout (uint8 device buffer; written into
        `out` when given, which must have the packed size)."""
        lib = _lib.load()
        dev, gs, w = self._device_consts()
        theta = theta.to(device=dev, dtype=torch.float64).contiguous()
        d = self.desc()
        nb = lib.nqs_packed_bytes(ctypes.byref(d))
        if out is not None:
            if out.numel() != nb or out.dtype != torch.uint8:
                raise ValueError("packed buffer of the wrong size")
            packed = out
        else:
            packed = torch.empty(int(nb), dtype=torch.uint8, device=dev)
        _lib.check(lib.nqs_pack_params(ctypes.byref(d), _lib.ptr(theta), _lib.ptr(gs), _lib.ptr(w),
                                       _lib.ptr(packed), _lib.stream()), "nqs_pack_params")
        return packed

    def apply(self, params, R, S):
        return make_logpsi(self, params)(R, S)


def _flatten(tree, prefix=()):
    if isinstance(tree, dict):
        out = []
        for k in sorted(tree.keys()):
            out += _flatten(tree[k], prefix + (k,))
        return out
    return [("/".join(prefix), tree)]


def ravel_params(params):
    """Nested dict -> flat float64 vector (complex leaves -> [re..., im...]), key-sorted."""
    flats = []
    for _, leaf in _flatten(params):
        t = torch.as_tensor(leaf)
        if t.is_complex():
            flats += [t.real.reshape(-1).to(torch.float64), t.imag.reshape(-1).to(torch.float64)]
        else:
            flats.append(t.reshape(-1).to(torch.float64))
    return torch.cat([f.cpu() for f in flats])


def as_walkers(R, S):
    """Accept (N,2)/(N,) or (W,N,2)/(W,N), numpy or torch; return device float64 / int32 batches."""
    dev = _lib.device()
    R = torch.as_tensor(np.asarray(R) if not torch.is_tensor(R) else R)
    S = torch.as_tensor(np.asarray(S) if not torch.is_tensor(S) else S)
    single = R.dim() == 2
    if single:
        R, S = R[None], S[None]
    R = R.to(device=dev, dtype=torch.float64).contiguous()
    S = S.to(device=dev, dtype=torch.int32).contiguous()
    return R, S, single


class LogPsi:
    """logpsi(R, S) bound to a model and a parameter vector (psiformer.py:192-196)."""

    def __init__(self, model, theta):
        self.model = model
        self.theta = theta.to(device=_lib.device(), dtype=torch.float64).contiguous()
        self.packed = model.pack(self.theta)

    def update(self, theta):
        """New parameters in the same device buffers (keeps addresses bound by a captured graph)."""
        self.theta.copy_(theta)
        self.model.pack(self.theta, out=self.packed)

    def __call__(self, R, S):
        lib = _lib.load()
        R, S, single = as_walkers(R, S)
        W = R.shape[0]
        d = self.model.desc()
        out = torch.empty(W, 2, dtype=torch.float64, device=R.device)
        nb = lib.nqs_workspace_bytes(ctypes.byref(d), _lib.OP_LOGPSI, W)
        ws = _lib.workspace(nb)
        _lib.check(lib.nqs_psiformer_fwd(ctypes.byref(d), _lib.ptr(self.packed), _lib.ptr(R), _lib.ptr(S), W,
                                         _lib.ptr(out), _lib.ptr(ws), ws.numel(), _lib.stream()),
                   "nqs_psiformer_fwd")
        v = torch.view_as_complex(out)
        return v[0] if single else v


def make_logpsi(model, params):
    """params: the nested dict from model.init (or a flat theta vector)."""
    theta = params if torch.is_tensor(params) and params.dim() == 1 else ravel_params(params)
    return LogPsi(model, theta)
