"""Oracle Psiformer log psi (test infrastructure), float64 / complex128 torch on the CPU.

Restates src/nqs/psiformer.py:37-196 for ONE configuration (R (N,2), S (N,)); batches go
through torch.func.vmap.  The parameter tree mirrors the flax module tree the reference
builds (psiformer.py:50-81, 118-188) and `real_ravel` (src/nqs/minsr.py:30-53) flattens it in
jax's key-sorted order with complex leaves split into [re..., im...].

Library definitions restated (flax 0.8.5 / jax 0.4.31 defaults the reference relies on):
  Dense      y = x @ kernel + bias
  LayerNorm  eps = 1e-6, mean = E[x], var = max(0, E[x^2] - E[x]^2),
             y = (x - mean) * (rsqrt(var + eps) * scale) + bias
  gelu       tanh form 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))
  softmax    exp(s - max s) / sum
  slogdet    LU with partial pivoting; log psi = log|det| + log(sign)
  logsumexp  complex: log(sum exp(a - max Re a)) + max Re a  (principal log)
"""
import math
import numpy as np
import torch

LN_EPS = 1e-6
GELU_K0 = math.sqrt(2.0 / math.pi)
GELU_K1 = 0.044715


# ---------------------------------------------------------------- architecture / parameters
class Arch:
    """Static architecture + physics inputs (psiformer.py:100-116)."""

    def __init__(self, n_elec, Gs, W, L, embed_dim=32, depth=3, n_heads=4, ff_mult=2, n_freq=4,
                 n_dets=1, learnable_env=False, use_jastrow=False, jastrow_kind="rbf", n_knots=10,
                 cusp=1.0):
        self.N = int(n_elec)
        self.Gs = np.asarray(Gs, float)
        self.W = np.asarray(W, complex)
        self.A = (np.array([[float(L), 0.0], [0.0, float(L)]]) if np.ndim(L) == 0
                  else np.asarray(L, float))
        self.d, self.depth, self.h, self.ff = int(embed_dim), int(depth), int(n_heads), int(ff_mult)
        self.F, self.K = int(n_freq), int(n_dets)
        self.learnable_env, self.use_jastrow = bool(learnable_env), bool(use_jastrow)
        self.jastrow_kind, self.n_knots, self.cusp = jastrow_kind, int(n_knots), float(cusp)


def param_shapes(a):
    """Nested dict of (shape, is_complex) in the flax tree layout (psiformer.py:50-188)."""
    d, N, K = a.d, a.N, a.K
    dense = lambda i, o: {"bias": ((o,), False), "kernel": ((i, o), False)}
    ln = {"bias": ((d,), False), "scale": ((d,), False)}
    t = {"LayerNorm_0": dict(ln), "embed": dense(4 * a.F + 2, d),
         "backflow": {"bias": ((K * N,), True), "kernel": ((d, K * N), True)}}
    for l in range(a.depth):
        t[f"TransformerBlock_{l}"] = {
            "Dense_0": dense(d, a.ff * d), "Dense_1": dense(a.ff * d, d),
            "LayerNorm_0": dict(ln), "LayerNorm_1": dict(ln),
            "attn": {"q": dense(d, d), "k": dense(d, d), "v": dense(d, d), "o": dense(d, d)}}
    if a.learnable_env:
        t["dW"] = (a.W.shape, True)
    if K > 1:
        t["det_logw"] = ((K,), True)
    if a.use_jastrow:
        if a.jastrow_kind == "rbf":
            t["jas_c_same"] = ((a.n_knots,), False)
            t["jas_c_opp"] = ((a.n_knots,), False)
        else:
            t["jastrow_loga"] = ((), False)
    return {"params": t}


def _leaves(tree, prefix=()):
    """Flatten a nested dict in jax's sorted-key order (jax.tree_util on dicts sorts keys)."""
    if isinstance(tree, dict):
        out = []
        for k in sorted(tree.keys()):
            out += _leaves(tree[k], prefix + (k,))
        return out
    return [(prefix, tree)]


def ravel_layout(a):
    """[(path, shape, is_complex, offset)] in real_ravel order (minsr.py:30-53) and total P."""
    out, off = [], 0
    for path, (shape, cplx) in _leaves(param_shapes(a)):
        size = int(np.prod(shape)) if len(shape) else 1
        out.append(("/".join(path), tuple(shape), cplx, off))
        off += size * (2 if cplx else 1)
    return out, off


def unravel(a, theta):
    """theta (P,) float64 -> dict path -> tensor (complex leaves recombined)."""
    lay, P = ravel_layout(a)
    assert theta.shape[-1] == P
    p = {}
    for path, shape, cplx, off in lay:
        n = int(np.prod(shape)) if len(shape) else 1
        if cplx:
            p[path] = torch.complex(theta[off:off + n], theta[off + n:off + 2 * n]).reshape(shape)
        else:
            p[path] = theta[off:off + n].reshape(shape)
    return p


def init_theta(a, seed=0, backflow_scale=0.0, scale_all=0.0):
    """Deterministic synthetic initial parameters (numpy Generator): Dense kernels
    ~ N(0, 1/fan_in) (lecun-normal scale), biases 0, LayerNorm scale 1 / bias 0, backflow and
    the optional leaves 0 (the reference zero-inits them, psiformer.py:124,139-141,157,180-181).
    backflow_scale > 0 fills the backflow kernel with (1+i) N(0, s^2); scale_all adds
    N(0, s^2) to every leaf (a perturbed, non-trivial network)."""
    rng = np.random.default_rng(seed)
    lay, P = ravel_layout(a)
    th = np.zeros(P)
    for path, shape, cplx, off in lay:
        n = int(np.prod(shape)) if len(shape) else 1
        leaf = path.split("/")[-1]
        if leaf == "kernel" and "backflow" not in path:
            th[off:off + n] = rng.standard_normal(n) / math.sqrt(shape[0])
        elif leaf == "scale":
            th[off:off + n] = 1.0
        elif "backflow/kernel" in path and backflow_scale > 0:
            z = rng.standard_normal(n) * backflow_scale
            th[off:off + n] = z
            th[off + n:off + 2 * n] = z
    if scale_all > 0:
        th = th + scale_all * rng.standard_normal(P)
    return torch.tensor(th, dtype=torch.float64)


# ---------------------------------------------------------------- the forward pass
def dense(x, p, name):
    return x @ p[name + "/kernel"] + p[name + "/bias"]


def layer_norm(x, p, name):
    mean = x.mean(-1, keepdim=True)
    mean2 = (x * x).mean(-1, keepdim=True)
    var = torch.clamp(mean2 - mean * mean, min=0.0)
    mul = torch.rsqrt(var + LN_EPS) * p[name + "/scale"]
    return (x - mean) * mul + p[name + "/bias"]


def gelu_tanh(x):
    return 0.5 * x * (1.0 + torch.tanh(GELU_K0 * (x + GELU_K1 * x ** 3)))


def mha(x, p, name, n_heads):
    """psiformer.py:50-64: contiguous head slices, softmax(QK^T / sqrt(hd)) V, then o-Dense."""
    N, d = x.shape
    hd = d // n_heads
    q = dense(x, p, name + "/q").reshape(N, n_heads, hd).transpose(0, 1)
    k = dense(x, p, name + "/k").reshape(N, n_heads, hd).transpose(0, 1)
    v = dense(x, p, name + "/v").reshape(N, n_heads, hd).transpose(0, 1)
    s = (q @ k.transpose(1, 2)) * (1.0 / math.sqrt(hd))
    s = s - s.max(-1, keepdim=True).values.detach()
    e = torch.exp(s)
    att = e / e.sum(-1, keepdim=True)
    out = (att @ v).transpose(0, 1).reshape(N, d)
    return dense(out, p, name + "/o")


def backbone(a, p, R, S):
    """Features -> embed -> pre-norm blocks -> final LayerNorm (psiformer.py:84-136)."""
    Ainv = torch.tensor(np.linalg.inv(a.A), dtype=torch.float64)
    s = R @ Ainv.T
    ph = 2.0 * math.pi * s
    m = torch.arange(1, a.F + 1, dtype=torch.float64)
    ang = ph[:, :, None] * m
    feats = torch.cat([torch.sin(ang), torch.cos(ang)], -1).reshape(R.shape[0], 4 * a.F)
    spin = torch.nn.functional.one_hot(S.long(), 2).to(torch.float64)
    x = dense(torch.cat([feats, spin], -1), p, "params/embed")
    for l in range(a.depth):
        b = f"params/TransformerBlock_{l}"
        x = x + mha(layer_norm(x, p, b + "/LayerNorm_0"), p, b + "/attn", a.h)
        h = layer_norm(x, p, b + "/LayerNorm_1")
        h = dense(gelu_tanh(dense(h, p, b + "/Dense_0")), p, b + "/Dense_1")
        x = x + h
    return layer_norm(x, p, "params/LayerNorm_0")


def logpsi_from_params(a, p, R, S):
    """log psi (complex128 scalar) of one configuration (psiformer.py:118-189)."""
    N, K = a.N, a.K
    x = backbone(a, p, R, S)
    bf = x.to(torch.complex128) @ p["params/backflow/kernel"] + p["params/backflow/bias"]
    bf = bf.reshape(N, K, N).permute(1, 2, 0)                        # (K, n_orb, N)
    Gs = torch.tensor(a.Gs, dtype=torch.float64)
    W = torch.tensor(a.W, dtype=torch.complex128)
    if a.learnable_env:
        W = W + p["params/dW"]
    phase = torch.exp(1j * (Gs @ R.T).to(torch.complex128))         # (n_G, N)
    Ws = W[:, :, S.long()]                                          # (n_orb, n_G, N)
    env = torch.einsum("gi,agi->ai", phase, Ws)
    M = env[None] * (1.0 + bf)
    sign, logabs = torch.linalg.slogdet(M)
    logdets = logabs.to(torch.complex128) + torch.log(sign)
    if K == 1:
        lp = logdets[0]
    else:
        z = logdets + p["params/det_logw"]
        amax = z.real.max().detach()
        lp = torch.log(torch.exp(z - amax).sum()) + amax
    if a.use_jastrow:
        lp = lp + jastrow(a, p, R, S).to(torch.complex128)
    return lp


def jastrow(a, p, R, S):
    """psiformer.py:162-188: minimum-image pair distance (round), RBF or cusp pair term, i<j.
    jastrow_kind "cusp_periodic" (no reference counterpart): the cusp form on the smooth periodic
    chord distance d^2 = sum_i (L_i/pi)^2 sin^2(pi x_i/L_i) of a rectangular cell."""
    A = torch.tensor(a.A, dtype=torch.float64)
    Ainv = torch.tensor(np.linalg.inv(a.A), dtype=torch.float64)
    d = R[:, None, :] - R[None, :, :]
    if a.jastrow_kind == "cusp_periodic":
        Ls = torch.tensor([a.A[0, 0], a.A[1, 1]], dtype=torch.float64)
        D = ((Ls / math.pi) ** 2 * torch.sin(math.pi * d / Ls) ** 2).sum(-1)
        r = torch.sqrt(D + 1e-12)
    else:
        dfrac = d @ Ainv.T
        dmin = (dfrac - torch.round(dfrac).detach()) @ A.T
        r = torch.sqrt((dmin * dmin).sum(-1) + 1e-12)
    same = S[:, None] == S[None, :]
    if a.jastrow_kind == "rbf":
        area = abs(float(np.linalg.det(a.A)))
        rcut = 0.5 * math.sqrt(area)
        mu = torch.tensor(np.linspace(0.0, rcut, a.n_knots), dtype=torch.float64)
        w = rcut / max(a.n_knots - 1, 1)
        rbf = torch.exp(-((r[..., None] - mu) / w) ** 2)
        u = torch.where(same, (rbf * p["params/jas_c_same"]).sum(-1), (rbf * p["params/jas_c_opp"]).sum(-1))
    else:
        aa = torch.nn.functional.softplus(p["params/jastrow_loga"]) + 0.05
        alpha = torch.where(same, torch.full_like(r, a.cusp / 3.0), torch.full_like(r, a.cusp))
        u = alpha * aa * (1.0 - torch.exp(-r / aa))
    iu = torch.triu_indices(a.N, a.N, 1)
    return u[iu[0], iu[1]].sum()


def logpsi(a, theta, R, S):
    return logpsi_from_params(a, unravel(a, theta), R, S)


def make_logpsi(a, theta):
    """psiformer.py:192-196 -- bind params -> logpsi(R, S) for one configuration."""
    p = unravel(a, theta)
    return lambda R, S: logpsi_from_params(a, p, R, S)


def batched_logpsi(a, theta, R, S):
    """(W,N,2), (W,N) -> (W,) complex128."""
    p = unravel(a, theta)
    return torch.stack([logpsi_from_params(a, p, R[w], S[w]) for w in range(R.shape[0])])


def slater_logpsi(Gs, W, R, S):
    """src/nqs/slater_reference.py:21-35 -- exact multi-G spinor Slater determinant."""
    Gs = torch.as_tensor(Gs, dtype=torch.float64)
    W = torch.as_tensor(W, dtype=torch.complex128)
    phase = torch.exp(1j * (Gs @ R.T).to(torch.complex128))
    env = torch.einsum("gi,agi->ai", phase, W[:, :, S.long()])
    sign, logabs = torch.linalg.slogdet(env)
    return logabs.to(torch.complex128) + torch.log(sign)
