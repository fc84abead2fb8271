"""Oracle Metropolis sampler (test infrastructure): src/nqs/sampler.py:19-98 restated with
CALLER-SUPPLIED random streams instead of jax.random keys, so accept/reject decisions can be
compared decision by decision.

Stream layout for n sweeps over W walkers with n_flip spin flips per sweep:
  xi     (n, W, N, 2) float64  standard normals of the joint position move
  u_pos  (n, W)       float64  uniforms for the position accept test
  flip_i (n, n_flip, W) int64  electron index of each spin-flip proposal
  u_flip (n, n_flip, W) float64 uniforms for the spin accept tests

One sweep (sampler.py:36-57):
  R' = wrap(R + step * xi),  wrap(r) = mod(r A^-T, 1) A^T
  accept iff u < exp(min(0, 2 Re L(R',S) - 2 Re L(R,S)))
  then n_flip times: S' = S with s_i -> 1 - s_i, same test on 2 Re L(R,S') - 2 Re L(R,S).
"""
import numpy as np


def wrap(R, A):
    """jnp.mod(r @ Ainv.T, 1.0) @ A.T with numpy's floating mod (sign of the divisor)."""
    A = np.asarray(A, float)
    Ainv = np.linalg.inv(A)
    x, y = R[..., 0], R[..., 1]
    sx = np.mod(x * Ainv[0, 0] + y * Ainv[0, 1], 1.0)      # elementwise: one rounding per op
    sy = np.mod(x * Ainv[1, 0] + y * Ainv[1, 1], 1.0)
    return np.stack([sx * A[0, 0] + sy * A[0, 1], sx * A[1, 0] + sy * A[1, 1]], -1)


def sweep(logpsi_b, R, S, A, step, xi, u_pos, flip_i, u_flip, lp_cur=None):
    """One sweep for all walkers.  logpsi_b(R (W,N,2), S (W,N)) -> complex (W,).
    Returns R, S, acc_pos (W,) bool, acc_spin (n_flip, W) bool, log|psi|^2 of the final state."""
    R = np.array(R, float)
    S = np.array(S, np.int64)
    logp = lambda R_, S_: 2.0 * np.real(np.asarray(logpsi_b(R_, S_)))
    lp = logp(R, S) if lp_cur is None else lp_cur
    Rp = wrap(R + step * xi, A)
    lpp = logp(Rp, S)
    acc = u_pos < np.exp(np.minimum(0.0, lpp - lp))
    R = np.where(acc[:, None, None], Rp, R)
    lp = np.where(acc, lpp, lp)
    accs = []
    for f in range(flip_i.shape[0]):
        Sp = S.copy()
        w = np.arange(S.shape[0])
        Sp[w, flip_i[f]] = 1 - S[w, flip_i[f]]
        lpp = logp(R, Sp)
        a = u_flip[f] < np.exp(np.minimum(0.0, lpp - lp))
        S = np.where(a[:, None], Sp, S)
        lp = np.where(a, lpp, lp)
        accs.append(a)
    return R, S, acc, np.array(accs), lp


def make_streams(n, W, N, n_flip, seed):
    """Seeded numpy streams in the layout above (the same arrays feed oracle and kernel)."""
    rng = np.random.default_rng(seed)
    return dict(xi=rng.standard_normal((n, W, N, 2)), u_pos=rng.random((n, W)),
                flip_i=rng.integers(0, N, (n, n_flip, W)), u_flip=rng.random((n, n_flip, W)))
