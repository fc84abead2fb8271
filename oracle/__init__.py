"""CPU oracle for the VMC hot path -- TEST INFRASTRUCTURE ONLY.

This package is a float64 / complex128 restatement, in torch on the CPU, of the reference's
per-step variational Monte Carlo path (wenqin-chen/continuum-nqs):

  cell.py          src/nqs/cell.py:14-46            cell helpers
  bands.py         src/nqs/rashba_reference.py:36-160, src/nqs/spiral_reference.py:24-75
                                                     closed-form Rashba bands / exact ground states
  psiformer.py     src/nqs/psiformer.py:37-196      log psi apply, flax parameter tree, real ravel order
  local_energy.py  src/nqs/local_energy.py:40-117   dense-Hessian E_loc (kinetic + SOC + Zeeman + ext field)
  ewald.py         src/nqs/ewald.py:26-73           2D Ewald Coulomb on a general cell
  sampler.py       src/nqs/sampler.py:25-98         Metropolis sweep with caller-supplied random streams
  minsr.py         src/nqs/minsr.py:30-159          per-sample O, MinSR, parameter-space SR, one step
  ed.py            (no reference counterpart)       plane-wave exact diagonalisation, N=2

Only tests, the smoke check and the CPU-baseline leg of the benchmark may import it, and only as
the checker.  The product (the HIP library and its Python host layer) never calls into it.

How each function is pinned (beyond retyping the reference):
  * bands: E_gs against the analytic band formula and brute-force eigh of the 2x2 Bloch
    Hamiltonian; spiral basis against its own boundary-leakage certificate.
  * psiformer: zero backflow == Slater determinant (closed form); antisymmetry; permutation
    equivariance of the backbone; the whole backbone (embed, pre-norm blocks, final LayerNorm)
    equal to torch.nn.TransformerEncoderLayer(norm_first=True) blocks and torch.nn.LayerNorm with
    the same weights at d = 128 / 8 heads, d = 32 / 4 heads and depth 0; mha equal to
    torch.nn.functional.multi_head_attention_forward; the backflow reshape against finite
    differences of a matrix built independently (tests/test_oracle.py).
  * local_energy: zero-variance certificates E_loc == E_gs (Rashba, polarized, spiral), which
    fail for a wrong sign / index in kinetic, SOC or Zeeman.
  * ewald: eta independence, periodicity, evenness, V(r)*r -> alpha, alpha linearity,
    pair-sum + Madelung, translation invariance, general == square cell.
  * sampler: accept rule checked on hand-built cases; uniform density of a plane-wave det.
  * minsr: push-through identity (sample space == parameter space); O against central
    finite differences.
  * ed: the non-interacting limit equals the closed-form band energy; at lambda_R = h_z = 0 the
    antisymmetrised spinful ground state equals the independently built relative-coordinate
    problem; the spectrum is invariant under the sign convention of the Rashba term; zeta_M
    matches the Ewald restatement.
The backbone's values are pinned against torch's library transformer layers (same definitions
as flax 0.8.5 LayerNorm / Dense and jax.nn.gelu's tanh form), not against flax itself, which is
not installable here.  cpu_step.py composes these pieces into the reference's step; the
two-electron trajectory it writes (tests/golden/) is the GPU path's whole-step parity target.
"""
