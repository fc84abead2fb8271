// Per-walker physics kernels: Ewald Coulomb energy, local-energy assembly (kinetic + Rashba SOC
// + Zeeman + external field) and the Metropolis propose/accept steps.  All float64.
#include "nqs_common.h"
#include "kernels.h"
#include "physics.h"

namespace nqs {

// Neumaier compensated accumulator (the Ewald energy is a sum of O(N^2 (n_img + n_G)) terms of
// both signs).
struct Comp {
  double s = 0, c = 0;
  __device__ inline void add(double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x))
      c += (s - t) + x;
    else
      c += (x - t) + s;
    s = t;
  }
  __device__ inline double val() const { return s + c; }
};

// ------------------------------------------------------------------ Ewald (ewald.py:47-65)
// One workgroup per walker.  Real space: every pair i<j, every image A n of the square index
// box, erfc(eta d)/d with the raw difference r_i - r_j (no minimum image), as the reference.
// Reciprocal space: sum_{i<j} cos(G.(r_i - r_j)) = (|rho_G|^2 - N)/2 with rho_G = sum_i e^{iG.r_i}.
__global__ __launch_bounds__(256) void k_ewald(nqs_ewald_desc e, const double* __restrict__ R, int N,
                                               double* __restrict__ out) {
  __shared__ double pos[2 * 64];
  __shared__ double red[2 * 8];
  const int64_t w = blockIdx.x;
  const int tid = threadIdx.x;
  for (int i = tid; i < 2 * N; i += blockDim.x) pos[i] = R[w * 2 * N + i];
  __syncthreads();
  Comp acc;
  const int npair = N * (N - 1) / 2;
  const int64_t nreal = (int64_t)npair * e.n_img;
  for (int64_t q = tid; q < nreal; q += blockDim.x) {
    const int pr = (int)(q / e.n_img), im = (int)(q % e.n_img);
    // pair index -> (i, j), i < j, row-major over the upper triangle
    int i = 0, rem = pr;
    while (rem >= N - 1 - i) {
      rem -= N - 1 - i;
      ++i;
    }
    const int j = i + 1 + rem;
    const double dx = pos[2 * i] - pos[2 * j] + e.img_dev[2 * im];
    const double dy = pos[2 * i + 1] - pos[2 * j + 1] + e.img_dev[2 * im + 1];
    const double d = sqrt(dx * dx + dy * dy);
    acc.add(erfc(e.eta * d) / d);
  }
  for (int g = tid; g < e.n_g; g += blockDim.x) {
    const double gx = e.g_dev[2 * g], gy = e.g_dev[2 * g + 1];
    double cr = 0, ci = 0;
    for (int i = 0; i < N; ++i) {
      double sn, cs;
      sincos(gx * pos[2 * i] + gy * pos[2 * i + 1], &sn, &cs);
      cr += cs;
      ci += sn;
    }
    acc.add(e.rg_dev[g] * 0.5 * ((cr * cr + ci * ci) - (double)N));
  }
  // workgroup reduction of the compensated pair (s, c)
  double s = acc.s, c = acc.c;
  for (int o = 32; o > 0; o >>= 1) {
    const double s2 = __shfl_xor(s, o, 64), c2 = __shfl_xor(c, o, 64);
    const double t = s + s2;
    c += c2 + ((fabs(s) >= fabs(s2)) ? (s - t) + s2 : (s2 - t) + s);
    s = t;
  }
  if ((tid & 63) == 0) {
    red[2 * (tid >> 6)] = s;
    red[2 * (tid >> 6) + 1] = c;
  }
  __syncthreads();
  if (tid == 0) {
    Comp tot;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
      tot.add(red[2 * k]);
      tot.c += red[2 * k + 1];
    }
    tot.add((double)npair * e.v_g0);
    out[w] = e.alpha * tot.val() + (double)N * e.zeta_m / 2.0;
  }
}

int launch_ewald(const nqs_ewald_desc& e, const double* R, int64_t W, int N, double* out, hipStream_t st) {
  if (W == 0) return NQS_OK;
  NQS_REQUIRE(N >= 1 && N <= 64, NQS_ERR_UNSUPPORTED, "Ewald kernel supports 1 <= N <= 64");
  hipLaunchKernelGGL(k_ewald, dim3((unsigned)W), dim3(256), 0, st, e, R, N, out);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// ------------------------------------------------------------------ local energy assembly (local_energy.py:63-115)
// m(r) = sum_t c_t cos(q_t.r) + s_t sin(q_t.r); f = m, or scale m / (|m| + 1e-12) for a
// normalised texture (pinning.py:29-37, 96-100)
__device__ inline void ext_field_at(const double* fld, int nf, double scale, double x, double y, double f[3]) {
  f[0] = f[1] = f[2] = 0;
  for (int t = 0; t < nf; ++t) {
    const double* q = fld + 8 * t;
    double sn, cs;
    sincos(q[0] * x + q[1] * y, &sn, &cs);
    for (int a = 0; a < 3; ++a) f[a] += q[2 + a] * cs + q[5 + a] * sn;
  }
  if (scale > 0) {
    const double k = scale / (sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]) + 1e-12);
    for (int a = 0; a < 3; ++a) f[a] *= k;
  }
}

// Gaussian pinning centres at (x, y): r - c reduced to its minimum image (fractional
// s = A^-1 (r - c), s -= rint(s)), then the 3 x 3 images around it, so the profile g is periodic
// for any positions; adds v0 g to *v and h g to f.
__device__ inline void pins_at(const double* pin, int np, const double A[4], double x, double y, double* v,
                               double f[3]) {
  const double idet = 1.0 / (A[0] * A[3] - A[1] * A[2]);
  for (int p = 0; p < np; ++p) {
    const double* c = pin + 8 * p;
    const double inv2w2 = 1.0 / (2.0 * c[2] * c[2]);
    const double ux = x - c[0], uy = y - c[1];
    const double sx = rint((A[3] * ux - A[1] * uy) * idet), sy = rint((A[0] * uy - A[2] * ux) * idet);
    const double rx = ux - (A[0] * sx + A[1] * sy), ry = uy - (A[2] * sx + A[3] * sy);
    double g = 0;
    for (int nx = -1; nx <= 1; ++nx)
      for (int ny = -1; ny <= 1; ++ny) {
        const double dx = rx - (A[0] * nx + A[1] * ny), dy = ry - (A[2] * nx + A[3] * ny);
        g += exp(-(dx * dx + dy * dy) * inv2w2);
      }
    *v += c[3] * g;
    for (int a = 0; a < 3; ++a) f[a] += c[4 + a] * g;
  }
}

__global__ void k_eloc_assemble(EAssembleArgs a) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= a.W) return;
  const int N = a.N;
  const cd L0 = mk(a.lp[2 * w], a.lp[2 * w + 1]);
  // kinetic: -1/(2m*) (Lap L + sum_c (d_c L)^2)
  cd g2 = mk(0, 0);
  for (int c = 0; c < 2 * N; ++c) {
    const cd g = mk(a.grad[(w * 2 * N + c) * 2], a.grad[(w * 2 * N + c) * 2 + 1]);
    g2 = g2 + g * g;
  }
  const cd lap = mk(a.lap[2 * w], a.lap[2 * w + 1]);
  const cd ekin = (-1.0 / (2.0 * a.m_star)) * (lap + g2);
  // Zeeman
  double sgs = 0;
  for (int i = 0; i < N; ++i) sgs += 1.0 - 2.0 * (double)a.S[w * N + i];
  double ez = -0.5 * a.h_z * sgs;
  // per electron: the field f(r_i) (ext field + pinning spin fields) and the pinning potential;
  // f_z enters the diagonal, f_x, f_y the flipped configurations (cfg = w N + i) with the SOC
  const bool fields = a.n_field > 0 || a.n_pin > 0;
  double epin = 0;
  cd esoc = mk(0, 0);
  for (int i = 0; i < N; ++i) {
    const int64_t cfg = w * N + i;
    const int s = a.S[w * N + i];
    double f[3] = {0, 0, 0};
    if (fields) {
      const double x = a.R[(w * N + i) * 2], y = a.R[(w * N + i) * 2 + 1];
      if (a.n_field > 0) ext_field_at(a.field, a.n_field, a.field_scale, x, y, f);
      if (a.n_pin > 0) pins_at(a.pin, a.n_pin, a.A, x, y, &epin, f);
      ez -= f[2] * (1.0 - 2.0 * (double)s);
    }
    if (!a.lpf) continue;   // no flipped configurations: no Rashba term, no in-plane fields
    const cd Lf = mk(a.lpf[2 * cfg], a.lpf[2 * cfg + 1]);
    const cd ratio = cexp_(Lf - L0);
    cd ax = ratio * mk(a.tanf[(cfg * 2 + 0) * 2], a.tanf[(cfg * 2 + 0) * 2 + 1]);
    cd ay = ratio * mk(a.tanf[(cfg * 2 + 1) * 2], a.tanf[(cfg * 2 + 1) * 2 + 1]);
    if (!(isfinite(ax.re) && isfinite(ax.im))) ax = mk(0, 0);
    if (!(isfinite(ay.re) && isfinite(ay.im))) ay = mk(0, 0);
    const cd sy = (s == 0) ? mk(0, -1) : mk(0, 1);   // (sigma_y)[s, 1-s]
    const cd mi = mk(0, -1);
    cd e = a.lambda_r * (mi * ay + sy * (mi * ax));
    if (a.n_field > 0 || f[0] != 0 || f[1] != 0) e = e - (mk(f[0], 0) + f[1] * sy) * ratio;
    esoc = esoc + e;
  }
  const cd tot = ekin + mk(ez + epin, 0) + esoc;
  a.out[2 * w] = tot.re;
  a.out[2 * w + 1] = tot.im;
}

int launch_eloc_assemble(const EAssembleArgs& a, hipStream_t st) {
  if (a.W == 0) return NQS_OK;
  hipLaunchKernelGGL(k_eloc_assemble, dim3((unsigned)cdiv(a.W, 128)), dim3(128), 0, st, a);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// ------------------------------------------------------------------ Metropolis (sampler.py:36-57)
// R' = wrap(R + step xi), wrap(r) = mod(r A^-T, 1) A^T.  Rounded operation by operation, no
// contraction, so that positions agree bit for bit with a host evaluation of the same formula.
#pragma clang fp contract(off)
__device__ inline double fmod_pos(double x) {
  double r = fmod(x, 1.0);
  if (r != 0.0 && r < 0.0) r += 1.0;
  return r;
}

__global__ void k_mcmc_propose(int64_t W, int N, const double* __restrict__ R, const double* __restrict__ xi,
                               double step, double a00, double a01, double a10, double a11, double i00,
                               double i01, double i10, double i11, double* __restrict__ Rp) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= W * N) return;
  // plain operators: the contract(off) region keeps each product and sum rounded on its own
  const double x = R[2 * q] + step * xi[2 * q];
  const double y = R[2 * q + 1] + step * xi[2 * q + 1];
  const double sx = fmod_pos(x * i00 + y * i01);
  const double sy = fmod_pos(x * i10 + y * i11);
  Rp[2 * q] = sx * a00 + sy * a01;
  Rp[2 * q + 1] = sx * a10 + sy * a11;
}

// Position accept: d = 2 Re L' - logp; accept iff u < exp(min(0, d)).
__global__ void k_mcmc_accept_pos(int64_t W, int N, double* __restrict__ R, const double* __restrict__ Rp,
                                  double* __restrict__ logp, const double* __restrict__ lp_new,
                                  const double* __restrict__ u, int32_t* __restrict__ acc) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  const double lpn = 2.0 * lp_new[2 * w];
  const double d = lpn - logp[w];
  // jnp.minimum propagates NaN (a NaN proposal is rejected); d > 0 clamps to 0
  if (u[w] < exp(d > 0.0 ? 0.0 : d)) {
    for (int i = 0; i < 2 * N; ++i) R[w * 2 * N + i] = Rp[w * 2 * N + i];
    logp[w] = lpn;
    if (acc) acc[w] += 1;
  }
}

// Spin flip proposal: S' = S with s_i flipped (stored in Sp); accept as above.  An index outside
// [0, N) proposes nothing (Sp = S) and the accept kernel rejects it, so a caller's bad stream can
// neither write outside the walker's spins nor count as a move.
__global__ void k_mcmc_flip_propose(int64_t W, int N, const int32_t* __restrict__ S,
                                    const int32_t* __restrict__ fi, int32_t* __restrict__ Sp) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= W * N) return;
  const int64_t w = q / N;
  const int i = (int)(q % N);
  const int s = S[q];
  Sp[q] = (i == fi[w]) ? 1 - s : s;   // never true for fi[w] outside [0, N)
}

__global__ void k_mcmc_accept_spin(int64_t W, int N, int32_t* __restrict__ S, const int32_t* __restrict__ fi,
                                   double* __restrict__ logp, const double* __restrict__ lp_new,
                                   const double* __restrict__ u, int32_t* __restrict__ acc) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  const double lpn = 2.0 * lp_new[2 * w];
  const double d = lpn - logp[w];
  const int i = fi[w];
  if ((unsigned)i >= (unsigned)N) return;   // out-of-range proposal: rejected
  if (u[w] < exp(d > 0.0 ? 0.0 : d)) {
    S[w * N + i] = 1 - S[w * N + i];
    logp[w] = lpn;
    if (acc) acc[w] += 1;
  }
}
#pragma clang fp contract(on)

int launch_mcmc_propose(int64_t W, int N, const double* R, const double* xi, double step, const double* A,
                        const double* Ainv, double* Rp, hipStream_t st) {
  const int64_t n = W * N;
  hipLaunchKernelGGL(k_mcmc_propose, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, st, W, N, R, xi, step, A[0],
                     A[1], A[2], A[3], Ainv[0], Ainv[1], Ainv[2], Ainv[3], Rp);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}
int launch_mcmc_accept_pos(int64_t W, int N, double* R, const double* Rp, double* logp, const double* lp_new,
                           const double* u, int32_t* acc, hipStream_t st) {
  hipLaunchKernelGGL(k_mcmc_accept_pos, dim3((unsigned)cdiv(W, 256)), dim3(256), 0, st, W, N, R, Rp, logp,
                     lp_new, u, acc);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}
int launch_mcmc_flip_propose(int64_t W, int N, const int32_t* S, const int32_t* fi, int32_t* Sp, hipStream_t st) {
  hipLaunchKernelGGL(k_mcmc_flip_propose, dim3((unsigned)cdiv(W * N, 256)), dim3(256), 0, st, W, N, S, fi, Sp);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}
int launch_mcmc_accept_spin(int64_t W, int N, int32_t* S, const int32_t* fi, double* logp, const double* lp_new,
                            const double* u, int32_t* acc, hipStream_t st) {
  hipLaunchKernelGGL(k_mcmc_accept_spin, dim3((unsigned)cdiv(W, 256)), dim3(256), 0, st, W, N, S, fi, logp,
                     lp_new, u, acc);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

}  // namespace nqs
