// Spin observables and ratio estimators (src/nqs/observables.py:31-128,
// src/nqs/manybody_chern.py:23-34, src/nqs/topology.py:114-122).
//
//   spin estimators   per electron i: ratio_i = exp(L(S with s_i flipped) - L(S)) (non-finite -> 0),
//                     sigma_x = ratio_i, sigma_y = -i (1 - 2 s_i) ratio_i, sigma_z = 1 - 2 s_i;
//   spin Fourier sums rho_a(q) = sum_i sigma_a^i exp(-i q.r_i) per walker, reduced over walkers to
//                     m_a(q) = <rho_a>/N, S_zz = <|rho_z|^2>/N, S_perp = <|rho_x|^2 + |rho_y|^2>/N;
//   spin texture      sums of sigma_a and counts per cell bin (fractional coordinates mod 1);
//   ratio statistics  r_w = phase_w exp(lb_w - la_w - shift): mean, mean |r|^2, effective sample
//                     fraction, standard deviation.
// All reductions run in a fixed order (strided per-thread sums, then an LDS tree), so results do
// not depend on scheduling.
#include "nqs_common.h"
#include "obs.h"

namespace nqs {

__global__ void k_spin_est(int64_t W, int N, const int32_t* __restrict__ S, const double* __restrict__ lp,
                           const double* __restrict__ lpf, double* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= W * N) return;
  const int64_t w = q / N;
  const cd d = mk(lpf[2 * q] - lp[2 * w], lpf[2 * q + 1] - lp[2 * w + 1]);
  cd r = cexp_(d);
  if (!(isfinite(r.re) && isfinite(r.im))) r = mk(0, 0);
  const double sz = 1.0 - 2.0 * (double)S[q];
  double* o = out + q * 6;
  o[0] = r.re;            // sigma_x
  o[1] = r.im;
  o[2] = sz * r.im;       // sigma_y = -i sz ratio
  o[3] = -sz * r.re;
  o[4] = sz;              // sigma_z
  o[5] = 0.0;
}

int launch_spin_est(int64_t W, int N, const int32_t* S, const double* lp, const double* lpf, double* out,
                    hipStream_t st) {
  if (W * N == 0) return NQS_OK;
  hipLaunchKernelGGL(k_spin_est, dim3((unsigned)cdiv(W * N, 256)), dim3(256), 0, st, W, N, S, lp, lpf, out);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

template <int NV>
__device__ inline void block_sum(double (&v)[NV], double* sh) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int k = 0; k < NV; ++k) sh[k * nt + tid] = v[k];
  __syncthreads();
  for (int s = nt / 2; s > 0; s >>= 1) {
    if (tid < s)
      for (int k = 0; k < NV; ++k) sh[k * nt + tid] += sh[k * nt + tid + s];
    __syncthreads();
  }
  for (int k = 0; k < NV; ++k) v[k] = sh[k * nt];
  __syncthreads();
}

// one block per wave vector q; threads stride over walkers
__global__ __launch_bounds__(256) void k_spin_fourier(int64_t W, int N, const double* __restrict__ R,
                                                      const double* __restrict__ sp, const double* __restrict__ qv,
                                                      double* __restrict__ out) {
  __shared__ double sh[8 * 256];
  const int iq = blockIdx.x;
  const double qx = qv[2 * iq], qy = qv[2 * iq + 1];
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t w = threadIdx.x; w < W; w += blockDim.x) {
    double rr[3] = {0, 0, 0}, ri[3] = {0, 0, 0};
    for (int i = 0; i < N; ++i) {
      const int64_t e = w * N + i;
      double sn, cs;
      sincos(qx * R[2 * e] + qy * R[2 * e + 1], &sn, &cs);
      for (int a = 0; a < 3; ++a) {
        rr[a] += sp[3 * e + a] * cs;   // exp(-i q.r) = cos - i sin
        ri[a] -= sp[3 * e + a] * sn;
      }
    }
    for (int a = 0; a < 3; ++a) {
      acc[2 * a] += rr[a];
      acc[2 * a + 1] += ri[a];
    }
    acc[6] += rr[2] * rr[2] + ri[2] * ri[2];
    acc[7] += rr[0] * rr[0] + ri[0] * ri[0] + rr[1] * rr[1] + ri[1] * ri[1];
  }
  block_sum<8>(acc, sh);
  if (threadIdx.x == 0) {
    const double inv = 1.0 / ((double)W * (double)N);
    for (int k = 0; k < 8; ++k) out[8 * iq + k] = acc[k] * inv;
  }
}

int launch_spin_fourier(int64_t W, int N, const double* R, const double* sp, const double* q, int nq, double* out,
                        hipStream_t st) {
  if (nq == 0) return NQS_OK;
  hipLaunchKernelGGL(k_spin_fourier, dim3((unsigned)nq), dim3(256), 0, st, W, N, R, sp, q, out);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// x mod 1 with the sign rule of numpy's remainder (result in [0, 1))
__device__ inline double mod1(double x) {
  double m = fmod(x, 1.0);
  if (m != 0.0) {
    if (m < 0.0) m += 1.0;
  } else {
    m = 0.0;
  }
  return m;
}

// one block per bin; threads stride over all electrons of all walkers
__global__ __launch_bounds__(256) void k_spin_texture(int64_t n_el, const double* __restrict__ R,
                                                      const double* __restrict__ sp, double a00, double a01,
                                                      double a10, double a11, int ngrid, double* __restrict__ m,
                                                      double* __restrict__ rho) {
  // fractional coordinates rounded operation by operation, as numpy evaluates them
#pragma clang fp contract(off)
  __shared__ double sh[4 * 256];
  const int bin = blockIdx.x;
  double acc[4] = {0, 0, 0, 0};
  for (int64_t e = threadIdx.x; e < n_el; e += blockDim.x) {
    const double x = R[2 * e], y = R[2 * e + 1];
    const double fx = mod1(x * a00 + y * a01), fy = mod1(x * a10 + y * a11);
    int ix = (int)(fx * (double)ngrid), iy = (int)(fy * (double)ngrid);
    ix = min(max(ix, 0), ngrid - 1);
    iy = min(max(iy, 0), ngrid - 1);
    if (ix * ngrid + iy == bin) {
      for (int a = 0; a < 3; ++a) acc[a] += sp[3 * e + a];
      acc[3] += 1.0;
    }
  }
  block_sum<4>(acc, sh);
  if (threadIdx.x == 0) {
    for (int a = 0; a < 3; ++a) m[3 * bin + a] = acc[a];
    rho[bin] = acc[3];
  }
}

int launch_spin_texture(int64_t n_el, const double* R, const double* sp, const double ai[4], int ngrid, double* m,
                        double* rho, hipStream_t st) {
  hipLaunchKernelGGL(k_spin_texture, dim3((unsigned)(ngrid * ngrid)), dim3(256), 0, st, n_el, R, sp, ai[0], ai[1],
                     ai[2], ai[3], ngrid, m, rho);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// single block: shift = max_w Re(lb - la) when requested; r_w = phase_w exp(lb - la - shift);
// out = [mean Re r, mean Im r, mean |r|^2, ess, shift, std r]
__global__ __launch_bounds__(1024) void k_ratio_stats(int64_t W, const double* __restrict__ la,
                                                      const double* __restrict__ lb, const double* __restrict__ ph,
                                                      int shift, double* __restrict__ out) {
  __shared__ double sh[4 * 1024];
  const int tid = threadIdx.x, nt = blockDim.x;
  double mx = -HUGE_VAL;
  if (shift)
    for (int64_t w = tid; w < W; w += nt) mx = fmax(mx, lb[2 * w] - la[2 * w]);
  sh[tid] = mx;
  __syncthreads();
  for (int s = nt / 2; s > 0; s >>= 1) {
    if (tid < s) sh[tid] = fmax(sh[tid], sh[tid + s]);
    __syncthreads();
  }
  const double sft = shift ? sh[0] : 0.0;
  __syncthreads();
  auto ratio = [&](int64_t w) {
    cd r = cexp_(mk(lb[2 * w] - la[2 * w] - sft, lb[2 * w + 1] - la[2 * w + 1]));
    if (ph) r = mk(ph[2 * w], ph[2 * w + 1]) * r;
    return r;
  };
  double acc[4] = {0, 0, 0, 0};
  for (int64_t w = tid; w < W; w += nt) {
    const cd r = ratio(w);
    const double a2 = r.re * r.re + r.im * r.im;
    acc[0] += r.re;
    acc[1] += r.im;
    acc[2] += a2;
    acc[3] += sqrt(a2);
  }
  block_sum<4>(acc, sh);
  const double mre = acc[0] / (double)W, mim = acc[1] / (double)W;
  double dv[1] = {0};
  for (int64_t w = tid; w < W; w += nt) {
    const cd r = ratio(w);
    dv[0] += (r.re - mre) * (r.re - mre) + (r.im - mim) * (r.im - mim);
  }
  block_sum<1>(dv, sh);
  if (tid == 0) {
    out[0] = mre;
    out[1] = mim;
    out[2] = acc[2] / (double)W;
    out[3] = acc[3] * acc[3] / acc[2] / (double)W;
    out[4] = sft;
    out[5] = sqrt(dv[0] / (double)W);
  }
}

int launch_ratio_stats(int64_t W, const double* la, const double* lb, const double* ph, int shift, double* out,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_ratio_stats, dim3(1), dim3(1024), 0, st, W, la, lb, ph, shift, out);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

}  // namespace nqs

using namespace nqs;

extern "C" {

int nqs_spin_fourier(const double* R, const double* spins, int64_t W, int32_t N, const double* q, int32_t n_q,
                     double* out, void* stream) {
  NQS_REQUIRE(W > 0 && N > 0 && n_q >= 0, NQS_ERR_ARG, "need walkers, electrons and n_q >= 0");
  NQS_REQUIRE(R && spins && out && (n_q == 0 || q), NQS_ERR_ARG, "null pointer");
  return launch_spin_fourier(W, N, R, spins, q, n_q, out, (hipStream_t)stream);
}

int nqs_spin_texture(const double* R, const double* spins, int64_t W, int32_t N, const double cell_inv[4],
                     int32_t ngrid, double* m, double* rho, void* stream) {
  NQS_REQUIRE(W >= 0 && N > 0 && ngrid >= 1 && ngrid <= 1024, NQS_ERR_ARG, "bad sizes (ngrid in [1, 1024])");
  NQS_REQUIRE(R && spins && cell_inv && m && rho, NQS_ERR_ARG, "null pointer");
  return launch_spin_texture(W * N, R, spins, cell_inv, ngrid, m, rho, (hipStream_t)stream);
}

int nqs_ratio_stats(const double* la, const double* lb, const double* phase, int64_t W, int32_t shift, double* out,
                    void* stream) {
  NQS_REQUIRE(W > 0, NQS_ERR_ARG, "need at least one walker");
  NQS_REQUIRE(la && lb && out, NQS_ERR_ARG, "null pointer");
  return launch_ratio_stats(W, la, lb, phase, shift ? 1 : 0, out, (hipStream_t)stream);
}

}  // extern "C"
