// Backward kernels for the per-walker parameter gradient O[w, k] = d log psi_w / d theta_k
// (src/nqs/minsr.py:56-65: grad of Re and Im of log psi with respect to the real ravel theta).
//
// log psi is complex and the backbone real, so every cotangent of a backbone activation is
// complex: two real channels (d Re L, d Im L) laid out [cfg][channel][token][feature].  Linear
// layers pass cotangents through the NT GEMM (gemm.hip) against the untransposed weights; the
// kernels here cover the determinant head, LayerNorm, GELU, attention (recomputing the softmax
// from the stored Q, K rather than keeping it per walker) and the per-walker weight / bias
// gradients, which are written straight into the rows of the O matrix:
//   row w            = Re O_w,      row W_total + w = Im O_w,      column = real_ravel index.
#include <algorithm>
#include "nqs_common.h"
#include "kernels.h"
#include "bwd.h"

namespace nqs {

// ------------------------------------------------------------------ per-walker weight gradient
// O[c][w][seg(j)] (i, j) = sum_t X[w][t][i] * G[w][c][t][j]   (X: value activations, G: cotangents)
// Output column j belongs to segment s = j / seg_w with its own theta base; inside a segment the
// leaf is (in, seg_w) row-major.
template <typename T>
__global__ __launch_bounds__(256) void k_wgrad(WgradArgs<T> a, int ob) {
  extern __shared__ char smem_raw[];
  T* Xs = reinterpret_cast<T*>(smem_raw);  // N x IB
  T* Gs = Xs + a.N * 16;                   // N x nb (output columns j0 .. j0 + nb)
  const int IB = 16;
  const int64_t cfg = blockIdx.y >> 1;
  const int c = blockIdx.y & 1;
  const int i0 = blockIdx.x * IB;
  const int j0 = blockIdx.z * ob;
  const int nb = min(ob, a.out - j0);
  const int tid = threadIdx.x;
  const int N = a.N;
  const T* X = a.X + cfg * (int64_t)N * a.ldx;
  const T* G = a.G + ((cfg * 2 + c) * (int64_t)N) * a.ldg + j0;
  for (int e = tid; e < N * IB; e += blockDim.x) {
    const int t = e / IB, ii = e % IB;
    Xs[e] = (i0 + ii < a.in) ? X[(int64_t)t * a.ldx + i0 + ii] : (T)0;
  }
  for (int e = tid; e < N * nb; e += blockDim.x) {
    const int t = e / nb, j = e % nb;
    Gs[e] = G[(int64_t)t * a.ldg + j];
  }
  __syncthreads();
  const int64_t row = (int64_t)c * a.W_total + a.w0 + cfg;
  T* O = a.O + row * a.ldo;
  const int nrow = min(IB, a.in - i0);
  for (int e = tid; e < nrow * nb; e += blockDim.x) {
    const int ii = e / nb, j = e % nb;
    T s = 0;
    for (int t = 0; t < N; ++t) s += Xs[t * IB + ii] * Gs[t * nb + j];
    const int sg = (j0 + j) / a.seg_w, jj = (j0 + j) % a.seg_w;
    O[a.seg_base[sg] + (int64_t)(i0 + ii) * a.seg_w + jj] = s;
  }
}

// Column form, used when the walker has at most 32 tokens: one workgroup per (cfg, channel, 256
// output columns [, chunk of rows i]); each thread keeps the N cotangents of its column j in
// registers and sweeps the rows i, reading X^T (staged in LDS, tokens padded to a multiple of 4)
// as broadcasts.  O is written once, a wave writing 64 consecutive columns of one row i; X and G
// are read once per workgroup.
template <typename T>
__global__ __launch_bounds__(256) void k_wgrad_cols(WgradArgs<T> a, int ib) {
  extern __shared__ char smem_raw[];
  T* Xt = reinterpret_cast<T*>(smem_raw);  // [ib][NP]
  const int N = a.N, NP = (N + 3) & ~3;
  const int64_t cfg = blockIdx.y >> 1;
  const int c = blockIdx.y & 1;
  const int i0 = blockIdx.z * ib;
  const int nrow = min(ib, a.in - i0);
  const T* X = a.X + cfg * (int64_t)N * a.ldx;
  for (int e = threadIdx.x; e < ib * NP; e += blockDim.x) {
    const int ii = e / NP, t = e % NP;
    Xt[e] = (t < N && ii < nrow) ? X[(int64_t)t * a.ldx + i0 + ii] : (T)0;
  }
  __syncthreads();
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= a.out) return;
  const T* G = a.G + ((cfg * 2 + c) * (int64_t)N) * a.ldg + j;
  T g[32];
#pragma unroll
  for (int t = 0; t < 32; ++t) g[t] = t < N ? G[(int64_t)t * a.ldg] : (T)0;
  const int64_t row = (int64_t)c * a.W_total + a.w0 + cfg;
  const int sg = j / a.seg_w, jj = j % a.seg_w;
  T* O = a.O + row * a.ldo + a.seg_base[sg] + jj;
  for (int ii = 0; ii < nrow; ++ii) {
    const T* xr = Xt + ii * NP;
    T s = 0;
#pragma unroll
    for (int t = 0; t < 32; t += 4) {
      if (t < NP) s += xr[t] * g[t] + xr[t + 1] * g[t + 1] + xr[t + 2] * g[t + 2] + xr[t + 3] * g[t + 3];
    }
    O[(int64_t)(i0 + ii) * a.seg_w] = s;
  }
}

// Float variant of the column form with two adjacent columns per thread and 8-byte stores.  A
// workgroup takes R rows of O (R = 1, 2 or 4: R x the pairs of one row <= 256 threads), rows
// 2 cfg + c for the cotangent channels c of its configurations, which share the staged inputs, so a
// 128-wide layer runs four waves per workgroup instead of one.  Taken when every segment offset,
// the segment width and the strides are even, so a column pair never straddles two leaves and the
// stores stay 8-byte aligned.  Each column's sum runs in the same order as in k_wgrad_cols.  NPT
// bounds the padded token count (16 or 32), so the cotangents take NPT registers per column; four
// rows i are swept per iteration (eight independent sums in flight), and O goes out with
// non-temporal stores (nothing reads it before the MinSR passes).
template <int NPT>
__global__ __launch_bounds__(256) void k_wgrad_cols2(WgradArgs<float> a, int ib, int pairs, int R) {
  typedef float f2 __attribute__((ext_vector_type(2)));
  extern __shared__ char smem_raw[];
  float* Xt = reinterpret_cast<float*>(smem_raw);  // [configurations of the workgroup][ib][NP]
  const int N = a.N, NP = (N + 3) & ~3;
  const int64_t row0 = (int64_t)blockIdx.y * R, cfg0 = row0 >> 1;
  const int ncfg = (int)min<int64_t>((R + 1) / 2, a.n_cfg - cfg0);
  const int i0 = blockIdx.z * ib;
  const int nrow = min(ib, a.in - i0);
  for (int e = threadIdx.x; e < ncfg * ib * NP; e += blockDim.x) {
    const int cl = e / (ib * NP), ii = (e / NP) % ib, t = e % NP;
    const float* X = a.X + (cfg0 + cl) * (int64_t)N * a.ldx;
    Xt[e] = (t < N && ii < nrow) ? X[(int64_t)t * a.ldx + i0 + ii] : 0.f;
  }
  __syncthreads();
  const int rl = threadIdx.x / pairs;
  const int64_t row = row0 + rl;
  const int j = 2 * (blockIdx.x * pairs + threadIdx.x % pairs);
  if (rl >= R || row >= 2 * a.n_cfg || j >= a.out) return;
  const int64_t cfg = row >> 1;
  const int c = (int)(row & 1);
  const float* Xc = Xt + (cfg - cfg0) * ib * NP;
  const float* G = a.G + ((cfg * 2 + c) * (int64_t)N) * a.ldg + j;
  float g0[NPT], g1[NPT];
#pragma unroll
  for (int t = 0; t < NPT; ++t) {
    const f2 v = t < N ? *reinterpret_cast<const f2*>(G + (int64_t)t * a.ldg) : (f2)0.f;
    g0[t] = v[0];
    g1[t] = v[1];
  }
  const int64_t orow = (int64_t)c * a.W_total + a.w0 + cfg;
  const int sg = j / a.seg_w, jj = j % a.seg_w;
  float* O = a.O + orow * a.ldo + a.seg_base[sg] + jj;
  auto rowsum = [&](int ii, float& s0, float& s1) {
    const float* xr = Xc + ii * NP;
    s0 = 0.f;
    s1 = 0.f;
#pragma unroll
    for (int t = 0; t < NPT; t += 4) {
      if (t < NP) {
        s0 += xr[t] * g0[t] + xr[t + 1] * g0[t + 1] + xr[t + 2] * g0[t + 2] + xr[t + 3] * g0[t + 3];
        s1 += xr[t] * g1[t] + xr[t + 1] * g1[t + 1] + xr[t + 2] * g1[t + 2] + xr[t + 3] * g1[t + 3];
      }
    }
  };
  int ii = 0;
  for (; ii + 4 <= nrow; ii += 4) {
    float s[4][2];
#pragma unroll
    for (int q = 0; q < 4; ++q) rowsum(ii + q, s[q][0], s[q][1]);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      __builtin_nontemporal_store((f2){s[q][0], s[q][1]},
                                  reinterpret_cast<f2*>(O + (int64_t)(i0 + ii + q) * a.seg_w));
  }
  for (; ii < nrow; ++ii) {
    float s0, s1;
    rowsum(ii, s0, s1);
    __builtin_nontemporal_store((f2){s0, s1}, reinterpret_cast<f2*>(O + (int64_t)(i0 + ii) * a.seg_w));
  }
}

template <typename T>
static bool wgrad_pairs_ok(const WgradArgs<T>& a) {
  if (sizeof(T) != 4 || a.seg_w % 2 || a.ldg % 2 || a.ldo % 2 || ((uintptr_t)a.O & 7) || ((uintptr_t)a.G & 7))
    return false;
  for (int s = 0; s < a.out / a.seg_w; ++s)
    if (a.seg_base[s] % 2) return false;
  return true;
}

// Output columns go in chunks of ob: all of them when the N x out cotangent tile takes at most
// 96 KB of LDS, else the largest multiple of 64 that keeps the tile within 64 KB.
template <typename T>
int launch_wgrad(const WgradArgs<T>& a, hipStream_t st) {
  if (a.n_cfg == 0) return NQS_OK;
  NQS_REQUIRE(a.out % a.seg_w == 0 && a.out / a.seg_w <= 3, NQS_ERR_ARG, "bad output segmentation");
  if (a.N <= 32) {
    const int NP = (a.N + 3) & ~3;
    int ib = (int)std::min<int64_t>(a.in, (48 * 1024) / ((int64_t)NP * sizeof(T)));
    const size_t lds = sizeof(T) * (size_t)ib * NP;
    if constexpr (sizeof(T) == 4) {
      if (wgrad_pairs_ok(a)) {
        // pairs of one row per workgroup (a multiple of 64, at most 256), and rows per workgroup
        const int pairs = (int)std::min<int64_t>(256, round_up(cdiv(a.out, 2), 64));
        int R = pairs <= 64 ? 4 : (pairs <= 128 ? 2 : 1);
        if (R == 4 && 2 * lds > 64 * 1024) R = 2;  // two configurations' inputs would not fit
        const size_t lds2 = lds * (size_t)((R + 1) / 2);
        const dim3 grid((unsigned)cdiv(a.out, 2 * pairs), (unsigned)cdiv(a.n_cfg * 2, R), (unsigned)cdiv(a.in, ib));
        if (a.N <= 16)
          hipLaunchKernelGGL(k_wgrad_cols2<16>, grid, dim3(pairs * R), lds2, st, a, ib, pairs, R);
        else
          hipLaunchKernelGGL(k_wgrad_cols2<32>, grid, dim3(pairs * R), lds2, st, a, ib, pairs, R);
        NQS_LAUNCH_CHECK();
        return NQS_OK;
      }
    }
    hipLaunchKernelGGL(k_wgrad_cols<T>, dim3((unsigned)cdiv(a.out, 256), (unsigned)(a.n_cfg * 2),
                                             (unsigned)cdiv(a.in, ib)),
                       dim3(256), lds, st, a, ib);
    NQS_LAUNCH_CHECK();
    return NQS_OK;
  }
  int ob = a.out;
  if (sizeof(T) * (size_t)a.N * (16 + ob) > 96 * 1024) {
    ob = ((int)(64 * 1024 / (sizeof(T) * (size_t)a.N)) - 16) / 64 * 64;
    NQS_REQUIRE(ob >= 64, NQS_ERR_UNSUPPORTED, "weight-gradient tile exceeds LDS");
  }
  const size_t lds = sizeof(T) * (size_t)a.N * (16 + ob);
  hipLaunchKernelGGL(k_wgrad<T>, dim3((unsigned)cdiv(a.in, 16), (unsigned)(a.n_cfg * 2), (unsigned)cdiv(a.out, ob)),
                     dim3(256), lds, st, a, ob);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// bias gradient: O[c][w][seg(j)] = sum_t G[w][c][t][j]
template <typename T>
__global__ void k_bgrad(WgradArgs<T> a) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.n_cfg * 2 * a.out) return;
  const int j = (int)(q % a.out);
  const int64_t cc = q / a.out;  // cfg * 2 + c
  const int64_t cfg = cc >> 1;
  const int c = (int)(cc & 1);
  const T* G = a.G + (cc * a.N) * a.ldg;
  T s = 0;
  for (int t = 0; t < a.N; ++t) s += G[(int64_t)t * a.ldg + j];
  const int64_t row = (int64_t)c * a.W_total + a.w0 + cfg;
  const int sg = j / a.seg_w, jj = j % a.seg_w;
  a.O[row * a.ldo + a.seg_base[sg] + jj] = s;
}

template <typename T>
int launch_bgrad(const WgradArgs<T>& a, hipStream_t st) {
  const int64_t n = a.n_cfg * 2 * a.out;
  if (n == 0) return NQS_OK;
  hipLaunchKernelGGL(k_bgrad<T>, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, st, a);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// ------------------------------------------------------------------ LayerNorm backward
// y = (x - mu) rs s + b.  For each cotangent channel: g = ybar s,
// xbar = rs (g - mean(g) - xhat mean(g xhat));  O[s] = sum_t ybar xhat;  O[b] = sum_t ybar.
// One wave per configuration, looping over its tokens; xbar is added to xb (accumulate = 1) or
// stored (accumulate = 0).
template <typename T, int MAXF>
__global__ __launch_bounds__(64) void k_ln_bwd(LnBwdArgs<T> a) {
  const int64_t cfg = blockIdx.x;
  const int lane = threadIdx.x;
  const int d = a.d, N = a.N;
  const T inv_d = (T)1 / (T)d;
  T ds[2][MAXF], db[2][MAXF], sc[MAXF];
#pragma unroll
  for (int q = 0; q < MAXF; ++q) {
    const int f = lane + 64 * q;
    sc[q] = (f < d) ? a.scale[f] : (T)0;
    ds[0][q] = ds[1][q] = db[0][q] = db[1][q] = 0;
  }
  for (int t = 0; t < N; ++t) {
    const T* x = a.x + (cfg * N + t) * (int64_t)d;
    T xv[MAXF];
    T s1 = 0, s2 = 0;
#pragma unroll
    for (int q = 0; q < MAXF; ++q) {
      const int f = lane + 64 * q;
      xv[q] = (f < d) ? x[f] : (T)0;
      s1 += xv[q];
      s2 += xv[q] * xv[q];
    }
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    const T mu = s1 * inv_d;
    T var = s2 * inv_d - mu * mu;
    var = var > (T)0 ? var : (T)0;
    const T rs = (T)1 / sqrt(var + (T)1e-6);
    T xh[MAXF];
#pragma unroll
    for (int q = 0; q < MAXF; ++q) xh[q] = (xv[q] - mu) * rs;
    for (int c = 0; c < 2; ++c) {
      const T* yb = a.ybar + ((cfg * 2 + c) * N + t) * (int64_t)d;
      T* xb = a.xbar + ((cfg * 2 + c) * N + t) * (int64_t)d;
      T yv[MAXF], g1 = 0, g2 = 0;
#pragma unroll
      for (int q = 0; q < MAXF; ++q) {
        const int f = lane + 64 * q;
        yv[q] = (f < d) ? yb[f] : (T)0;
        const T g = yv[q] * sc[q];
        g1 += g;
        g2 += g * xh[q];
        ds[c][q] += yv[q] * xh[q];
        db[c][q] += yv[q];
      }
      g1 = wave_sum(g1) * inv_d;
      g2 = wave_sum(g2) * inv_d;
#pragma unroll
      for (int q = 0; q < MAXF; ++q) {
        const int f = lane + 64 * q;
        if (f < d) {
          const T v = rs * (yv[q] * sc[q] - g1 - xh[q] * g2);
          xb[f] = a.accumulate ? xb[f] + v : v;
        }
      }
    }
  }
  for (int c = 0; c < 2; ++c) {
    T* O = a.O + ((int64_t)c * a.W_total + a.w0 + cfg) * a.ldo;
#pragma unroll
    for (int q = 0; q < MAXF; ++q) {
      const int f = lane + 64 * q;
      if (f < d) {
        O[a.off_scale + f] = ds[c][q];
        O[a.off_bias + f] = db[c][q];
      }
    }
  }
}

template <typename T>
int launch_ln_bwd(const LnBwdArgs<T>& a, hipStream_t st) {
  if (a.n_cfg == 0) return NQS_OK;
  if (a.d <= 128)
    hipLaunchKernelGGL((k_ln_bwd<T, 2>), dim3((unsigned)a.n_cfg), dim3(64), 0, st, a);
  else if (a.d <= 256)
    hipLaunchKernelGGL((k_ln_bwd<T, 4>), dim3((unsigned)a.n_cfg), dim3(64), 0, st, a);
  else
    hipLaunchKernelGGL((k_ln_bwd<T, 8>), dim3((unsigned)a.n_cfg), dim3(64), 0, st, a);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// ------------------------------------------------------------------ GELU backward: ubar = gbar * gelu'(u)
template <typename T>
__global__ void k_gelu_bwd(int64_t n_cfg, int64_t per_cfg, const T* __restrict__ u, const T* __restrict__ gbar,
                           T* __restrict__ ubar) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_cfg * 2 * per_cfg) return;
  const int64_t cc = q / per_cfg, r = q % per_cfg;
  const T x = u[(cc >> 1) * per_cfg + r];
  const T k0 = (T)0.7978845608028654, k1 = (T)0.044715;
  const T th = tanh(k0 * (x + k1 * x * x * x));
  const T g1 = (T)0.5 * ((T)1 + th) + (T)0.5 * x * ((T)1 - th * th) * k0 * ((T)1 + (T)3 * k1 * x * x);
  ubar[q] = gbar[q] * g1;
}

template <typename T>
int launch_gelu_bwd(int64_t n_cfg, int64_t per_cfg, const T* u, const T* gbar, T* ubar, hipStream_t st) {
  const int64_t n = n_cfg * 2 * per_cfg;
  if (n == 0) return NQS_OK;
  hipLaunchKernelGGL(k_gelu_bwd<T>, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, st, n_cfg, per_cfg, u, gbar, ubar);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// ------------------------------------------------------------------ attention backward (per cfg, head)
// Recomputes P = softmax(Q K^T sc) from the stored qkv, then per cotangent channel:
//   Vbar = P^T Obar,  Pbar = Obar V^T,  Sbar = P (Pbar - rowdot(Pbar, P)),
//   Qbar = sc Sbar K,  Kbar = sc Sbar^T Q.
template <typename T>
__global__ __launch_bounds__(256) void k_attn_bwd(int64_t n_cfg, int N, int d, int nh, const T* __restrict__ qkv,
                                                  const T* __restrict__ obar, T* __restrict__ qkvbar) {
  extern __shared__ char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  const int hd = d / nh, NN = N * N, NH = N * hd;
  T* Q = sm;
  T* K = Q + NH;
  T* V = K + NH;
  T* Ob = V + NH;
  T* P = Ob + NH;
  T* Sb = P + NN;
  T* rd = Sb + NN;  // N
  const int64_t cfg = blockIdx.x / nh;
  const int h = blockIdx.x % nh;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int64_t rs = 3LL * d;
  const T* in = qkv + cfg * (int64_t)N * rs;
  const T sc = (T)1 / sqrt((T)hd);
  for (int e = tid; e < NH; e += nthr) {
    const int t = e / hd, q = e % hd;
    Q[e] = in[t * rs + h * hd + q];
    K[e] = in[t * rs + d + h * hd + q];
    V[e] = in[t * rs + 2 * d + h * hd + q];
  }
  __syncthreads();
  for (int e = tid; e < NN; e += nthr) {
    const int a = e / N, b = e % N;
    T s = 0;
    for (int q = 0; q < hd; ++q) s += Q[a * hd + q] * K[b * hd + q];
    P[e] = s * sc;
  }
  __syncthreads();
  for (int a = tid; a < N; a += nthr) {
    T mx = P[a * N];
    for (int b = 1; b < N; ++b) mx = P[a * N + b] > mx ? P[a * N + b] : mx;
    T z = 0;
    for (int b = 0; b < N; ++b) {
      const T e_ = exp(P[a * N + b] - mx);
      P[a * N + b] = e_;
      z += e_;
    }
    for (int b = 0; b < N; ++b) P[a * N + b] /= z;
  }
  __syncthreads();
  for (int c = 0; c < 2; ++c) {
    const T* ob = obar + ((cfg * 2 + c) * N) * (int64_t)d;
    T* out = qkvbar + ((cfg * 2 + c) * N) * rs;
    for (int e = tid; e < NH; e += nthr) {
      const int t = e / hd, q = e % hd;
      Ob[e] = ob[t * d + h * hd + q];
    }
    __syncthreads();
    // Vbar and Pbar
    for (int e = tid; e < NH; e += nthr) {
      const int b = e / hd, q = e % hd;
      T s = 0;
      for (int a = 0; a < N; ++a) s += P[a * N + b] * Ob[a * hd + q];
      out[b * rs + 2 * d + h * hd + q] = s;
    }
    for (int e = tid; e < NN; e += nthr) {
      const int a = e / N, b = e % N;
      T s = 0;
      for (int q = 0; q < hd; ++q) s += Ob[a * hd + q] * V[b * hd + q];
      Sb[e] = s;
    }
    __syncthreads();
    for (int a = tid; a < N; a += nthr) {
      T s = 0;
      for (int b = 0; b < N; ++b) s += Sb[a * N + b] * P[a * N + b];
      rd[a] = s;
    }
    __syncthreads();
    for (int e = tid; e < NN; e += nthr) Sb[e] = P[e] * (Sb[e] - rd[e / N]) * sc;
    __syncthreads();
    for (int e = tid; e < NH; e += nthr) {
      const int a = e / hd, q = e % hd;
      T sq = 0, sk = 0;
      for (int b = 0; b < N; ++b) {
        sq += Sb[a * N + b] * K[b * hd + q];
        sk += Sb[b * N + a] * Q[b * hd + q];
      }
      out[a * rs + h * hd + q] = sq;
      out[a * rs + d + h * hd + q] = sk;
    }
    __syncthreads();
  }
}

template <typename T>
int launch_attn_bwd(int64_t n_cfg, int N, int d, int nh, const T* qkv, const T* obar, T* qkvbar, hipStream_t st) {
  if (n_cfg == 0) return NQS_OK;
  const int hd = d / nh;
  const size_t lds = sizeof(T) * (size_t)(4 * N * hd + 2 * N * N + N);
  hipLaunchKernelGGL(k_attn_bwd<T>, dim3((unsigned)(n_cfg * nh)), dim3(256), lds, st, n_cfg, N, d, nh, qkv, obar,
                     qkvbar);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// ------------------------------------------------------------------ vector-Jacobian product
template <typename T>
__global__ void k_vjp_fold(int64_t n_cfg, int64_t per_cfg, int N, int64_t w0, const double* __restrict__ cot,
                           T* __restrict__ ybar) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_cfg * per_cfg) return;
  const int64_t cfg = q / per_cfg, r = q % per_cfg;
  T* y0 = ybar + cfg * 2 * per_cfg + r;
  T* y1 = y0 + per_cfg;
  const double gr = cot[2 * (w0 + cfg)], gi = cot[2 * (w0 + cfg) + 1];
  *y0 = (T)(gr * (double)*y0 + gi * (double)*y1);
  *y1 = (T)0;
}

template <typename T>
int launch_vjp_fold(int64_t n_cfg, int N, int width, int64_t w0, const double* cot, T* ybar, hipStream_t st) {
  const int64_t per = (int64_t)N * width;
  if (n_cfg * per == 0) return NQS_OK;
  hipLaunchKernelGGL(k_vjp_fold<T>, dim3((unsigned)cdiv(n_cfg * per, 256)), dim3(256), 0, st, n_cfg, per, N, w0, cot,
                     ybar);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

int64_t vjp_slices(int64_t n_cfg) { return std::max<int64_t>(1, std::min<int64_t>(256, n_cfg)); }

// part[slice][i][j] = sum over cfg in the slice, t of X[cfg][t][i] G[cfg][0][t][j]; one thread per
// output column j and 16 rows i (X values are wave-uniform and come through the scalar cache).
template <typename T>
__global__ __launch_bounds__(256) void k_vjp_wsum(WgradArgs<T> a, T* __restrict__ part, int64_t per_slice) {
  constexpr int IB = 16;
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int i0 = blockIdx.y * IB;
  const int64_t s = blockIdx.z;
  const int64_t c0 = s * per_slice, c1 = std::min(a.n_cfg, c0 + per_slice);
  const int N = a.N;
  const int in = a.X ? a.in : 1;
  T acc[IB];
#pragma unroll
  for (int ii = 0; ii < IB; ++ii) acc[ii] = (T)0;
  if (j < a.out) {
    for (int64_t cfg = c0; cfg < c1; ++cfg) {
      const T* G = a.G + (cfg * 2 * N) * a.ldg + j;
      const T* X = a.X ? a.X + (cfg * N) * a.ldx + i0 : nullptr;
      for (int t = 0; t < N; ++t) {
        const T gv = G[(int64_t)t * a.ldg];
        if (X) {
#pragma unroll
          for (int ii = 0; ii < IB; ++ii)
            if (i0 + ii < in) acc[ii] += X[(int64_t)t * a.ldx + ii] * gv;
        } else {
          acc[0] += gv;
        }
      }
    }
  }
  if (j >= a.out) return;
#pragma unroll
  for (int ii = 0; ii < IB; ++ii)
    if (i0 + ii < in) part[(s * in + i0 + ii) * (int64_t)a.out + j] = acc[ii];
}

template <typename T>
__global__ void k_vjp_wsum_reduce(WgradArgs<T> a, const T* __restrict__ part, int64_t n_slices, double* __restrict__ g) {
  const int in = a.X ? a.in : 1;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)in * a.out) return;
  const int i = (int)(q / a.out), j = (int)(q % a.out);
  double t = 0.0;
  for (int64_t s = 0; s < n_slices; ++s) t += (double)part[(s * in + i) * (int64_t)a.out + j];
  const int sg = j / a.seg_w, jj = j % a.seg_w;
  g[a.seg_base[sg] + (int64_t)i * a.seg_w + jj] += t;
}

template <typename T>
int launch_vjp_wsum(const WgradArgs<T>& a, T* part, int64_t n_slices, double* g, hipStream_t st) {
  if (a.n_cfg == 0) return NQS_OK;
  NQS_REQUIRE(a.out % a.seg_w == 0 && a.out / a.seg_w <= 3, NQS_ERR_ARG, "bad output segmentation");
  const int in = a.X ? a.in : 1;
  const int64_t per = cdiv(a.n_cfg, n_slices);
  const int64_t ns = cdiv(a.n_cfg, per);
  hipLaunchKernelGGL(k_vjp_wsum<T>, dim3((unsigned)cdiv(a.out, 256), (unsigned)cdiv(in, 16), (unsigned)ns), dim3(256),
                     0, st, a, part, per);
  NQS_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_vjp_wsum_reduce<T>, dim3((unsigned)cdiv((int64_t)in * a.out, 256)), dim3(256), 0, st, a, part,
                     ns, g);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

template <typename T>
__global__ void k_vjp_osm_reduce(const T* __restrict__ osm, int64_t ldsm, int64_t n_cfg, OsmMap map,
                                 const double* __restrict__ cot, int64_t w0, double* __restrict__ g) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= map.start[map.n]) return;
  int sgm = 0;
  while (sgm + 1 < map.n && q >= map.start[sgm + 1]) ++sgm;
  double t = 0.0;
  for (int64_t cfg = 0; cfg < n_cfg; ++cfg) {
    const double v0 = (double)osm[cfg * ldsm + q];
    if (cot) {
      const double v1 = (double)osm[(n_cfg + cfg) * ldsm + q];
      t += cot[2 * (w0 + cfg)] * v0 + cot[2 * (w0 + cfg) + 1] * v1;
    } else {
      t += v0;
    }
  }
  g[map.dst[sgm] + q - map.start[sgm]] += t;
}

template <typename T>
int launch_vjp_osm_reduce(const T* osm, int64_t ldsm, int64_t n_cfg, const OsmMap& map, const double* cot,
                          int64_t w0, double* g, hipStream_t st) {
  const int64_t ncol = map.n > 0 ? map.start[map.n] : 0;
  if (n_cfg == 0 || ncol == 0) return NQS_OK;
  hipLaunchKernelGGL(k_vjp_osm_reduce<T>, dim3((unsigned)cdiv(ncol, 64)), dim3(64), 0, st, osm, ldsm, n_cfg, map, cot,
                     w0, g);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

#define NQS_INST_B(T)                                                                                          \
  template int launch_wgrad<T>(const WgradArgs<T>&, hipStream_t);                                            \
  template int launch_bgrad<T>(const WgradArgs<T>&, hipStream_t);                                            \
  template int launch_ln_bwd<T>(const LnBwdArgs<T>&, hipStream_t);                                           \
  template int launch_gelu_bwd<T>(int64_t, int64_t, const T*, const T*, T*, hipStream_t);                    \
  template int launch_attn_bwd<T>(int64_t, int, int, int, const T*, const T*, T*, hipStream_t);                \
  template int launch_vjp_fold<T>(int64_t, int, int, int64_t, const double*, T*, hipStream_t);               \
  template int launch_vjp_wsum<T>(const WgradArgs<T>&, T*, int64_t, double*, hipStream_t);                   \
  template int launch_vjp_osm_reduce<T>(const T*, int64_t, int64_t, const OsmMap&, const double*, int64_t,   \
                                        double*, hipStream_t);
NQS_INST_B(double)
NQS_INST_B(float)

}  // namespace nqs
