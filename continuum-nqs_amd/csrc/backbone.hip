// Psiformer backbone jet kernels (value + tangents + Laplacian) for gfx950.
//
// A "jet" carries, per configuration, the value of an activation, its derivatives along nt
// tangent directions (positions of electrons) and, optionally, its Laplacian summed over all
// 2N position coordinates.  Layout [cfg][channel][token][feature]; linear layers act on every
// channel alike (bias on the value channel only, see gemm.hip), the kernels here carry the
// non-linear ops with the first- and second-order chain rule:
//   f(u):      df = f'(u) du,          Lap f = f'(u) Lap u + f''(u) sum_c du_c^2
//   a*b:       d(ab) = a db + b da,    Lap(ab) = a Lap b + b Lap a + 2 sum_c da_c db_c
// Reference ops (src/nqs/psiformer.py): features 84-97, LayerNorm 76-77,136 (flax eps 1e-6,
// fast variance), attention 50-64, tanh-GELU 79.
#include <algorithm>
#include "nqs_common.h"
#include "kernels.h"
#include "tune.h"
#include "attn.h"

namespace nqs {

// ------------------------------------------------------------------ features (psiformer.py:84-97,128-130)
// Row of token t: [x: sin m=1..F, cos m=1..F, y: sin, cos, onehot(s_t) (2), zero pad to kpad].
template <typename T>
__global__ void k_features(JetGeom j, const double* __restrict__ R, const int32_t* __restrict__ S,
                           double ai00, double ai01, double ai10, double ai11, int F, int kpad,
                           int64_t w0, T* __restrict__ X) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int N = j.N;
  if (gid >= j.n_cfg * N) return;
  const int64_t cfg = gid / N;
  const int t = (int)(gid % N);
  int64_t w;
  int flip = -1;
  if (j.mode == 2) {
    w = w0 + cfg / N;
    flip = (int)(cfg % N);
  } else {
    w = w0 + cfg;
  }
  const double rx = R[(w * N + t) * 2 + 0], ry = R[(w * N + t) * 2 + 1];
  int s = S[w * N + t];
  if (t == flip) s = 1 - s;
  const double sa[2] = {rx * ai00 + ry * ai01, rx * ai10 + ry * ai11};
  const double ainv[2][2] = {{ai00, ai01}, {ai10, ai11}};
  const int C = j.C;
  T* base = X + (cfg * C) * (int64_t)N * kpad;
  const int64_t chs = (int64_t)N * kpad;  // channel stride
  // value channel
  T* v = base + (int64_t)t * kpad;
  for (int a = 0; a < 2; ++a) {
    const double ph = 2.0 * M_PI * sa[a];
    for (int m = 1; m <= F; ++m) {
      double sn, cs;
      sincos(ph * (double)m, &sn, &cs);
      v[a * 2 * F + (m - 1)] = (T)sn;
      v[a * 2 * F + F + (m - 1)] = (T)cs;
      // tangents
      for (int c = 1; c <= j.nt; ++c) {
        int e, b;
        if (j.mode == 2) {
          e = flip;
          b = c - 1;
        } else {
          e = (c - 1) >> 1;
          b = (c - 1) & 1;
        }
        T* tc = base + c * chs + (int64_t)t * kpad;
        const double dang = (t == e) ? 2.0 * M_PI * m * ainv[a][b] : 0.0;
        tc[a * 2 * F + (m - 1)] = (T)(cs * dang);
        tc[a * 2 * F + F + (m - 1)] = (T)(-sn * dang);
      }
      if (j.has_lap) {
        T* lc = base + (j.nt + 1) * chs + (int64_t)t * kpad;
        const double w2 = (2.0 * M_PI * m) * (2.0 * M_PI * m) * (ainv[a][0] * ainv[a][0] + ainv[a][1] * ainv[a][1]);
        lc[a * 2 * F + (m - 1)] = (T)(-sn * w2);
        lc[a * 2 * F + F + (m - 1)] = (T)(-cs * w2);
      }
    }
  }
  for (int c = 0; c < C; ++c) {
    T* row = base + c * chs + (int64_t)t * kpad;
    row[4 * F] = (c == 0) ? (T)(s == 0 ? 1 : 0) : (T)0;
    row[4 * F + 1] = (c == 0) ? (T)(s == 1 ? 1 : 0) : (T)0;
    for (int k = 4 * F + 2; k < kpad; ++k) row[k] = (T)0;
  }
}

template <typename T>
int launch_features(const JetGeom& j, const double* R, const int32_t* S, const double* Ainv,
                    int n_freq, int kpad, int64_t w0, T* X, hipStream_t st) {
  const int64_t n = j.n_cfg * j.N;
  if (n == 0) return NQS_OK;
  hipLaunchKernelGGL(k_features<T>, dim3((unsigned)cdiv(n, 128)), dim3(128), 0, st, j, R, S, Ainv[0],
                     Ainv[1], Ainv[2], Ainv[3], n_freq, kpad, w0, X);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// ------------------------------------------------------------------ LayerNorm jet
// One wave per (cfg, token).  mu = E[x], var = E[x^2] - mu^2, rs = (var + eps)^-1/2,
// y = (x - mu) rs scale + bias.
// F features per lane of a 16-lane row group, loaded / stored as 16-byte vectors when the row is
// exactly 16 F wide (zero past d otherwise)
template <typename T, int F>
__device__ __forceinline__ void ln_load(const T* __restrict__ p, int j, int d, T (&v)[F]) {
  constexpr int VN = 16 / sizeof(T);
  if (F % VN == 0 && d == 16 * F) {
    typedef T vt __attribute__((ext_vector_type(VN)));
#pragma unroll
    for (int q = 0; q < F / VN; ++q) {
      const vt w = *reinterpret_cast<const vt*>(p + F * j + VN * q);
#pragma unroll
      for (int e = 0; e < VN; ++e) v[VN * q + e] = w[e];
    }
  } else {
#pragma unroll
    for (int e = 0; e < F; ++e) v[e] = (F * j + e < d) ? p[F * j + e] : (T)0;
  }
}
template <typename T, int F>
__device__ __forceinline__ void ln_store(T* __restrict__ p, int j, int d, const T (&v)[F]) {
  constexpr int VN = 16 / sizeof(T);
  if (F % VN == 0 && d == 16 * F) {
    typedef T vt __attribute__((ext_vector_type(VN)));
#pragma unroll
    for (int q = 0; q < F / VN; ++q) {
      vt w;
#pragma unroll
      for (int e = 0; e < VN; ++e) w[e] = v[VN * q + e];
      *reinterpret_cast<vt*>(p + F * j + VN * q) = w;
    }
  } else {
#pragma unroll
    for (int e = 0; e < F; ++e)
      if (F * j + e < d) p[F * j + e] = v[e];
  }
}

// LayerNorm jet: four token rows per 64-lane block, one 16-lane row group per token (ln_value's
// layout: lane j holds features F j .. F j + F - 1), every channel of the token in turn.  First-order
// tangents use ln_tangent; with a Laplacian channel the same tangent rule is written out with the
// second-order accumulators beside it.
template <typename T, int F>
__global__ __launch_bounds__(64) void k_layernorm_jet(JetGeom jg, int d, const T* __restrict__ x,
                                                      const T* __restrict__ scale, const T* __restrict__ bias,
                                                      T* __restrict__ y, int64_t rows) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x, j = lane & 15;
  const int64_t row = (int64_t)blockIdx.x * 4 + (lane >> 4);  // cfg * N + t
  if (row >= rows) return;  // whole 16-lane groups leave together
  const int64_t cfg = row / jg.N;
  const int t = (int)(row % jg.N);
  const int64_t chs = (int64_t)jg.N * d;
  const T* xb = x + cfg * jg.C * chs + (int64_t)t * d;
  T* yb = y + cfg * jg.C * chs + (int64_t)t * d;
  const T inv_d = (T)1 / (T)d;
  T xv[F], sc[F], y0[F];
  ln_load<T, F>(xb, j, d, xv);
  T mu, rs;
  ln_value<T, F>(xv, j, d, scale, bias, sc, y0, mu, rs);
  ln_store<T, F>(yb, j, d, y0);
  if (!jg.has_lap) {
    for (int c = 1; c <= jg.nt; ++c) {
      T dx[F], dy[F];
      ln_load<T, F>(xb + c * chs, j, d, dx);
      ln_tangent<T, F>(xv, dx, sc, d, mu, rs, dy);
      ln_store<T, F>(yb + c * chs, j, d, dy);
    }
    return;
  }
  T s1[F];
#pragma unroll
  for (int e = 0; e < F; ++e) s1[e] = 0;
  T acc_dx2 = 0, acc_dmu2 = 0, acc_dvar2 = 0;
  for (int c = 1; c <= jg.nt; ++c) {
    T dx[F], dy[F];
    ln_load<T, F>(xb + c * chs, j, d, dx);
    T sd = 0, sxd = 0, sdd = 0;
#pragma unroll
    for (int e = 0; e < F; ++e) {
      sd += dx[e];
      sxd += xv[e] * dx[e];
      sdd += dx[e] * dx[e];
    }
    sd = row_sum16(sd);
    sxd = row_sum16(sxd);
    const T dmu = sd * inv_d;
    const T dvar = (T)2 * (sxd * inv_d - mu * dmu);
    const T drs = (T)(-0.5) * rs * rs * rs * dvar;
#pragma unroll
    for (int e = 0; e < F; ++e) {
      const T u = dx[e] - dmu;
      dy[e] = (u * rs + (xv[e] - mu) * drs) * sc[e];
      s1[e] += u * drs;
    }
    ln_store<T, F>(yb + c * chs, j, d, dy);
    acc_dx2 += row_sum16(sdd) * inv_d;
    acc_dmu2 += dmu * dmu;
    acc_dvar2 += dvar * dvar;
  }
  const int cl = jg.nt + 1;
  T lx[F], ly[F];
  ln_load<T, F>(xb + cl * chs, j, d, lx);
  T sl = 0, sxl = 0;
#pragma unroll
  for (int e = 0; e < F; ++e) {
    sl += lx[e];
    sxl += xv[e] * lx[e];
  }
  sl = row_sum16(sl);
  sxl = row_sum16(sxl);
  const T lmu = sl * inv_d;
  const T lvar = (T)2 * (sxl * inv_d) + (T)2 * acc_dx2 - (T)2 * mu * lmu - (T)2 * acc_dmu2;
  const T rs3 = rs * rs * rs;
  const T lrs = (T)(-0.5) * rs3 * lvar + (T)0.75 * rs3 * rs * rs * acc_dvar2;
#pragma unroll
  for (int e = 0; e < F; ++e) ly[e] = ((lx[e] - lmu) * rs + (xv[e] - mu) * lrs + (T)2 * s1[e]) * sc[e];
  ln_store<T, F>(yb + cl * chs, j, d, ly);
}

template <typename T>
int launch_layernorm_jet(const JetGeom& j, int d, const T* x, const T* scale, const T* bias, T* y,
                         hipStream_t st) {
  const int64_t rows = j.n_cfg * j.N;
  if (rows == 0) return NQS_OK;
  NQS_REQUIRE(d <= 512, NQS_ERR_UNSUPPORTED, "LayerNorm width above 512 is not supported");
  const dim3 grid((unsigned)cdiv(rows, 4));
  if (d <= 32)
    hipLaunchKernelGGL((k_layernorm_jet<T, 2>), grid, dim3(64), 0, st, j, d, x, scale, bias, y, rows);
  else if (d <= 64)
    hipLaunchKernelGGL((k_layernorm_jet<T, 4>), grid, dim3(64), 0, st, j, d, x, scale, bias, y, rows);
  else if (d <= 128)
    hipLaunchKernelGGL((k_layernorm_jet<T, 8>), grid, dim3(64), 0, st, j, d, x, scale, bias, y, rows);
  else if (d <= 256)
    hipLaunchKernelGGL((k_layernorm_jet<T, 16>), grid, dim3(64), 0, st, j, d, x, scale, bias, y, rows);
  else
    hipLaunchKernelGGL((k_layernorm_jet<T, 32>), grid, dim3(64), 0, st, j, d, x, scale, bias, y, rows);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// ------------------------------------------------------------------ tanh-GELU jet

template <typename T>
__global__ void k_gelu_jet(JetGeom j, int width, const T* __restrict__ x, T* __restrict__ y) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per_cfg = (int64_t)j.N * width;
  if (gid >= j.n_cfg * per_cfg) return;
  const int64_t cfg = gid / per_cfg, rem = gid % per_cfg;
  const int64_t base = cfg * j.C * per_cfg + rem;
  const T xv = x[base];
  T g, g1, g2;
  gelu_derivs(xv, g, g1, g2);
  y[base] = g;
  T s2 = 0;
  for (int c = 1; c <= j.nt; ++c) {
    const T dx = x[base + c * per_cfg];
    y[base + c * per_cfg] = g1 * dx;
    s2 += dx * dx;
  }
  if (j.has_lap) {
    const int64_t o = base + (int64_t)(j.nt + 1) * per_cfg;
    y[o] = g1 * x[o] + g2 * s2;
  }
}

template <typename T>
int launch_gelu_jet(const JetGeom& j, int width, const T* x, T* y, hipStream_t st) {
  const int64_t n = j.n_cfg * j.N * width;
  if (n == 0) return NQS_OK;
  hipLaunchKernelGGL(k_gelu_jet<T>, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, st, j, width, x, y);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// ------------------------------------------------------------------ attention jet (psiformer.py:50-64)
// One workgroup per (cfg, head).  qkv rows have 3d columns [q | k | v], head h owns columns
// [h hd, (h+1) hd) of each.  S = Q K^T / sqrt(hd), P = softmax_rows(S), O = P V.
//   dS_c = (dQ_c K^T + Q dK_c^T) sc
//   dP_c = P (dS_c - abar_c),            abar_c = rowdot(P, dS_c)
//   Lap S = (LapQ K^T + Q LapK^T + 2 sum_c dQ_c dK_c^T) sc
//   Lap P = P (Lap S - rowdot(P, Lap S) + sum_c [(dS_c - abar_c)^2 - rowdot(P, (dS_c - abar_c)^2)])
//   dO_c = dP_c V + P dV_c,   Lap O = Lap P V + P Lap V + 2 sum_c dP_c dV_c
// Wave-level variant: one 64-lane wave per (cfg, head), `wpg` waves per workgroup, each with its
// own LDS region; phases inside a wave are ordered by wave-scope fences (LDS operations of one
// wave complete in issue order), so no workgroup barrier is needed.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

template <typename T>
__global__ __launch_bounds__(256) void k_attention_jet_wave(JetGeom j, int d, int nh, const T* __restrict__ qkv,
                                                            T* __restrict__ out, int wpg, int64_t npairs) {
  extern __shared__ char smem_raw[];
  const int N = j.N, hd = d / nh;
  const int NN = N * N, NH = N * hd;
  const int per_wave = 7 * NH + 5 * NN + 2 * N;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t pair = (int64_t)blockIdx.x * wpg + wave;
  if (wave >= wpg || pair >= npairs) return;
  T* sm = reinterpret_cast<T*>(smem_raw) + (size_t)wave * per_wave;
  T* Q0 = sm;
  T* K0 = Q0 + NH;
  T* V0 = K0 + NH;
  T* Qc = V0 + NH;
  T* Kc = Qc + NH;
  T* Vc = Kc + NH;
  T* P = Vc + NH;
  T* Sc = P + NN;
  T* accS = Sc + NN;
  T* accP = accS + NN;
  T* accPV = accP + NN;
  T* ab = accPV + NH;
  T* ab2 = ab + N;
  const int64_t cfg = pair / nh;
  const int h = (int)(pair % nh);
  const int64_t rs = 3 * (int64_t)d;
  const int64_t chs_in = (int64_t)N * rs;
  const int64_t chs_out = (int64_t)N * d;
  const T* in0 = qkv + cfg * j.C * chs_in;
  T* out0 = out + cfg * j.C * chs_out;
  const T sc = (T)1 / sqrt((T)hd);
  auto load = [&](int c, T* q, T* k, T* v) {
    const T* b = in0 + c * chs_in;
    for (int e = lane; e < NH; e += 64) {
      const int t = e / hd, q_ = e % hd;
      q[e] = b[t * rs + h * hd + q_];
      k[e] = b[t * rs + d + h * hd + q_];
      v[e] = b[t * rs + 2 * d + h * hd + q_];
    }
  };
  load(0, Q0, K0, V0);
  wave_sync();
  for (int e = lane; e < NN; e += 64) {
    const int a = e / N, b = e % N;
    T s = 0;
    for (int q = 0; q < hd; ++q) s += Q0[a * hd + q] * K0[b * hd + q];
    P[e] = s * sc;
  }
  wave_sync();
  for (int a = lane; a < N; a += 64) {
    T mx = P[a * N];
    for (int b = 1; b < N; ++b) mx = P[a * N + b] > mx ? P[a * N + b] : mx;
    T z = 0;
    for (int b = 0; b < N; ++b) {
      const T e_ = exp(P[a * N + b] - mx);
      P[a * N + b] = e_;
      z += e_;
    }
    const T iz = (T)1 / z;
    for (int b = 0; b < N; ++b) P[a * N + b] *= iz;
  }
  wave_sync();
  for (int e = lane; e < NH; e += 64) {
    const int a = e / hd, q = e % hd;
    T o = 0;
    for (int b = 0; b < N; ++b) o += P[a * N + b] * V0[b * hd + q];
    out0[a * d + h * hd + q] = o;
  }
  if (j.nt == 0 && !j.has_lap) return;
  for (int e = lane; e < NN; e += 64) {
    accS[e] = 0;
    accP[e] = 0;
  }
  for (int e = lane; e < NH; e += 64) accPV[e] = 0;
  for (int c = 1; c <= j.nt; ++c) {
    load(c, Qc, Kc, Vc);
    wave_sync();
    for (int e = lane; e < NN; e += 64) {
      const int a = e / N, b = e % N;
      T s = 0, s2 = 0;
      for (int q = 0; q < hd; ++q) {
        s += Qc[a * hd + q] * K0[b * hd + q] + Q0[a * hd + q] * Kc[b * hd + q];
        s2 += Qc[a * hd + q] * Kc[b * hd + q];
      }
      Sc[e] = s * sc;
      accS[e] += s2;
    }
    wave_sync();
    for (int a = lane; a < N; a += 64) {
      T m = 0;
      for (int b = 0; b < N; ++b) m += P[a * N + b] * Sc[a * N + b];
      T m2 = 0;
      for (int b = 0; b < N; ++b) {
        const T u = Sc[a * N + b] - m;
        m2 += P[a * N + b] * u * u;
      }
      ab[a] = m;
      ab2[a] = m2;
    }
    wave_sync();
    for (int e = lane; e < NN; e += 64) {
      const int a = e / N;
      const T u = Sc[e] - ab[a];
      accP[e] += u * u - ab2[a];
      Sc[e] = P[e] * u;
    }
    wave_sync();
    T* oc = out0 + c * chs_out;
    for (int e = lane; e < NH; e += 64) {
      const int a = e / hd, q = e % hd;
      T o = 0, o2 = 0;
      for (int b = 0; b < N; ++b) {
        o += Sc[a * N + b] * V0[b * hd + q] + P[a * N + b] * Vc[b * hd + q];
        o2 += Sc[a * N + b] * Vc[b * hd + q];
      }
      oc[a * d + h * hd + q] = o;
      accPV[e] += o2;
    }
    wave_sync();
  }
  if (!j.has_lap) return;
  const int cl = j.nt + 1;
  load(cl, Qc, Kc, Vc);
  wave_sync();
  for (int e = lane; e < NN; e += 64) {
    const int a = e / N, b = e % N;
    T s = 0;
    for (int q = 0; q < hd; ++q) s += Qc[a * hd + q] * K0[b * hd + q] + Q0[a * hd + q] * Kc[b * hd + q];
    Sc[e] = (s + (T)2 * accS[e]) * sc;
  }
  wave_sync();
  for (int a = lane; a < N; a += 64) {
    T m = 0;
    for (int b = 0; b < N; ++b) m += P[a * N + b] * Sc[a * N + b];
    ab[a] = m;
  }
  wave_sync();
  for (int e = lane; e < NN; e += 64) {
    const int a = e / N;
    Sc[e] = P[e] * (Sc[e] - ab[a] + accP[e]);
  }
  wave_sync();
  T* oc = out0 + cl * chs_out;
  for (int e = lane; e < NH; e += 64) {
    const int a = e / hd, q = e % hd;
    T o = 0;
    for (int b = 0; b < N; ++b) o += Sc[a * N + b] * V0[b * hd + q] + P[a * N + b] * Vc[b * hd + q];
    oc[a * d + h * hd + q] = o + (T)2 * accPV[e];
  }
}


template <typename T, int NQ>
__device__ __forceinline__ void attn_store(T* __restrict__ o, int lane, int N, int hd, int d, int h,
                                           const typename Mma16<T>::acc_t (&acc)[NQ], T scale2,
                                           const typename Mma16<T>::acc_t (&add)[NQ]) {
  const int li = lane & 15;
#pragma unroll
  for (int t = 0; t < NQ; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int a = Mma16<T>::row(lane, r), q = 16 * t + li;
      if (a < N && q < hd) o[a * d + h * hd + q] = acc[t][r] + scale2 * add[t][r];
    }
}

template <typename T, int NQ>
__global__ __launch_bounds__(512) void k_attention_jet_mfma(JetGeom j, int d, int nh, const T* __restrict__ qkv,
                                                            T* __restrict__ out, int64_t npairs) {
  // no FMA contraction: a first-order tangent channel rounds as attn_tangent does in the fused forward
#pragma clang fp contract(off)
  typedef typename Mma16<T>::acc_t acc_t;
  const int lane = threadIdx.x & 63;
  const int64_t pair = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (pair >= npairs) return;
  const int N = j.N, hd = d / nh;
  const int64_t cfg = pair / nh;
  const int h = (int)(pair % nh);
  const int64_t rs = 3 * (int64_t)d;
  const int64_t chs_in = (int64_t)N * rs, chs_out = (int64_t)N * d;
  const T* in0 = qkv + cfg * j.C * chs_in;
  T* out0 = out + cfg * j.C * chs_out;
  const T sc = (T)1 / sqrt((T)hd);
  const acc_t zacc = (acc_t)(T)0;
  acc_t zq[NQ];
#pragma unroll
  for (int t = 0; t < NQ; ++t) zq[t] = zacc;
  // ---- value channel
  AttnRegs<T, NQ> x0;
  attn_load<T, NQ>(in0, lane, N, hd, d, h, rs, x0);
  T P[4];
  {
    acc_t o[NQ];
    attn_value<T, NQ>(x0, lane, N, sc, P, o);
    attn_store<T, NQ>(out0, lane, N, hd, d, h, o, (T)0, zq);
  }
  if (j.nt == 0 && !j.has_lap) return;
  // ---- tangent channels
  acc_t accS = zacc, accPV[NQ];
  T accP[4] = {0, 0, 0, 0};
#pragma unroll
  for (int t = 0; t < NQ; ++t) accPV[t] = zacc;
  for (int c = 1; c <= j.nt; ++c) {
    AttnRegs<T, NQ> xc;
    attn_load<T, NQ>(in0 + c * chs_in, lane, N, hd, d, h, rs, xc);
    acc_t sa = zacc;
#pragma unroll
    for (int s = 0; s < 4 * NQ; ++s) {
      sa = Mma16<T>::mma(x0.k[s], xc.q[s], sa);
      sa = Mma16<T>::mma(xc.k[s], x0.q[s], sa);
      accS = Mma16<T>::mma(xc.k[s], xc.q[s], accS);
    }
    T u[4], Pc[4], m = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) m += P[r] * (sa[r] * sc);
    m = xsum16(m);
    T m2 = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      u[r] = sa[r] * sc - m;
      m2 += P[r] * u[r] * u[r];
    }
    m2 = xsum16(m2);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      accP[r] += u[r] * u[r] - m2;
      Pc[r] = P[r] * u[r];
    }
    acc_t o[NQ];
#pragma unroll
    for (int t = 0; t < NQ; ++t) {
      o[t] = zacc;
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) {
        o[t] = Mma16<T>::mma(Pc[kb], x0.v[4 * t + kb], o[t]);
        o[t] = Mma16<T>::mma(P[kb], xc.v[4 * t + kb], o[t]);
        accPV[t] = Mma16<T>::mma(Pc[kb], xc.v[4 * t + kb], accPV[t]);
      }
    }
    attn_store<T, NQ>(out0 + c * chs_out, lane, N, hd, d, h, o, (T)0, zq);
  }
  if (!j.has_lap) return;
  // ---- Laplacian channel
  const int cl = j.nt + 1;
  AttnRegs<T, NQ> xl;
  attn_load<T, NQ>(in0 + cl * chs_in, lane, N, hd, d, h, rs, xl);
  acc_t sl = zacc;
#pragma unroll
  for (int s = 0; s < 4 * NQ; ++s) {
    sl = Mma16<T>::mma(x0.k[s], xl.q[s], sl);
    sl = Mma16<T>::mma(xl.k[s], x0.q[s], sl);
  }
  T w[4], ml = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    w[r] = (sl[r] + (T)2 * accS[r]) * sc;
    ml += P[r] * w[r];
  }
  ml = xsum16(ml);
  T Pl[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) Pl[r] = P[r] * (w[r] - ml + accP[r]);
  acc_t o[NQ];
#pragma unroll
  for (int t = 0; t < NQ; ++t) {
    o[t] = zacc;
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      o[t] = Mma16<T>::mma(Pl[kb], x0.v[4 * t + kb], o[t]);
      o[t] = Mma16<T>::mma(P[kb], xl.v[4 * t + kb], o[t]);
    }
  }
  attn_store<T, NQ>(out0 + cl * chs_out, lane, N, hd, d, h, o, (T)2, accPV);
}

// 16 < N <= 32: the same register/MFMA scheme with two key blocks, one wave per (cfg, head,
// query block of 16).  Keys 16 kb + row(lane, r) sit in slot (kb, r); the softmax of a query
// runs over both key blocks; P V sums the two blocks.  Head width hd == 16 NQ.
template <typename T, int NQ>
__global__ __launch_bounds__(512) void k_attention_jet_mfma32(JetGeom j, int d, int nh, const T* __restrict__ qkv,
                                                              T* __restrict__ out, int64_t nwork) {
  typedef typename Mma16<T>::acc_t acc_t;
  constexpr int NKB = 2;
  const int lane = threadIdx.x & 63, li = lane & 15, g = lane >> 4;
  const int64_t work = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (work >= nwork) return;
  const int N = j.N, hd = d / nh;
  const int qb = (int)(work % NKB);
  const int64_t pair = work / NKB;
  const int64_t cfg = pair / nh;
  const int h = (int)(pair % nh);
  const int64_t rs = 3 * (int64_t)d;
  const int64_t chs_in = (int64_t)N * rs, chs_out = (int64_t)N * d;
  const T* in0 = qkv + cfg * j.C * chs_in;
  T* out0 = out + cfg * j.C * chs_out;
  const T sc = (T)1 / sqrt((T)hd);
  const acc_t zacc = (acc_t)(T)0;
  // q of this wave's 16 queries, k and v of all 32 keys (zeros past N)
  auto load = [&](const T* base, T (&q)[4 * NQ], T (&k)[NKB][4 * NQ], T (&v)[NKB][4 * NQ]) {
    const int tq = 16 * qb + li;
    const int tqc = tq < N ? tq : N - 1;
#pragma unroll
    for (int s = 0; s < 4 * NQ; ++s) {
      const int f = 16 * (s >> 2) + 4 * g + (s & 3);
      const T vq = base[tqc * rs + h * hd + f];
      q[s] = tq < N ? vq : (T)0;
#pragma unroll
      for (int kb = 0; kb < NKB; ++kb) {
        const int tk = 16 * kb + li, tkc = tk < N ? tk : N - 1;
        const T vk = base[tkc * rs + d + h * hd + f];
        k[kb][s] = tk < N ? vk : (T)0;
        const int b = 16 * kb + Mma16<T>::row(lane, s & 3), bc = b < N ? b : N - 1;
        const T vv = base[bc * rs + 2 * d + h * hd + 16 * (s >> 2) + li];
        v[kb][s] = b < N ? vv : (T)0;
      }
    }
  };
  auto store = [&](T* o, const acc_t (&acc)[NQ], T scale2, const acc_t (&add)[NQ]) {
#pragma unroll
    for (int t = 0; t < NQ; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int a = 16 * qb + Mma16<T>::row(lane, r), f = 16 * t + li;
        if (a < N) o[a * d + h * hd + f] = acc[t][r] + scale2 * add[t][r];
      }
  };
  acc_t zq[NQ];
#pragma unroll
  for (int t = 0; t < NQ; ++t) zq[t] = zacc;
  // ---- value channel
  T q0[4 * NQ], k0[NKB][4 * NQ], v0[NKB][4 * NQ];
  load(in0, q0, k0, v0);
  T P[NKB][4];
  {
    T sv[NKB][4], mx = -INFINITY;
#pragma unroll
    for (int kb = 0; kb < NKB; ++kb) {
      acc_t s0 = zacc;
#pragma unroll
      for (int s = 0; s < 4 * NQ; ++s) s0 = Mma16<T>::mma(k0[kb][s], q0[s], s0);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        sv[kb][r] = 16 * kb + Mma16<T>::row(lane, r) < N ? s0[r] * sc : -INFINITY;
        mx = sv[kb][r] > mx ? sv[kb][r] : mx;
      }
    }
    T o = __shfl_xor(mx, 16);
    mx = o > mx ? o : mx;
    o = __shfl_xor(mx, 32);
    mx = o > mx ? o : mx;
    T z = 0;
#pragma unroll
    for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        P[kb][r] = exp(sv[kb][r] - mx);
        z += P[kb][r];
      }
    const T iz = (T)1 / xsum16(z);
#pragma unroll
    for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
      for (int r = 0; r < 4; ++r) P[kb][r] *= iz;
  }
  {
    acc_t o[NQ];
#pragma unroll
    for (int t = 0; t < NQ; ++t) {
      o[t] = zacc;
#pragma unroll
      for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) o[t] = Mma16<T>::mma(P[kb][s4], v0[kb][4 * t + s4], o[t]);
    }
    store(out0, o, (T)0, zq);
  }
  if (j.nt == 0 && !j.has_lap) return;
  // ---- tangent channels
  acc_t accS[NKB], accPV[NQ];
  T accP[NKB][4];
#pragma unroll
  for (int kb = 0; kb < NKB; ++kb) {
    accS[kb] = zacc;
#pragma unroll
    for (int r = 0; r < 4; ++r) accP[kb][r] = 0;
  }
#pragma unroll
  for (int t = 0; t < NQ; ++t) accPV[t] = zacc;
  for (int c = 1; c <= j.nt; ++c) {
    T qc[4 * NQ], kc[NKB][4 * NQ], vc[NKB][4 * NQ];
    load(in0 + c * chs_in, qc, kc, vc);
    T sa[NKB][4];
#pragma unroll
    for (int kb = 0; kb < NKB; ++kb) {
      acc_t a = zacc;
#pragma unroll
      for (int s = 0; s < 4 * NQ; ++s) {
        a = Mma16<T>::mma(k0[kb][s], qc[s], a);
        a = Mma16<T>::mma(kc[kb][s], q0[s], a);
        accS[kb] = Mma16<T>::mma(kc[kb][s], qc[s], accS[kb]);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) sa[kb][r] = a[r] * sc;
    }
    T m = 0;
#pragma unroll
    for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
      for (int r = 0; r < 4; ++r) m += P[kb][r] * sa[kb][r];
    m = xsum16(m);
    T u[NKB][4], m2 = 0;
#pragma unroll
    for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        u[kb][r] = sa[kb][r] - m;
        m2 += P[kb][r] * u[kb][r] * u[kb][r];
      }
    m2 = xsum16(m2);
    T Pc[NKB][4];
#pragma unroll
    for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        accP[kb][r] += u[kb][r] * u[kb][r] - m2;
        Pc[kb][r] = P[kb][r] * u[kb][r];
      }
    acc_t o[NQ];
#pragma unroll
    for (int t = 0; t < NQ; ++t) {
      o[t] = zacc;
#pragma unroll
      for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
          o[t] = Mma16<T>::mma(Pc[kb][s4], v0[kb][4 * t + s4], o[t]);
          o[t] = Mma16<T>::mma(P[kb][s4], vc[kb][4 * t + s4], o[t]);
          accPV[t] = Mma16<T>::mma(Pc[kb][s4], vc[kb][4 * t + s4], accPV[t]);
        }
    }
    store(out0 + c * chs_out, o, (T)0, zq);
  }
  if (!j.has_lap) return;
  // ---- Laplacian channel
  const int cl = j.nt + 1;
  T ql[4 * NQ], kl[NKB][4 * NQ], vl[NKB][4 * NQ];
  load(in0 + cl * chs_in, ql, kl, vl);
  T w[NKB][4], ml = 0;
#pragma unroll
  for (int kb = 0; kb < NKB; ++kb) {
    acc_t a = zacc;
#pragma unroll
    for (int s = 0; s < 4 * NQ; ++s) {
      a = Mma16<T>::mma(k0[kb][s], ql[s], a);
      a = Mma16<T>::mma(kl[kb][s], q0[s], a);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      w[kb][r] = (a[r] + (T)2 * accS[kb][r]) * sc;
      ml += P[kb][r] * w[kb][r];
    }
  }
  ml = xsum16(ml);
  T Pl[NKB][4];
#pragma unroll
  for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
    for (int r = 0; r < 4; ++r) Pl[kb][r] = P[kb][r] * (w[kb][r] - ml + accP[kb][r]);
  acc_t o[NQ];
#pragma unroll
  for (int t = 0; t < NQ; ++t) {
    o[t] = zacc;
#pragma unroll
    for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        o[t] = Mma16<T>::mma(Pl[kb][s4], v0[kb][4 * t + s4], o[t]);
        o[t] = Mma16<T>::mma(P[kb][s4], vl[kb][4 * t + s4], o[t]);
      }
  }
  store(out0 + cl * chs_out, o, (T)2, accPV);
}

template <typename T>
int launch_attention_jet(const JetGeom& j, int d, int n_heads, const T* qkv, T* out, hipStream_t st) {
  const int64_t npairs = j.n_cfg * n_heads;
  if (npairs == 0) return NQS_OK;
  const int N = j.N, hd = d / n_heads;
  if (N <= 16 && hd <= 32) {
    // launch shape (waves per workgroup) chosen by the autotuner per shape class
    static const char* const names[4] = {"1wave", "2waves", "4waves", "8waves"};
    const TuneKey key{TUNE_ATTENTION, (int64_t)sizeof(T) * 1000 + N * 64 + hd, (int64_t)j.nt * 2 + j.has_lap,
                      tune_bucket(npairs)};
    return tune_run(key, 4, 2, names, [&](int c) -> int {
      const int wpg = 1 << c;
      const dim3 grid((unsigned)cdiv(npairs, wpg)), block(64 * wpg);
      if (hd <= 16)
        hipLaunchKernelGGL((k_attention_jet_mfma<T, 1>), grid, block, 0, st, j, d, n_heads, qkv, out, npairs);
      else
        hipLaunchKernelGGL((k_attention_jet_mfma<T, 2>), grid, block, 0, st, j, d, n_heads, qkv, out, npairs);
      NQS_LAUNCH_CHECK();
      return NQS_OK;
    }, st);
  }
  if (N <= 32 && (hd == 16 || (hd == 32 && sizeof(T) == 4))) {  // float64 with hd 32 would spill
    const int64_t nwork = npairs * 2;
    const int wpg = 4;
    if (hd == 16)
      hipLaunchKernelGGL((k_attention_jet_mfma32<T, 1>), dim3((unsigned)cdiv(nwork, wpg)), dim3(64 * wpg), 0, st, j, d,
                         n_heads, qkv, out, nwork);
    else
      hipLaunchKernelGGL((k_attention_jet_mfma32<T, 2>), dim3((unsigned)cdiv(nwork, wpg)), dim3(64 * wpg), 0, st, j, d,
                         n_heads, qkv, out, nwork);
    NQS_LAUNCH_CHECK();
    return NQS_OK;
  }
  const size_t per_wave = sizeof(T) * (size_t)(7 * N * hd + 5 * N * N + 2 * N);
  NQS_REQUIRE(per_wave <= 160 * 1024, NQS_ERR_UNSUPPORTED, "attention tile exceeds LDS");
  int wpg = (int)std::min<size_t>(4, (64 * 1024) / per_wave);
  if (wpg < 1) wpg = 1;
  const size_t lds = per_wave * wpg;
  if (lds > 64 * 1024) {
    static bool attr[2] = {false, false};
    if (!attr[sizeof(T) == 8]) {
      NQS_CHECK_HIP(hipFuncSetAttribute((const void*)k_attention_jet_wave<T>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
      attr[sizeof(T) == 8] = true;
    }
  }
  hipLaunchKernelGGL(k_attention_jet_wave<T>, dim3((unsigned)cdiv(npairs, wpg)), dim3(64 * wpg), lds, st, j, d,
                     n_heads, qkv, out, wpg, npairs);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

#define NQS_INST(T)                                                                                   \
  template int launch_features<T>(const JetGeom&, const double*, const int32_t*, const double*, int, \
                                  int, int64_t, T*, hipStream_t);                                   \
  template int launch_layernorm_jet<T>(const JetGeom&, int, const T*, const T*, const T*, T*,       \
                                       hipStream_t);                                                \
  template int launch_gelu_jet<T>(const JetGeom&, int, const T*, T*, hipStream_t);                  \
  template int launch_attention_jet<T>(const JetGeom&, int, int, const T*, T*, hipStream_t);
NQS_INST(double)
NQS_INST(float)

}  // namespace nqs
