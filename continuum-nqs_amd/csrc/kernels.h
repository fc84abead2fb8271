// Internal launcher declarations shared by the kernel files and the C ABI layer.
#pragma once
#include "nqs_common.h"

namespace nqs {

// v_mfma_*_16x16x4 wrappers: A[lane%16][lane/16], B[lane/16][lane%16]; accumulator slot r of
// a lane holds C[row(lane, r)][lane%16], with row = lane/16 + 4 r (f64) or 4 (lane/16) + r (f32).
template <typename T>
struct Mma16;
template <>
struct Mma16<double> {
  typedef d4v acc_t;
  static __device__ inline acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mma16<float> {
  typedef f4v acc_t;
  static __device__ inline acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// float32 pair -> bf16 hi = bf16(x) and lo = bf16(x - hi), both round-to-nearest-even, packed two
// per 32-bit word (element 0 in the low half): the operand split of the bf16x3 products.
typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));

// No FMA contraction: x - hi is never fused with a product that computed x, whichever kernel inlines
// the split, so every caller gets the same lo.
__device__ inline void split_bf16x2(float x0, float x1, uint32_t& hi, uint32_t& lo) {
#pragma clang fp contract(off)
  const f2v v = {x0, x1};
  hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf2v));
  const float h0 = __builtin_bit_cast(float, hi << 16), h1 = __builtin_bit_cast(float, hi & 0xffff0000u);
  const f2v l = {x0 - h0, x1 - h1};
  lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(l, bf2v));
}

// tanh-GELU (jax.nn.gelu, approximate=True) and its first two derivatives.  No FMA contraction:
// every caller gets the same roundings whichever of the outputs it uses (the fused forward and the
// layer-by-layer kernels agree bit for bit).
// tanh for the GELU jets: float64 uses the library tanh; float32 uses 1 - 2 / (1 + e^{2u}) with the
// hardware exp2 and reciprocal (absolute error a few 1e-7, below the bf16x3 products' rounding;
// +-1 exactly once e^{2u} over- or underflows).
template <typename T>
__device__ inline T gelu_tanh(T u) {
  return tanh(u);
}
template <>
__device__ inline float gelu_tanh<float>(float u) {
#pragma clang fp contract(off)
  const float e = __builtin_amdgcn_exp2f(u * 2.8853900817779268f);  // 2 log2(e)
  return 1.f - 2.f * __builtin_amdgcn_rcpf(1.f + e);
}

template <typename T>
__device__ inline void gelu_derivs(T x, T& g, T& g1, T& g2) {
#pragma clang fp contract(off)
  const T k0 = (T)0.7978845608028654, k1 = (T)0.044715;
  const T u = k0 * (x + k1 * x * x * x);
  const T th = gelu_tanh<T>(u);
  const T sech2 = (T)1 - th * th;
  const T du = k0 * ((T)1 + (T)3 * k1 * x * x);
  const T ddu = (T)6 * k0 * k1 * x;
  g = (T)0.5 * x * ((T)1 + th);
  g1 = (T)0.5 * ((T)1 + th) + (T)0.5 * x * sech2 * du;
  g2 = sech2 * du + (T)0.5 * x * (sech2 * ddu - (T)2 * sech2 * th * du * du);
}

template <typename T>
struct GemmArgs {
  const T* A;  // M x K, row stride lda
  const T* B;  // N x K, row stride ldb
  T* C;        // M x N, row stride ldc
  int64_t M, N, K, lda, ldb, ldc;
  int64_t strideA, strideB, strideC;  // batch strides (gridDim.z)
  T alpha, beta;
  const T* bias;        // N or NULL
  int64_t bias_period;  // bias added to rows with (row % bias_period) < bias_len
  int64_t bias_len;
  int tri;     // 0 full, 1 lower-triangular tiles only (col <= row)
  // set by the launcher: 1 = a square lower triangle (tri, M == N) on a 1-D grid in super-tile
  // order, each XCD taking a contiguous stretch of it, so that an XCD's concurrent tiles share
  // row and column panels in its L2
  int tile_order = 0;
  int mirror;  // also store C[col][row]
  int split;   // float only: 0 float32 MFMA, 1 bf16 MFMA, 3 bf16x3 (hi/lo split) MFMA
  // Optional M x ldc buffer the tile-shape autotuner may overwrite: with beta != 0 the trial
  // launches write A B there (beta = 0) instead of accumulating into C.  NULL: a beta != 0 call
  // uses the tile shape already chosen for its shape class (or the default) without trials.
  T* tune_scratch = nullptr;
  // Optional first-order GELU jet applied to A as it is staged (bf16 float kernels only), for jets
  // without a Laplacian channel laid out [cfg][channel][token][feature] with gelu_cn = channels x
  // tokens rows per configuration and gelu_n tokens: A holds the pre-activation u; the value rows
  // (channel 0) become gelu(u), a tangent row du becomes gelu'(u of its token's value row) du.
  // 0 = no prologue.
  int gelu_cn = 0, gelu_n = 0;
  // float64 only (the Cholesky factorisation): when potrf_linv is set, the workgroup of tile (0, 0)
  // factors the 64 x 64 block at C's origin after storing its tile (potrf_diag64: the block is final
  // once this product is added), writing the block's inverse to potrf_linv + z potrf_sl and its first
  // non-positive pivot to potrf_info[z] (potrf_j0 = the block's column in the factored matrix)
  double* potrf_linv = nullptr;
  int* potrf_info = nullptr;
  int64_t potrf_j0 = 0, potrf_sl = 0;
};

template <typename T>
int gemm_nt(const GemmArgs<T>& g, int batch, hipStream_t st);

__device__ __forceinline__ double readlane_f64(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

// One wave (lane i = threadIdx.x % 64) factors the 64 x 64 diagonal block at A (row stride lda) in
// place (lower; the upper triangle is set to zero) and writes its inverse (lower, row-major) to Linv.
// Lane i keeps row i in registers; column k's multipliers reach the other lanes by v_readlane, so no
// step needs a barrier.  The registers of the upper triangle (k > i) are updated along with the rest
// but never read by another lane nor stored, so the steps need no masks.  The inverse is formed by
// forward substitution, column c in lane c, reading L and the reciprocal pivots from the LDS scratch
// Ls (64 x 65 doubles) and rdiag (64 doubles), the same address on every lane.  info[0] = jg + first
// non-positive pivot + 1 (jg: the block's column in the factored matrix; the pivot is then taken as
// 1 and the factor goes on).
__device__ __forceinline__ void potrf_diag64(double* __restrict__ A, int64_t lda, int64_t jg, double* __restrict__ Linv,
                                             int* __restrict__ info, double (*Ls)[65], double* rdiag) {
  constexpr int NB = 64;
  const int i = threadIdx.x & 63;
  double a[NB];
  double* row = A + (int64_t)i * lda;
#pragma unroll
  for (int k = 0; k < NB; ++k) a[k] = row[k];
  int bad = 0;
  double myinv = 0.0;
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    double d = readlane_f64(a[k], k);
    if (!(d > 0.0)) {
      if (!bad) bad = k + 1;
      d = 1.0;
    }
    const double lkk = sqrt(d), inv = 1.0 / lkk;
    a[k] = i == k ? lkk : a[k] * inv;
    myinv = i == k ? inv : myinv;
#pragma unroll
    for (int j = k + 1; j < NB; ++j) a[j] -= a[k] * readlane_f64(a[k], j);
  }
  if (bad && i == 0 && info[0] == 0) info[0] = (int)(jg + bad);
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    const double v = k <= i ? a[k] : 0.0;
    row[k] = v;
    Ls[i][k] = v;
  }
  rdiag[i] = myinv;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const int c = i;
  double x[NB];
#pragma unroll
  for (int r = 0; r < NB; ++r) {
    double s0 = r == c ? 1.0 : 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
    for (int k = 0; k + 3 < r; k += 4) {
      s0 -= Ls[r][k] * x[k];
      s1 -= Ls[r][k + 1] * x[k + 1];
      s2 -= Ls[r][k + 2] * x[k + 2];
      s3 -= Ls[r][k + 3] * x[k + 3];
    }
#pragma unroll
    for (int k = r & ~3; k < r; ++k) s0 -= Ls[r][k] * x[k];
    x[r] = ((s0 + s1) + (s2 + s3)) * rdiag[r];
  }
#pragma unroll
  for (int r = 0; r < NB; ++r) Linv[r * NB + c] = x[r];
}
// The same as a called function, for kernels whose own register budget it would otherwise set.
__device__ __attribute__((noinline)) void potrf_diag64_call(double* __restrict__ A, int64_t lda, int64_t jg,
                                                            double* __restrict__ Linv, int* __restrict__ info,
                                                            double (*Ls)[65], double* rdiag);

// Jet geometry: activations are laid out [cfg][channel][token][feature]; channel 0 is the
// value, channels 1..nt the tangents, channel nt+1 the Laplacian when has_lap.
struct JetGeom {
  int64_t n_cfg;  // configurations in this launch
  int N;          // tokens (electrons)
  int C;          // channels = 1 + nt + has_lap
  int nt;         // tangent channels
  int has_lap;
  int mode;       // 0 value, 1 full Laplacian (nt = 2N), 2 spin-flip (cfg = w*N + i, nt = 2)
};

template <typename T>
int launch_features(const JetGeom& j, const double* R, const int32_t* S, const double* Ainv,
                    int n_freq, int kpad, int64_t w0, T* X, hipStream_t st);
template <typename T>
int launch_layernorm_jet(const JetGeom& j, int d, const T* x, const T* scale, const T* bias, T* y,
                         hipStream_t st);
template <typename T>
int launch_gelu_jet(const JetGeom& j, int width, const T* x, T* y, hipStream_t st);
template <typename T>
int launch_attention_jet(const JetGeom& j, int d, int n_heads, const T* qkv, T* out,
                         hipStream_t st);

// Determinant head: envelope, K spinor determinants, log-sum-exp, Jastrow.  Outputs per cfg:
// logpsi (complex), tangents (nt complex), Laplacian (complex, when has_lap).
struct DetParams {
  int N, n_g, K, use_jastrow, jastrow_kind, n_knots;
  double A[4], Ainv[4];
  double cusp, rbf_rcut, rbf_w;
  const double* gs;       // (n_g, 2)
  const double* wenv;     // (N, n_g, 2) complex interleaved (includes dW)
  const double* det_logw; // (K) complex or NULL
  const double* jas;      // jastrow params: rbf [c_same(n_knots), c_opp(n_knots)]; cusp [loga]
};
// Sum over the 16 lanes of a DPP row (lanes 16 g .. 16 g + 15), the same bits in each of them:
// the first four steps of wave_sum.
template <typename T>
__device__ __forceinline__ T row_sum16(T v) {
  v += row_dpp<0xB1>(v);
  v += row_dpp<0x4E>(v);
  v += row_dpp<0x141>(v);
  v += row_dpp<0x140>(v);
  return v;
}

// LayerNorm rows on 16-lane groups: the row's d features are spread over the 16 lanes of one DPP
// row, lane j (= lane % 16) holding features F j .. F j + F - 1 in xv (zero for f >= d), so a wave
// normalises four rows at a time.  Sums run over a lane's F values in order, then over the 16
// lanes (row_sum16); no FMA contraction, so every kernel that inlines these rounds alike.
//
// Value row (flax: mean, fast variance E[x^2] - mu^2, eps 1e-6): the scales sc, the outputs
// y = (x - mu) rs scale + bias (0 for f >= d), mu and rs = (var + eps)^-1/2.
template <typename T, int F>
__device__ __forceinline__ void ln_value(const T (&xv)[F], int j, int d, const T* __restrict__ scale,
                                         const T* __restrict__ bias, T (&sc)[F], T (&y)[F], T& mu, T& rs) {
#pragma clang fp contract(off)
  const T inv_d = (T)1 / (T)d;
  T sum = 0, sum2 = 0;
#pragma unroll
  for (int e = 0; e < F; ++e) {
    const int f = F * j + e;
    sc[e] = (f < d) ? scale[f] : (T)0;
    sum += xv[e];
    sum2 += xv[e] * xv[e];
  }
  sum = row_sum16(sum);
  sum2 = row_sum16(sum2);
  mu = sum * inv_d;
  T var = sum2 * inv_d - mu * mu;
  var = var > (T)0 ? var : (T)0;
  rs = (T)1 / sqrt(var + (T)1e-6);
#pragma unroll
  for (int e = 0; e < F; ++e) {
    const int f = F * j + e;
    y[e] = (f < d) ? (xv[e] - mu) * rs * sc[e] + bias[f] : (T)0;
  }
}

// First-order tangent row, given the value row's xv, mu, rs and scales sc (from ln_value) and the
// tangent dx: dmu = mean(dx), drs = -rs^3 (mean(x dx) - mu dmu),
// dy = ((dx - dmu) rs + (x - mu) drs) scale.
template <typename T, int F>
__device__ __forceinline__ void ln_tangent(const T (&xv)[F], const T (&dx)[F], const T (&sc)[F], int d, T mu, T rs,
                                           T (&dy)[F]) {
#pragma clang fp contract(off)
  const T inv_d = (T)1 / (T)d;
  T sd = 0, sxd = 0;
#pragma unroll
  for (int e = 0; e < F; ++e) {
    sd += dx[e];
    sxd += xv[e] * dx[e];
  }
  sd = row_sum16(sd);
  sxd = row_sum16(sxd);
  const T dmu = sd * inv_d;
  const T dvar = (T)2 * (sxd * inv_d - mu * dmu);
  const T drs = (T)(-0.5) * rs * rs * rs * dvar;
#pragma unroll
  for (int e = 0; e < F; ++e) dy[e] = ((dx[e] - dmu) * rs + (xv[e] - mu) * drs) * sc[e];
}

// Lower-triangular tile grids in super-tile order: tiles grouped in SUPER x SUPER super-tiles (lower
// triangle of super-tiles, row-major; a diagonal super-tile keeps its lower triangle), row-major
// inside a super-tile.  Position t of the sequence -> tile (bm, bn) of the padded grid; positions
// whose tile lies outside the lower triangle of the nt x nt grid are skipped by the caller.
constexpr int SUPER = 8;
__device__ inline void supertile_of(int64_t t, int& bm, int& bn) {
  // super row I holds I full super-tiles (SUPER^2 tiles each) and one diagonal one
  // (SUPER (SUPER + 1) / 2 tiles): prefix(I) = SUPER^2 I (I - 1) / 2 + SUPER (SUPER + 1) / 2 I
  constexpr int64_t F = SUPER * SUPER, D = SUPER * (SUPER + 1) / 2;
  int I = 0;
  while (F * (int64_t)(I + 1) * I / 2 + D * (I + 1) <= t) ++I;
  const int64_t o = t - (F * (int64_t)I * (I - 1) / 2 + D * I);
  if (o < F * I) {
    const int J = (int)(o / F), u = (int)(o % F);
    bm = SUPER * I + u / SUPER;
    bn = SUPER * J + u % SUPER;
  } else {
    const int u = (int)(o - F * I);  // lower triangle of the diagonal super-tile, row-major
    int r = 0;
    while ((r + 1) * (r + 2) / 2 <= u) ++r;
    bm = SUPER * I + r;
    bn = SUPER * I + (u - r * (r + 1) / 2);
  }
}
__host__ __device__ inline int64_t supertile_count(int64_t nt) {
  const int64_t ns = (nt + SUPER - 1) / SUPER;
  return (int64_t)SUPER * SUPER * ns * (ns - 1) / 2 + (int64_t)SUPER * (SUPER + 1) / 2 * ns;
}

template <typename T>
// scratch (may be NULL): device memory the head may use for per-determinant results, at least
// n_cfg * K * (nt + 2) * 16 bytes for the value-only and spin-flip passes' row-per-lane kernels.
int launch_det_jet(const JetGeom& j, const DetParams& p, const double* R, const int32_t* S,
                   int64_t w0, const T* bf, double* out_logpsi, double* out_tan, double* out_lap,
                   void* scratch, size_t scratch_bytes, hipStream_t st);

int launch_ewald(const nqs_ewald_desc& e, const double* R, int64_t W, int N, double* out,
                 hipStream_t st);

}  // namespace nqs
