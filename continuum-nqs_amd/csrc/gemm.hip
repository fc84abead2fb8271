// MFMA GEMMs for gfx950.
//
//   gemm_nt<T>:  C[m][n] = alpha * sum_k A[m][k] * B[n][k] + beta * C[m][n] + bias[n]*rowmask(m)
//
// Both operands are K-contiguous ("NT"): activations are row-major (rows x features) and the
// weights are stored transposed (out x in) by the parameter packer, so one kernel serves the
// dense layers (bias on the value-channel rows only), the float64 Gram O O^T and the
// Cholesky trailing update (lower-triangular tiles).  T = double uses v_mfma_f64_16x16x4f64,
// T = float uses v_mfma_f32_16x16x4f32; both take A[lane%16][lane/16], B[lane/16][lane%16]
// and differ in the accumulator row map: f64 row = lane/16 + 4 r, f32 row = 4 (lane/16) + r.
#include "nqs_common.h"
#include "kernels.h"

namespace nqs {

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

// Block tile 64 x 64, 4 waves in a 2 x 2 arrangement, 32 x 32 per wave (2 x 2 MFMA blocks).
template <typename T, int BK>
__global__ __launch_bounds__(256) void k_gemm_nt(GemmArgs<T> g) {
  constexpr int BM = 64, BN = 64, PAD = 1;
  __shared__ T As[BM][BK + PAD];
  __shared__ T Bs[BN][BK + PAD];
  const int bm = blockIdx.y, bn = blockIdx.x, bz = blockIdx.z;
  if (g.tri == 1 && bn > bm) return;  // lower-triangular tiles only
  const T* A = g.A + (int64_t)bz * g.strideA;
  const T* B = g.B + (int64_t)bz * g.strideB;
  T* C = g.C + (int64_t)bz * g.strideC;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int64_t m0 = (int64_t)bm * BM, n0 = (int64_t)bn * BN;
  typedef typename Mma16<T>::acc_t acc_t;
  acc_t acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int r = 0; r < 4; ++r) acc[i][j][r] = 0;

  constexpr int LPR = BK;                 // loads per row
  constexpr int ROWS_PER_PASS = 256 / LPR;
  for (int64_t k0 = 0; k0 < g.K; k0 += BK) {
    for (int rr = tid / LPR; rr < BM; rr += ROWS_PER_PASS) {
      const int kk = tid % LPR;
      const int64_t gm = m0 + rr, gk = k0 + kk;
      As[rr][kk] = (gm < g.M && gk < g.K) ? A[gm * g.lda + gk] : (T)0;
      const int64_t gn = n0 + rr;
      Bs[rr][kk] = (gn < g.N && gk < g.K) ? B[gn * g.ldb + gk] : (T)0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[wm * 32 + i * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[wn * 32 + j * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = Mma16<T>::mma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm * 32 + i * 16 + Mma16<T>::row(lane, r);
        const int64_t col = n0 + wn * 32 + j * 16 + (lane & 15);
        if (row >= g.M || col >= g.N) continue;
        if (g.tri == 1 && col > row) continue;
        T v = g.alpha * acc[i][j][r];
        if (g.beta != (T)0) v += g.beta * C[row * g.ldc + col];
        if (g.bias && (row % g.bias_period) < g.bias_len) v += g.bias[col];
        C[row * g.ldc + col] = v;
        if (g.mirror && row != col) C[col * g.ldc + row] = v;
      }
}

// v2: block tile 128 x 128 x BK, 4 waves in 2 x 2, 64 x 64 per wave (4 x 4 MFMA blocks), operands
// moved in 16-byte vectors through registers into a double-buffered LDS whose row pitch
// (BK + 16 bytes) keeps the 16 rows a fragment read touches on distinct banks.
// Requires lda, ldb multiples of 16 bytes and 16-byte aligned A, B; rows/columns beyond M, N
// and k beyond K are zero-filled in the tile.
template <typename T>
struct Vec16;
template <>
struct Vec16<double> {
  typedef double t __attribute__((ext_vector_type(2)));
  static constexpr int n = 2;
};
template <>
struct Vec16<float> {
  typedef float t __attribute__((ext_vector_type(4)));
  static constexpr int n = 4;
};

template <typename T, int BK>
__global__ __launch_bounds__(256) void k_gemm_nt_v2(GemmArgs<T> g) {
  constexpr int BM = 128, VN = Vec16<T>::n, PAD = VN, LD = BK + PAD;
  constexpr int CH = BK / VN;         // 16-byte chunks per tile row
  constexpr int PER = 128 * CH / 256; // chunks per thread per operand
  typedef typename Vec16<T>::t vt;
  __shared__ T As[2][BM][LD];
  __shared__ T Bs[2][BM][LD];
  const int bm = blockIdx.y, bn = blockIdx.x, bz = blockIdx.z;
  if (g.tri == 1 && bn > bm) return;
  const T* A = g.A + (int64_t)bz * g.strideA;
  const T* B = g.B + (int64_t)bz * g.strideB;
  T* C = g.C + (int64_t)bz * g.strideC;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int64_t m0 = (int64_t)bm * BM, n0 = (int64_t)bn * BM;
  typedef typename Mma16<T>::acc_t acc_t;
  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = 0;
  // Branch-free staging: addresses are clamped into the matrices and out-of-range vectors are
  // replaced by zeros with a select, so the loads of a stage issue back to back and the
  // staging arrays stay in registers.  K is a multiple of VN (checked by the launcher).
  vt ra[PER], rb[PER];
  const vt zero = (vt)(T)0;
  const int64_t nk = cdiv_dev(g.K, BK);
  // raw loads here; the out-of-range zeroing happens at the LDS store, so no wait is forced
  // between issuing a stage's loads and the MFMAs of the current stage
#define NQS_GEMM_LOAD(k0)                                                                         \
  _Pragma("unroll") for (int q = 0; q < PER; ++q) {                                               \
    const int idx = tid + 256 * q, row = idx / CH, ch = idx % CH;                                 \
    const int64_t gk = (k0) + ch * VN;                                                            \
    const int64_t gm = m0 + row, gn = n0 + row;                                                   \
    const int64_t ck = gk < g.K ? gk : g.K - VN;                                                  \
    ra[q] = *reinterpret_cast<const vt*>(A + (gm < g.M ? gm : g.M - 1) * g.lda + ck);             \
    rb[q] = *reinterpret_cast<const vt*>(B + (gn < g.N ? gn : g.N - 1) * g.ldb + ck);             \
  }
#define NQS_GEMM_STORE(buf, k0)                                                                   \
  _Pragma("unroll") for (int q = 0; q < PER; ++q) {                                               \
    const int idx = tid + 256 * q, row = idx / CH, ch = idx % CH;                                 \
    const bool okk = (k0) + ch * VN < g.K;                                                        \
    *reinterpret_cast<vt*>(&As[buf][row][ch * VN]) = (okk && m0 + row < g.M) ? ra[q] : zero;      \
    *reinterpret_cast<vt*>(&Bs[buf][row][ch * VN]) = (okk && n0 + row < g.N) ? rb[q] : zero;      \
  }
  NQS_GEMM_LOAD(0)
  NQS_GEMM_STORE(0, 0)
  __syncthreads();
  for (int64_t kb = 0; kb < nk; ++kb) {
    const int cur = kb & 1;
    const int64_t knext = (kb + 1 < nk ? kb + 1 : kb) * BK;
    NQS_GEMM_LOAD(knext)
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[cur][wm * 64 + i * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[cur][wn * 64 + j * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = Mma16<T>::mma(a[i], b[j], acc[i][j]);
    }
    NQS_GEMM_STORE(cur ^ 1, knext)
    __syncthreads();
  }
#undef NQS_GEMM_LOAD
#undef NQS_GEMM_STORE
  // Epilogue in two phases: every C / bias operand is read first (clamped addresses, uniform
  // branches only), then the tile is combined and stored under the per-element predicate.
  const bool has_beta = g.beta != (T)0, has_bias = g.bias != nullptr;
  T addv[4][4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) addv[i][j][r] = 0;
  if (has_beta) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t row = m0 + wm * 64 + i * 16 + Mma16<T>::row(lane, r);
          const int64_t col = n0 + wn * 64 + j * 16 + (lane & 15);
          const int64_t rr = row < g.M ? row : g.M - 1, cc = col < g.N ? col : g.N - 1;
          addv[i][j][r] = g.beta * C[rr * g.ldc + cc];
        }
  }
  if (has_bias) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t col = n0 + wn * 64 + j * 16 + (lane & 15);
      const T bv = g.bias[col < g.N ? col : g.N - 1];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const uint32_t row = (uint32_t)(m0 + wm * 64 + i * 16 + Mma16<T>::row(lane, r));
          addv[i][j][r] += (row % (uint32_t)g.bias_period < (uint32_t)g.bias_len) ? bv : (T)0;
        }
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm * 64 + i * 16 + Mma16<T>::row(lane, r);
        const int64_t col = n0 + wn * 64 + j * 16 + (lane & 15);
        const bool in = row < g.M && col < g.N && !(g.tri == 1 && col > row);
        const T v = g.alpha * acc[i][j][r] + addv[i][j][r];
        if (in) {
          C[row * g.ldc + col] = v;
          if (g.mirror && row != col) C[col * g.ldc + row] = v;
        }
      }
}

template <typename T>
int gemm_nt(const GemmArgs<T>& g, int batch, hipStream_t st) {
  if (g.M <= 0 || g.N <= 0) return NQS_OK;
  constexpr int VN = Vec16<T>::n;
  const bool vec_ok = (g.K % VN == 0) && (g.lda % VN == 0) && (g.ldb % VN == 0) && (((uintptr_t)g.A & 15) == 0) &&
                      (((uintptr_t)g.B & 15) == 0) && (g.strideA % VN == 0) && (g.strideB % VN == 0);
  if (vec_ok) {
    dim3 grid((unsigned)cdiv(g.N, 128), (unsigned)cdiv(g.M, 128), (unsigned)batch);
    if (sizeof(T) == 8)
      hipLaunchKernelGGL((k_gemm_nt_v2<T, 16>), grid, dim3(256), 0, st, g);
    else
      hipLaunchKernelGGL((k_gemm_nt_v2<T, 32>), grid, dim3(256), 0, st, g);
    NQS_LAUNCH_CHECK();
    return NQS_OK;
  }
  dim3 grid((unsigned)cdiv(g.N, 64), (unsigned)cdiv(g.M, 64), (unsigned)batch);
  if (sizeof(T) == 8)
    hipLaunchKernelGGL((k_gemm_nt<T, 16>), grid, dim3(256), 0, st, g);
  else
    hipLaunchKernelGGL((k_gemm_nt<T, 32>), grid, dim3(256), 0, st, g);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

template int gemm_nt<double>(const GemmArgs<double>&, int, hipStream_t);
template int gemm_nt<float>(const GemmArgs<float>&, int, hipStream_t);

}  // namespace nqs
