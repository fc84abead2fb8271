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

template <typename T>
int gemm_nt(const GemmArgs<T>& g, int batch, hipStream_t st) {
  if (g.M <= 0 || g.N <= 0) return NQS_OK;
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
