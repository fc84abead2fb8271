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
#include "tune.h"


namespace nqs {

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
__global__ __launch_bounds__(256, 2) void k_gemm_nt_v2(GemmArgs<T> g) {
  constexpr int BM = 128, VN = Vec16<T>::n, PAD = VN, LD = BK + PAD;
  constexpr int CH = BK / VN;         // 16-byte chunks per tile row
  constexpr int PER = 128 * CH / 256; // chunks per thread per operand
  typedef typename Vec16<T>::t vt;
  __shared__ T As[2][BM][LD];
  __shared__ T Bs[2][BM][LD];
  int bm = blockIdx.y, bn = blockIdx.x;
  const int bz = blockIdx.z;
  if (g.tile_order == 1) {
    const int64_t nt = cdiv_dev(g.M, (int64_t)BM), ntiles = supertile_count(nt), per_xcd = (ntiles + 7) / 8;
    const int64_t t = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    if (t >= ntiles) return;
    supertile_of(t, bm, bn);
    if (bm >= nt) return;
  }
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
  // Epilogue in two phases per 16-row block i (so that only 16 operands are live at a time):
  // every C / bias operand of the block is read first (clamped addresses, uniform branches only),
  // then the block is combined and stored under the per-element predicate.
  const bool has_beta = g.beta != (T)0, has_bias = g.bias != nullptr;
  T bv[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int64_t col = n0 + wn * 64 + j * 16 + (lane & 15);
    bv[j] = has_bias ? g.bias[col < g.N ? col : g.N - 1] : (T)0;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    T addv[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) addv[j][r] = 0;
    if (has_beta) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t row = m0 + wm * 64 + i * 16 + Mma16<T>::row(lane, r);
          const int64_t col = n0 + wn * 64 + j * 16 + (lane & 15);
          const int64_t rr = row < g.M ? row : g.M - 1, cc = col < g.N ? col : g.N - 1;
          addv[j][r] = g.beta * C[rr * g.ldc + cc];
        }
    }
    if (has_bias) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const uint32_t row = (uint32_t)(m0 + wm * 64 + i * 16 + Mma16<T>::row(lane, r));
          addv[j][r] += (row % (uint32_t)g.bias_period < (uint32_t)g.bias_len) ? bv[j] : (T)0;
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm * 64 + i * 16 + Mma16<T>::row(lane, r);
        const int64_t col = n0 + wn * 64 + j * 16 + (lane & 15);
        const bool in = row < g.M && col < g.N && !(g.tri == 1 && col > row);
        const T v = g.alpha * acc[i][j][r] + addv[j][r];
        if (in) {
          C[row * g.ldc + col] = v;
          if (g.mirror && row != col) C[col * g.ldc + row] = v;
        }
      }
  }
  if constexpr (sizeof(T) == 8) {
    if (g.potrf_linv != nullptr && bm == 0 && bn == 0) {
      // the block at C's origin is this tile's and final: every wave's stores complete and are
      // visible to wave 0 (the fence on both sides of the barrier), which factors it with the
      // staging LDS as scratch
      __threadfence();
      __syncthreads();
      if (wave == 0) {
        __threadfence();
        potrf_diag64_call(reinterpret_cast<double*>(C), g.ldc, g.potrf_j0, g.potrf_linv + (int64_t)bz * g.potrf_sl,
                     g.potrf_info + bz, reinterpret_cast<double (*)[65]>(&As[0][0][0]),
                     reinterpret_cast<double*>(&Bs[0][0][0]));
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// float32 operands on bf16 MFMA.  Each float32 x is split into hi = bf16(x) and lo = bf16(x - hi)
// (both round-to-nearest-even) as it is staged into LDS; SPLIT = 3 accumulates
// hi*hi + hi*lo + lo*hi in float32 (products carry about 16 significant bits; the dropped lo*lo
// term is below 2^-16 relative), SPLIT = 1 keeps hi*hi only (plain bf16).  Block tile
// 128 x 128 x 32, 4 waves in 2 x 2, 64 x 64 per wave as 2 x 2 blocks of
// v_mfma_f32_32x32x16_bf16 (A[lane%32][8 (lane/32) + e], C row 8 (r/4) + 4 (lane/32) + r%4,
// col lane%32).  Tiles are dealt XCD-aware: workgroup b runs on XCD b % 8, and the column
// tiles of one row panel go to the same XCD back to back, so the panel is read from HBM once.

constexpr int BF_BK = 32, BF_LD = BF_BK + 8;
// staging geometry: 8-float chunks per tile row, rows per pass of the 256 threads
constexpr int BF_CPR = BF_BK / 8, BF_RPP = 256 / BF_CPR;

// Tile shapes: 4 waves in WM x WN, each wave BI x BJ blocks of 32 x 32, so the workgroup tile is
// (32 WM BI) x (32 WN BJ).  Every shape accumulates each output element over K in the same order
// (16-wide k chunks, lo*hi, hi*lo, hi*hi per chunk) with the same epilogue, so all shapes give
// bit-identical results and the autotuner may pick any of them per shape class.
template <int WM, int WN, int BI, int BJ>
struct BfTile {
  static constexpr int BM = 32 * WM * BI, BN = 32 * WN * BJ;
  static constexpr int NQA = BM / BF_RPP, NQB = BN / BF_RPP;
  static_assert(WM * WN == 4 && NQA >= 1 && NQB >= 1, "four waves, tiles of at least 64 rows");
};

template <int SPLIT, int BM, int BN>
constexpr size_t bf_lds_bytes() {
  return sizeof(__bf16) * (size_t)(SPLIT == 1 ? 1 : 2) * (BM + BN) * BF_LD;
}

// Staging: a thread moves rows tid / 4 + 64 q, k-chunk tid % 4 (8 floats, two 16-byte loads) and
// writes its 8 hi and 8 lo bf16 as 16-byte LDS stores; 4 lanes fill one 64-byte row, stored at an
// 80-byte row pitch.  The epilogue keeps one 32-row block of addends live at a time.  One LDS
// buffer per operand and plane; the registers carry stage kb + 1 across the loop back edge
// (loaded at the end of iteration kb - 1), so global latency overlaps a stage of MFMA work.
// PRO = 1: the first-order GELU jet prologue on A (GemmArgs::gelu_cn): each staged A row also loads
// its token's value row (the same row for the value channel) and the pair is turned into the GELU
// jet before the bf16 split.
template <int SPLIT, int WM, int WN, int BI, int BJ, int MINB, int PRO = 0>
__global__ __launch_bounds__(256, MINB) void k_gemm_nt_bf(GemmArgs<float> g, int nbn, int64_t ntiles) {
  typedef BfTile<WM, WN, BI, BJ> TL;
  constexpr int BM = TL::BM, BN = TL::BN, NQA = TL::NQA, NQB = TL::NQB;
  constexpr int NP = SPLIT == 1 ? 1 : 2;
  constexpr int APL = BM * BF_LD, BPL = BN * BF_LD;  // A planes (hi, lo), then B planes
  extern __shared__ __bf16 bfs[];
  __bf16* As = bfs;
  __bf16* Bs = bfs + NP * APL;
  const int64_t per_xcd = (ntiles + 7) / 8;
  const int64_t t = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (t >= ntiles) return;
  const int64_t bm = t / nbn;
  const int bn = (int)(t % nbn);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const int64_t m0 = bm * BM, n0 = (int64_t)bn * BN;
  f16v acc[BI][BJ];
#pragma unroll
  for (int i = 0; i < BI; ++i)
#pragma unroll
    for (int j = 0; j < BJ; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
  const int srow = tid / BF_CPR, sk = (tid % BF_CPR) * 8;
  const float* pa[NQA];
  const float* pb[NQB];
#pragma unroll
  for (int q = 0; q < NQA; ++q) {
    const int64_t gm = m0 + srow + BF_RPP * q;
    pa[q] = g.A + (gm < g.M ? gm : g.M - 1) * g.lda;
  }
#pragma unroll
  for (int q = 0; q < NQB; ++q) {
    const int64_t gn = n0 + srow + BF_RPP * q;
    pb[q] = g.B + (gn < g.N ? gn : g.N - 1) * g.ldb;
  }
  f4v ra[NQA][2], rb[NQB][2];
  // PRO: the value row of each staged A row, and whether the row is a value row itself
  const float* pv[PRO ? NQA : 1];
  bool isval[PRO ? NQA : 1];
  f4v rv[PRO ? NQA : 1][2];
  if constexpr (PRO) {
#pragma unroll
    for (int q = 0; q < NQA; ++q) {
      const int64_t gm0 = m0 + srow + BF_RPP * q, gm = gm0 < g.M ? gm0 : g.M - 1;
      const int64_t c = (gm % g.gelu_cn) / g.gelu_n;
      isval[q] = c == 0;
      pv[q] = g.A + (gm - c * g.gelu_n) * g.lda;
    }
  }
  const f4v zero = (f4v)0.f;
  const int64_t nk = cdiv_dev(g.K, BF_BK);
  // K is a multiple of 4: each 4-wide half of a chunk is either wholly inside K or wholly out
#define NQS_BF_LOAD(kb_)                                                                \
  {                                                                                     \
    const int64_t g0 = (int64_t)(kb_) * BF_BK + sk;                                     \
    const int64_t c0 = g0 < g.K ? g0 : g.K - 4, c1 = g0 + 4 < g.K ? g0 + 4 : g.K - 4;    \
    _Pragma("unroll") for (int q = 0; q < NQA; ++q) {                                   \
      ra[q][0] = *reinterpret_cast<const f4v*>(pa[q] + c0);                             \
      ra[q][1] = *reinterpret_cast<const f4v*>(pa[q] + c1);                             \
      if constexpr (PRO) {                                                              \
        rv[q][0] = *reinterpret_cast<const f4v*>(pv[q] + c0);                           \
        rv[q][1] = *reinterpret_cast<const f4v*>(pv[q] + c1);                           \
      }                                                                                 \
    }                                                                                   \
    _Pragma("unroll") for (int q = 0; q < NQB; ++q) {                                   \
      rb[q][0] = *reinterpret_cast<const f4v*>(pb[q] + c0);                             \
      rb[q][1] = *reinterpret_cast<const f4v*>(pb[q] + c1);                             \
    }                                                                                   \
  }
#define NQS_BF_SPLIT8(x0, x1, hi, lo)                                             \
  {                                                                               \
    uint32_t h_[4], l_[4];                                                        \
    split_bf16x2(x0[0], x0[1], h_[0], l_[0]);                                     \
    split_bf16x2(x0[2], x0[3], h_[1], l_[1]);                                     \
    split_bf16x2(x1[0], x1[1], h_[2], l_[2]);                                     \
    split_bf16x2(x1[2], x1[3], h_[3], l_[3]);                                     \
    hi = (u4v){h_[0], h_[1], h_[2], h_[3]};                                       \
    lo = (u4v){l_[0], l_[1], l_[2], l_[3]};                                       \
  }
#define NQS_BF_STORE(kb_)                                                                     \
  {                                                                                           \
    const int64_t g0 = (int64_t)(kb_) * BF_BK + sk;                                           \
    const bool ok0 = g0 < g.K, ok1 = g0 + 4 < g.K;                                            \
    _Pragma("unroll") for (int q = 0; q < NQA; ++q) {                                         \
      const int row = srow + BF_RPP * q;                                                      \
      const bool am = m0 + row < g.M;                                                         \
      f4v a0 = (ok0 && am) ? ra[q][0] : zero, a1 = (ok1 && am) ? ra[q][1] : zero;             \
      if constexpr (PRO) {                                                                    \
        _Pragma("unroll") for (int e_ = 0; e_ < 8; ++e_) {                                    \
          const float u_ = e_ < 4 ? a0[e_] : a1[e_ - 4];                                      \
          const float v_ = e_ < 4 ? rv[q][0][e_] : rv[q][1][e_ - 4];                          \
          float g0_, g1_, g2_;                                                                \
          gelu_derivs<float>(isval[q] ? u_ : v_, g0_, g1_, g2_);                              \
          const float y_ = isval[q] ? g0_ : g1_ * u_;                                         \
          if (e_ < 4) a0[e_] = y_; else a1[e_ - 4] = y_;                                      \
        }                                                                                     \
      }                                                                                       \
      u4v ah, al;                                                                             \
      NQS_BF_SPLIT8(a0, a1, ah, al)                                                           \
      *reinterpret_cast<u4v*>(As + row * BF_LD + sk) = ah;                                    \
      if (NP == 2) *reinterpret_cast<u4v*>(As + APL + row * BF_LD + sk) = al;                 \
    }                                                                                         \
    _Pragma("unroll") for (int q = 0; q < NQB; ++q) {                                         \
      const int row = srow + BF_RPP * q;                                                      \
      const bool bnn = n0 + row < g.N;                                                        \
      const f4v b0 = (ok0 && bnn) ? rb[q][0] : zero, b1 = (ok1 && bnn) ? rb[q][1] : zero;     \
      u4v bh, bl;                                                                             \
      NQS_BF_SPLIT8(b0, b1, bh, bl)                                                           \
      *reinterpret_cast<u4v*>(Bs + row * BF_LD + sk) = bh;                                    \
      if (NP == 2) *reinterpret_cast<u4v*>(Bs + BPL + row * BF_LD + sk) = bl;                 \
    }                                                                                         \
  }
  NQS_BF_LOAD(0)
  NQS_BF_STORE(0)
  if (nk > 1) NQS_BF_LOAD(1)
  __syncthreads();
  for (int64_t kb = 0; kb < nk; ++kb) {
#pragma unroll
    for (int kk = 0; kk < BF_BK; kk += 16) {
      bf8v ah[BI], al[BI], bh[BJ], bl[BJ];
      const int ko = kk + 8 * (lane >> 5);
#pragma unroll
      for (int i = 0; i < BI; ++i) {
        const int ra_ = (wm * BI * 32 + i * 32 + (lane & 31)) * BF_LD + ko;
        ah[i] = *reinterpret_cast<const bf8v*>(As + ra_);
        if (NP == 2) al[i] = *reinterpret_cast<const bf8v*>(As + APL + ra_);
      }
#pragma unroll
      for (int j = 0; j < BJ; ++j) {
        const int rb_ = (wn * BJ * 32 + j * 32 + (lane & 31)) * BF_LD + ko;
        bh[j] = *reinterpret_cast<const bf8v*>(Bs + rb_);
        if (NP == 2) bl[j] = *reinterpret_cast<const bf8v*>(Bs + BPL + rb_);
      }
#pragma unroll
      for (int i = 0; i < BI; ++i)
#pragma unroll
        for (int j = 0; j < BJ; ++j) {
          if (NP == 2) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          }
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
    }
    if (kb + 1 < nk) {  // uniform
      __syncthreads();
      NQS_BF_STORE(kb + 1)
      if (kb + 2 < nk) NQS_BF_LOAD(kb + 2)
      __syncthreads();
    }
  }
#undef NQS_BF_LOAD
#undef NQS_BF_STORE
#undef NQS_BF_SPLIT8
  // two-phase epilogue as in k_gemm_nt_v2, one 32-row block i at a time (16 BJ operands live)
  const bool has_beta = g.beta != 0.f, has_bias = g.bias != nullptr;
  float bv[BJ];
#pragma unroll
  for (int j = 0; j < BJ; ++j) {
    const int64_t col = n0 + wn * BJ * 32 + j * 32 + (lane & 31);
    bv[j] = has_bias ? g.bias[col < g.N ? col : g.N - 1] : 0.f;
  }
#pragma unroll
  for (int i = 0; i < BI; ++i) {
    float addv[BJ][16];
#pragma unroll
    for (int j = 0; j < BJ; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) addv[j][r] = 0.f;
    if (has_beta) {
#pragma unroll
      for (int j = 0; j < BJ; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int64_t row = m0 + wm * BI * 32 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
          const int64_t col = n0 + wn * BJ * 32 + j * 32 + (lane & 31);
          const int64_t rr = row < g.M ? row : g.M - 1, cc = col < g.N ? col : g.N - 1;
          addv[j][r] = g.beta * g.C[rr * g.ldc + cc];
        }
    }
    if (has_bias) {
#pragma unroll
      for (int j = 0; j < BJ; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const uint32_t row = (uint32_t)(m0 + wm * BI * 32 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3));
          addv[j][r] += (row % (uint32_t)g.bias_period < (uint32_t)g.bias_len) ? bv[j] : 0.f;
        }
    }
#pragma unroll
    for (int j = 0; j < BJ; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t row = m0 + wm * BI * 32 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int64_t col = n0 + wn * BJ * 32 + j * 32 + (lane & 31);
        const float v = g.alpha * acc[i][j][r] + addv[j][r];
        if (row < g.M && col < g.N) g.C[row * g.ldc + col] = v;
      }
  }
}

template <int SPLIT, int WM, int WN, int BI, int BJ, int MINB, int PRO>
static int launch_gemm_bf_tile(const GemmArgs<float>& g, hipStream_t st) {
  typedef BfTile<WM, WN, BI, BJ> TL;
  constexpr size_t lds = bf_lds_bytes<SPLIT, TL::BM, TL::BN>();
  auto kern = k_gemm_nt_bf<SPLIT, WM, WN, BI, BJ, MINB, PRO>;
  if (lds > 64 * 1024) {
    static bool attr = false;
    if (!attr) {
      NQS_CHECK_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      attr = true;
    }
  }
  const int nbn = (int)cdiv(g.N, TL::BN);
  const int64_t ntiles = cdiv(g.M, TL::BM) * nbn;
  NQS_REQUIRE(ntiles < (int64_t)1 << 31, NQS_ERR_UNSUPPORTED, "GEMM too large for one launch");
  hipLaunchKernelGGL(kern, dim3((unsigned)(8 * ((ntiles + 7) / 8))), dim3(256), lds, st, g, nbn, ntiles);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// Tile shapes the autotuner chooses among for the bf16 dense-layer GEMM (first = default).
static const char* const kBfTileNames[5] = {"128x128", "64x128", "128x64", "64x128w", "256x64"};
template <int SPLIT, int PRO>
static int launch_gemm_bf_choice(const GemmArgs<float>& g, int c, hipStream_t st) {
  // with the GELU prologue the staging holds twice the A registers: one workgroup fewer per CU
  switch (c) {
    case 1: return launch_gemm_bf_tile<SPLIT, 1, 4, 2, 1, PRO ? 3 : 4, PRO>(g, st);   // waves side by side, 64 x 32 each
    case 2: return launch_gemm_bf_tile<SPLIT, 4, 1, 1, 2, PRO ? 3 : 4, PRO>(g, st);   // waves stacked, 32 x 64 each
    case 3: return launch_gemm_bf_tile<SPLIT, 2, 2, 1, 2, PRO ? 3 : 4, PRO>(g, st);   // 2 x 2 waves of 32 x 64
    case 4: return launch_gemm_bf_tile<SPLIT, 4, 1, 2, 2, 2, PRO>(g, st);             // waves stacked, 64 x 64 each
    default: return launch_gemm_bf_tile<SPLIT, 2, 2, 2, 2, PRO ? 2 : 3, PRO>(g, st);  // 2 x 2 waves of 64 x 64
  }
}

// The dense-layer GEMM with its tile shape chosen per shape class (split, N, K, log2 M) by the
// autotuner.  A call with beta != 0 accumulates into C, so its trials write to g.tune_scratch
// with beta = 0; without a scratch buffer it uses the shape already chosen for its class.
template <int SPLIT>
static int launch_gemm_bf(const GemmArgs<float>& g, hipStream_t st) {
  const bool pro = g.gelu_cn > 0;
  const TuneKey key{TUNE_GEMM_BF, (pro ? 1000000 : 0) + SPLIT * 100000 + g.N, g.K, tune_bucket(g.M)};
  const bool acc = g.beta != 0.f;
  auto run = [&](const GemmArgs<float>& a, int c) {
    return pro ? launch_gemm_bf_choice<SPLIT, 1>(a, c, st) : launch_gemm_bf_choice<SPLIT, 0>(a, c, st);
  };
  return tune_run_trials(key, 5, 0, kBfTileNames, [&](int c, bool trial) -> int {
    if (trial && acc) {
      GemmArgs<float> t = g;
      t.C = g.tune_scratch;
      t.ldc = g.ldc;
      t.beta = 0.f;
      return run(t, c);
    }
    return run(g, c);
  }, !acc || g.tune_scratch != nullptr, st);
}

__device__ __attribute__((noinline)) void potrf_diag64_call(double* __restrict__ A, int64_t lda, int64_t jg,
                                                            double* __restrict__ Linv, int* __restrict__ info,
                                                            double (*Ls)[65], double* rdiag) {
  potrf_diag64(A, lda, jg, Linv, info, Ls, rdiag);
}

template <typename T>
int gemm_nt(const GemmArgs<T>& g, int batch, hipStream_t st) {
  if (g.M <= 0 || g.N <= 0) return NQS_OK;
  constexpr int VN = Vec16<T>::n;
  const bool vec_ok = (g.K % VN == 0) && (g.lda % VN == 0) && (g.ldb % VN == 0) && (((uintptr_t)g.A & 15) == 0) &&
                      (((uintptr_t)g.B & 15) == 0) && (g.strideA % VN == 0) && (g.strideB % VN == 0);
  if constexpr (sizeof(T) == 4) {
    // shapes the bf16 kernel cannot take (K or a leading dimension not a multiple of 4, batched or
    // triangular calls) run on the float32 kernels below
    NQS_REQUIRE(g.split == 0 || g.split == 1 || g.split == 3, NQS_ERR_ARG, "bf16 GEMM split must be 0, 1 or 3");
    if (g.split != 0 && vec_ok && batch == 1 && g.tri == 0 && !g.mirror)
      return g.split == 3 ? launch_gemm_bf<3>(g, st) : launch_gemm_bf<1>(g, st);
    NQS_REQUIRE(g.gelu_cn == 0, NQS_ERR_UNSUPPORTED, "GELU prologue needs the bf16 GEMM path");
  } else {
    NQS_REQUIRE(g.gelu_cn == 0, NQS_ERR_UNSUPPORTED, "GELU prologue needs the bf16 GEMM path");
  }
  NQS_REQUIRE(g.potrf_linv == nullptr || (vec_ok && g.M >= 64 && g.N >= 64), NQS_ERR_ARG,
              "a fused diagonal factor needs the vector float64 kernel and a 64 x 64 block at C's origin");
  if (vec_ok) {
    dim3 grid((unsigned)cdiv(g.N, 128), (unsigned)cdiv(g.M, 128), (unsigned)batch);
    GemmArgs<T> gl = g;
    gl.tile_order = 0;
    if (g.tri == 1 && g.M == g.N && cdiv(g.M, 128) >= 16) {
      // the lower triangle in super-tile order, 1-D grid
      gl.tile_order = 1;
      grid = dim3((unsigned)(8 * cdiv(supertile_count(cdiv(g.M, 128)), 8)), 1, (unsigned)batch);
    }
    if constexpr (sizeof(T) == 8)
      hipLaunchKernelGGL((k_gemm_nt_v2<T, 16>), grid, dim3(256), 0, st, gl);
    else
      hipLaunchKernelGGL((k_gemm_nt_v2<T, 32>), grid, dim3(256), 0, st, gl);
    NQS_LAUNCH_CHECK();
    return NQS_OK;
  }
  dim3 grid((unsigned)cdiv(g.N, 64), (unsigned)cdiv(g.M, 64), (unsigned)batch);
  if constexpr (sizeof(T) == 8)
    hipLaunchKernelGGL((k_gemm_nt<T, 16>), grid, dim3(256), 0, st, g);
  else
    hipLaunchKernelGGL((k_gemm_nt<T, 32>), grid, dim3(256), 0, st, g);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

template int gemm_nt<double>(const GemmArgs<double>&, int, hipStream_t);
template int gemm_nt<float>(const GemmArgs<float>&, int, hipStream_t);

}  // namespace nqs
