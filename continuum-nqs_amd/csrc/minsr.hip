// MinSR sample-space solve (src/nqs/minsr.py:68-81):
//   Otil = [Re dO; Im dO] (2W x P, columns centred over walkers), etil = [Re ec; Im ec]
//   T = Otil Otil^T,  (T + lam I) y = etil,  dtheta = Otil^T y
// Kernels: column means / centring (+ bf16 copy), the Gram as a lower-triangular SYRK on MFMA
// (bf16 in, f32 accumulate; or the f64 / f32 MFMA GEMM), a float64 blocked right-looking
// Cholesky (diagonal block factor + inverse in one workgroup, panel and trailing update on the
// f64 MFMA GEMM), blocked triangular solves, one step of iterative refinement and the
// push-through GEMV.  Also the median/MAD clip of the local energies (minsr.py:142-152).
#include <mutex>

#include "nqs_common.h"
#include "kernels.h"
#include "minsr.h"

namespace nqs {

// ------------------------------------------------------------------ column sums, centring
// Streaming passes over O (2W x ldo) read 4 consecutive columns per thread (16 or 32 bytes) and
// split the rows into kSplit chunks per half (rows [0, W) hold Re O, [W, 2W) Im O); each chunk
// writes its partial column sums, which a second kernel adds up in a fixed order.
constexpr int kSplit = 8;  // row chunks per half

// part[s][c] = sum over rows r of chunk s of x(r) O[r][c]  (x = 1 when x is null), over the first
// Wr <= W rows of each half
template <typename T>
__global__ __launch_bounds__(256) void k_colsum_part(const T* __restrict__ O, int64_t ldo, int64_t W, int64_t Wr,
                                                     const double* __restrict__ x, double* __restrict__ part) {
  typedef T v4 __attribute__((ext_vector_type(4)));
  const int64_t c0 = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 4;
  if (c0 >= ldo) return;
  const int s = blockIdx.y, half = s / kSplit, chunk = s % kSplit;
  const int64_t len = (Wr + kSplit - 1) / kSplit;
  const int64_t r0 = half * W + chunk * len;
  const int64_t r1 = min(half * W + Wr, r0 + len);
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  const T* p = O + r0 * ldo + c0;
  int64_t r = r0;
  // four rows' loads in flight per step; the sums keep the row order
  for (; r + 4 <= r1; r += 4, p += 4 * ldo) {
    v4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = *reinterpret_cast<const v4*>(p + u * ldo);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double w = x ? x[r + u] : 1.0;
      a0 += w * (double)v[u][0];
      a1 += w * (double)v[u][1];
      a2 += w * (double)v[u][2];
      a3 += w * (double)v[u][3];
    }
  }
  for (; r < r1; ++r, p += ldo) {
    const v4 v = *reinterpret_cast<const v4*>(p);
    const double w = x ? x[r] : 1.0;
    a0 += w * (double)v[0];
    a1 += w * (double)v[1];
    a2 += w * (double)v[2];
    a3 += w * (double)v[3];
  }
  double* q = part + (int64_t)s * ldo + c0;
  q[0] = a0;
  q[1] = a1;
  q[2] = a2;
  q[3] = a3;
}

// mode 0: mean[h][c] = scale * sum_{chunks of half h} part   (mean is [2][P])
// mode 1: out[c] = sum_{all chunks} part - (mean ? mean[0][c] ysum[0] + mean[1][c] ysum[1] : 0),
//         i.e. (O - 1 mean^T)^T x for the row-wise centred O without forming it
//         written for c < nout; entries P <= c < nout are set to zero (a padded vector)
__global__ void k_colsum_final(const double* __restrict__ part, int64_t ldo, int64_t P, int mode, double scale,
                               const double* __restrict__ mean, const double* __restrict__ ysum,
                               double* __restrict__ out, int64_t nout) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (mode == 1 && c >= P && c < nout) out[c] = 0.0;
  if (c >= P) return;
  double h0 = 0, h1 = 0;
  for (int s = 0; s < kSplit; ++s) h0 += part[(int64_t)s * ldo + c];
  for (int s = kSplit; s < 2 * kSplit; ++s) h1 += part[(int64_t)s * ldo + c];
  if (mode == 0) {
    out[c] = h0 * scale;
    out[P + c] = h1 * scale;
  } else {
    double v = h0 + h1;
    if (mean) v -= mean[c] * ysum[0] + mean[P + c] * ysum[1];
    out[c] = v;
  }
}

// ysum = [sum_{r < W} y_r, sum_{r >= W} y_r]; two workgroups
__global__ __launch_bounds__(256) void k_half_sums(const double* __restrict__ y, int64_t W, double* __restrict__ ysum) {
  __shared__ double red[4];
  const int h = blockIdx.x;
  double s = 0;
  for (int64_t r = threadIdx.x; r < W; r += 256) s += y[h * W + r];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) ysum[h] = red[0] + red[1] + red[2] + red[3];
}

// x[h W + i] -= (1/W) sum_i' x[h W + i'] for both halves (x <- Pi x, Pi the projector off the
// centring null space); one workgroup per half, the sum in a fixed order
__global__ __launch_bounds__(256) void k_proj_halves(double* __restrict__ x, int64_t W) {
  __shared__ double red[4];
  const int h = blockIdx.x;
  double* xh = x + h * W;
  double s = 0;
  for (int64_t r = threadIdx.x; r < W; r += 256) s += xh[r];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  const double m = ((red[0] + red[1]) + (red[2] + red[3])) / (double)W;
  for (int64_t r = threadIdx.x; r < W; r += 256) xh[r] -= m;
}

int proj_halves(double* x, int64_t W, hipStream_t st) {
  hipLaunchKernelGGL(k_proj_halves, dim3(2), dim3(256), 0, st, x, W);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// mdot[h] = mean[h] . u over P columns; two workgroups of 1024 threads
__global__ __launch_bounds__(1024) void k_mean_dot(const double* __restrict__ mean, int64_t P,
                                                   const double* __restrict__ u, double* __restrict__ mdot) {
  __shared__ double red[16];
  const int h = blockIdx.x;
  double s = 0;
  for (int64_t c = threadIdx.x; c < P; c += 1024) s += mean[h * P + c] * u[c];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int q = 0; q < 16; ++q) t += red[q];
    mdot[h] = t;
  }
}

#define CT_ROWS 128  // rows per k_center workgroup

// Centring.  In place (O -= 1 mean^T per half) for the float64 / float32 Gram, which reads O;
// for the bf16 Gram only the centred bf16 copy is written and O is left as it was (the
// refinement and push-through apply the mean correction on the fly).  A thread takes 8
// consecutive columns of one row, in turn for four groups of 32 rows (128 rows x 64 columns per
// workgroup).  With `blocked`, the
// bf16 copy is laid out in 256-row x 64-column blocks, [row panel][column block][256][64], each
// 32 KB contiguous, so the Gram streams a panel slice as one contiguous read instead of 256
// scattered 128-byte pieces.
template <typename T>
__global__ __launch_bounds__(256) void k_center(T* __restrict__ O, int64_t ldo, int64_t W, int64_t P,
                                                const double* __restrict__ mean, __bf16* __restrict__ Ob,
                                                int blocked, int inplace) {
  typedef T v4 __attribute__((ext_vector_type(4)));
  const int64_t k0 = (int64_t)blockIdx.x * 64 + (threadIdx.x & 7) * 8;
  if (k0 >= ldo) return;
  const int64_t nkb = ldo / 64;
  // the 8 means of this thread's columns for both halves (zero past P: padded columns stay zero)
  double mu[2][8];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int e = 0; e < 8; ++e) mu[h][e] = k0 + e < P ? mean[(int64_t)h * P + k0 + e] : 0.0;
#pragma unroll
  for (int q = 0; q < CT_ROWS / 32; ++q) {
    const int64_t r = (int64_t)blockIdx.y * CT_ROWS + q * 32 + (threadIdx.x >> 3);
    if (r >= 2 * W) break;
    const int half = r >= W ? 1 : 0;
    T* orow = O + r * ldo + k0;
    v4 x0 = *reinterpret_cast<const v4*>(orow), x1 = *reinterpret_cast<const v4*>(orow + 4);
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const T raw = e < 4 ? x0[e & 3] : x1[e & 3];
      const T c = k0 + e < P ? (T)((double)raw - mu[half][e]) : (T)0;
      v[e] = (float)c;
      if (e < 4) x0[e & 3] = c;
      else x1[e & 3] = c;
    }
    if (inplace) {
      *reinterpret_cast<v4*>(orow) = x0;
      *reinterpret_cast<v4*>(orow + 4) = x1;
    }
    if (!Ob) continue;
    bf8v b;
#pragma unroll
    for (int e = 0; e < 8; ++e) b[e] = (__bf16)v[e];
    const int64_t off = blocked ? (((r >> 8) * nkb + (k0 >> 6)) * 256 + (r & 255)) * 64 + (k0 & 63) : r * ldo + k0;
    *reinterpret_cast<bf8v*>(Ob + off) = b;
  }
}

bool gram_bf16_blocked(int64_t n, int64_t K) { return n % 256 == 0 && K % 64 == 0; }

int64_t colsum_part_elems(int64_t ldo) { return 2LL * kSplit * ldo; }

template <typename T>
int minsr_center(T* O, int64_t ldo, int64_t W, int64_t P, double* mean, __bf16* Ob, int inplace, int64_t mean_rows,
                 double* part, hipStream_t st) {
  NQS_REQUIRE(ldo % 64 == 0, NQS_ERR_ARG, "ldo must be a multiple of 64");
  NQS_REQUIRE(mean_rows >= 1 && mean_rows <= W, NQS_ERR_ARG, "bad centring row count");
  hipLaunchKernelGGL(k_colsum_part<T>, dim3((unsigned)cdiv(ldo, 1024), 2 * kSplit), dim3(256), 0, st, O, ldo, W,
                     mean_rows, nullptr, part);
  NQS_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_colsum_final, dim3((unsigned)cdiv(P, 256)), dim3(256), 0, st, part, ldo, P, 0,
                     1.0 / (double)mean_rows, nullptr, nullptr, mean, P);
  NQS_LAUNCH_CHECK();
  if (!Ob && !inplace) return NQS_OK;
  const int blocked = gram_bf16_blocked(2 * W, ldo) ? 1 : 0;
  hipLaunchKernelGGL(k_center<T>, dim3((unsigned)cdiv(ldo, 64), (unsigned)cdiv(2 * W, CT_ROWS)), dim3(256), 0, st,
                     O, ldo, W, P, mean, Ob, blocked, inplace);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// out = (O - 1 mean^T)^T y over the 2W rows (mean null: O already centred).  ysum: 2 doubles.
// out gets nout entries, zero beyond P.
template <typename T>
int gemv_cols_c(const T* O, int64_t ldo, int64_t W, int64_t P, const double* y, const double* mean, double* part,
                double* ysum, double* out, int64_t nout, hipStream_t st) {
  hipLaunchKernelGGL(k_colsum_part<T>, dim3((unsigned)cdiv(ldo, 1024), 2 * kSplit), dim3(256), 0, st, O, ldo, W, W,
                     y, part);
  NQS_LAUNCH_CHECK();
  if (mean) {
    hipLaunchKernelGGL(k_half_sums, dim3(2), dim3(256), 0, st, y, W, ysum);
    NQS_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_colsum_final, dim3((unsigned)cdiv(nout > P ? nout : P, 256)), dim3(256), 0, st, part, ldo, P, 1,
                     1.0, mean, ysum, out, nout);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// v[r] = sum_{c < P} O[r][c] u[c] - (mean ? mdot[half(r)] : 0), 8 rows per workgroup so that u
// is read once per 8 rows; columns in 4-wide vectors (pad columns masked out).
template <typename T>
__global__ __launch_bounds__(256) void k_gemv_rows8(const T* __restrict__ O, int64_t ldo, int64_t nrow, int64_t W,
                                                    int64_t P, const double* __restrict__ u,
                                                    const double* __restrict__ mdot, double* __restrict__ v) {
  typedef T v4 __attribute__((ext_vector_type(4)));
  typedef double d4 __attribute__((ext_vector_type(4)));
  __shared__ double red[4][8];
  const int64_t r0 = (int64_t)blockIdx.x * 8;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t c = (int64_t)threadIdx.x * 4; c < ldo; c += 1024) {
    const d4 uu = *reinterpret_cast<const d4*>(u + c);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int64_t r = r0 + q < nrow ? r0 + q : nrow - 1;
      const v4 x = *reinterpret_cast<const v4*>(O + r * ldo + c);
      const double x0 = c < P ? (double)x[0] : 0.0, x1 = c + 1 < P ? (double)x[1] : 0.0;
      const double x2 = c + 2 < P ? (double)x[2] : 0.0, x3 = c + 3 < P ? (double)x[3] : 0.0;
      acc[q] += x0 * uu[0] + x1 * uu[1] + x2 * uu[2] + x3 * uu[3];
    }
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const double s = wave_sum(acc[q]);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][q] = s;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    const int64_t r = r0 + threadIdx.x;
    if (r < nrow) {
      double t = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
      if (mdot) t -= mdot[r >= W ? 1 : 0];
      v[r] = t;
    }
  }
}

// v = (O - 1 mean^T) u; u must hold ldo entries (zero beyond P) and be 32-byte aligned.
template <typename T>
int gemv_rows_c(const T* O, int64_t ldo, int64_t W, int64_t P, const double* u, const double* mean, double* mdot,
                double* v, hipStream_t st) {
  if (mean) {
    hipLaunchKernelGGL(k_mean_dot, dim3(2), dim3(1024), 0, st, mean, P, u, mdot);
    NQS_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_gemv_rows8<T>, dim3((unsigned)cdiv(2 * W, 8)), dim3(256), 0, st, O, ldo, 2 * W, W, P, u,
                     mean ? mdot : nullptr, v);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// etil = [Re(e - mean e); Im(e - mean e)] (one workgroup)
__global__ void k_etil(const double* __restrict__ e, int64_t W, double* __restrict__ et) {
  __shared__ double red[2 * 16];
  double sr = 0, si = 0;
  for (int64_t w = threadIdx.x; w < W; w += blockDim.x) {
    sr += e[2 * w];
    si += e[2 * w + 1];
  }
  sr = wave_sum(sr);
  si = wave_sum(si);
  if ((threadIdx.x & 63) == 0) {
    red[2 * (threadIdx.x >> 6)] = sr;
    red[2 * (threadIdx.x >> 6) + 1] = si;
  }
  __syncthreads();
  double mr = 0, mi = 0;
  for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
    mr += red[2 * q];
    mi += red[2 * q + 1];
  }
  mr /= (double)W;
  mi /= (double)W;
  for (int64_t w = threadIdx.x; w < W; w += blockDim.x) {
    et[w] = e[2 * w] - mr;
    et[W + w] = e[2 * w + 1] - mi;
  }
}

// ------------------------------------------------------------------ bf16 Gram (lower-triangular SYRK)
// C[i][j] (i >= j) = sum_k A[i][k] A[j][k]; A bf16 row-major (n x K), C f32.  Block tile
// 128 x 128 x 64 on 4 waves (64 x 64 each = 2 x 2 of v_mfma_f32_32x32x16_bf16), LDS double
// buffer with a 16-byte row pad, global loads of 16 bytes per lane.  Only tiles on or below
// the diagonal are launched (linear tile index -> (bm, bn)).
#define GB_BM 128
#define GB_BK 64
#define GB_LD (GB_BK + 8)
__global__ __launch_bounds__(256) void k_syrk_bf16(const __bf16* __restrict__ A, int64_t lda, int64_t K,
                                                   float* __restrict__ C, int64_t ldc) {
  __shared__ __bf16 As[2][GB_BM][GB_LD];
  __shared__ __bf16 Bs[2][GB_BM][GB_LD];
  const int64_t t = blockIdx.x;
  int bm = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((int64_t)(bm + 1) * (bm + 2) / 2 <= t) ++bm;
  while ((int64_t)bm * (bm + 1) / 2 > t) --bm;
  const int bn = (int)(t - (int64_t)bm * (bm + 1) / 2);
  const int64_t m0 = (int64_t)bm * GB_BM, n0 = (int64_t)bn * GB_BM;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  f16v acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
  // each thread moves 4 x 16 bytes of A and of B per stage: idx = tid + 256 q, row = idx / 8, chunk = idx % 8
  uint4 ra[4], rb[4];
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int idx = tid + 256 * q, row = idx >> 3, ch = idx & 7;
      ra[q] = *reinterpret_cast<const uint4*>(A + (m0 + row) * lda + k0 + ch * 8);
      rb[q] = *reinterpret_cast<const uint4*>(A + (n0 + row) * lda + k0 + ch * 8);
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int idx = tid + 256 * q, row = idx >> 3, ch = idx & 7;
      *reinterpret_cast<uint4*>(&As[buf][row][ch * 8]) = ra[q];
      *reinterpret_cast<uint4*>(&Bs[buf][row][ch * 8]) = rb[q];
    }
  };
  const int64_t nk = K / GB_BK;
  gload(0);
  sstore(0);
  __syncthreads();
  for (int64_t kb = 0; kb < nk; ++kb) {
    const int cur = kb & 1;
    if (kb + 1 < nk) gload((kb + 1) * GB_BK);
#pragma unroll
    for (int kk = 0; kk < GB_BK; kk += 16) {
      bf8v a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
        a[i] = *reinterpret_cast<const bf8v*>(&As[cur][wm * 64 + i * 32 + (lane & 31)][kk + 8 * (lane >> 5)]);
#pragma unroll
      for (int j = 0; j < 2; ++j)
        b[j] = *reinterpret_cast<const bf8v*>(&Bs[cur][wn * 64 + j * 32 + (lane & 31)][kk + 8 * (lane >> 5)]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (kb + 1 < nk) sstore(cur ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t row = m0 + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int64_t col = n0 + wn * 64 + j * 32 + (lane & 31);
        if (row >= col) C[row * ldc + col] = acc[i][j][r];
      }
}

// 256 x 256 x 64 block tile on 8 waves (4 x 2), each owning a 64 x 128 sub-tile (4 x 8 blocks of
// v_mfma_f32_16x16x32_bf16, 128 accumulators per lane: the same operand traffic per FLOP as
// 32x32x16 blocks, in shorter MFMAs); the operand stages go from global memory straight to LDS
// (double-buffered, 128 KB), so staging costs no registers.  On random data at the full K the
// two changes together cut the Gram from 145 to 123 ms (the kernel is power-limited there, and
// both spend less energy per FLOP).
//
// Tile order.  Every tile streams two 256-row panels over all of K, so HBM traffic is set by how
// many concurrently running tiles share panels in an XCD's L2.  Tiles are grouped in 8 x 8
// super-tiles (lower triangle of super-tiles, row-major; a diagonal super-tile keeps its lower
// triangle), row-major inside a super-tile, and each XCD (workgroup b runs on XCD b % 8) takes a
// contiguous stretch of that sequence: the 32 workgroups an XCD runs at once then cover a 4 x 8
// block reading 12 panels, where a row-major walk of the triangle reads 33.
#define G2_BM 256
#define G2_BK 64

// Split of the last, partly filled round of tiles: tiles [t0, t0 + ntail) (super-tile order) are
// cut into nsplit chunks of kchunk k-steps, one workgroup per (tile, chunk), each writing a
// 256 x 256 partial into part[(tile - t0) * nsplit + chunk]; k_gram_tail_sum adds the chunks in
// a fixed order.  part == nullptr: whole tiles.
struct GramTail {
  float* part;
  int64_t t0, ntail, kchunk;
  int nsplit;
};

// one launch computes the Grams of up to kGramBatch systems of the same size (blockIdx.y = system)
typedef __attribute__((address_space(3))) void* lds_ptr_t;

__global__ __launch_bounds__(512) void k_syrk_bf16_256(GramBatch gb, int64_t ldc, int64_t ntiles, int nt,
                                                       GramTail tail) {
  const __bf16* __restrict__ A = gb.A[blockIdx.y];
  const int64_t K = gb.K[blockIdx.y];
  float* __restrict__ C = gb.C[blockIdx.y];
  // LDS: 2 stages x (row panel, column panel) x 256 rows x 128 bytes (64 bf16 of one k-step),
  // the 16-byte slot s of row r holding k-chunk s ^ (r & 7) (conflict-free fragment reads)
  extern __shared__ __bf16 g2s[];
  char* lds = reinterpret_cast<char*>(g2s);
  int64_t t, kb0 = 0, kb1 = K / G2_BK;
  if (tail.part == nullptr) {
    const int64_t per_xcd = (ntiles + 7) / 8;
    t = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    if (t >= ntiles) return;
  } else {
    const int64_t item = blockIdx.x;
    if (item >= tail.ntail * tail.nsplit) return;
    t = tail.t0 + item / tail.nsplit;
    kb0 = (item % tail.nsplit) * tail.kchunk;
    kb1 = kb0 + tail.kchunk < kb1 ? kb0 + tail.kchunk : kb1;
    C = tail.part + item * (int64_t)(G2_BM * G2_BM);
    ldc = G2_BM;
  }
  int bm, bn;
  supertile_of(t, bm, bn);
  if (bm >= nt || bn > bm) return;
  const int64_t m0 = (int64_t)bm * G2_BM, n0 = (int64_t)bn * G2_BM;
  // output origin: the tile in C, or the partial's own 256 x 256 square
  const int64_t om = tail.part == nullptr ? m0 : 0, on = tail.part == nullptr ? n0 : 0;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave >> 1, wn = wave & 1;
  f4v acc[4][8];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[i][j] = (f4v)0.f;
  // Staging by direct global -> LDS loads (global_load_lds_dwordx4: lane l's 16 bytes land at the
  // wave's LDS base + 16 l), no registers in between.  A is in the blocked layout of k_center, so
  // the k-step kb of a panel is the contiguous block A + (panel * nkb + kb) * 256 * 64; waves 0-3
  // load the row panel, 4-7 the column panel, 64 rows each as 8 loads of 8 rows; lane l fetches the
  // k-chunk (l & 7) ^ (l >> 3) of its row so that it lands in the swizzled slot.
  const int64_t nkb = K / G2_BK;
  const int op = wave >> 2;
  const __bf16* src = A + ((int64_t)(op ? bn : bm) * nkb + kb0) * (G2_BM * G2_BK) +
                      ((wave & 3) * 64 + (lane >> 3)) * G2_BK + (((lane & 7) ^ (lane >> 3)) * 8);
  const int dst0 = op * 32768 + (wave & 3) * 64 * 128;
  const int64_t nk = kb1 - kb0;  // >= 1: chunks are never empty
#define NQS_G2_DL(stage, kb_)                                                                           \
  {                                                                                                     \
    const __bf16* s_ = src + (int64_t)(kb_) * (G2_BM * G2_BK);                                          \
    _Pragma("unroll") for (int j = 0; j < 8; ++j) __builtin_amdgcn_global_load_lds(                     \
        s_ + j * 8 * G2_BK, (lds_ptr_t)(lds + (stage) * 65536 + dst0 + j * 1024), 16, 0, 0);             \
  }
  NQS_G2_DL(0, 0)
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  for (int64_t kb = 0; kb < nk; ++kb) {
    const int cur = kb & 1;
    // stage kb + 1 into the other buffer (its last readers passed the barrier that ended kb - 1)
    if (kb + 1 < nk) NQS_G2_DL(cur ^ 1, kb + 1)
    const char* base = lds + cur * 65536;
    // the fragments of both 32-deep halves in their own registers; the second half's 12 reads go out
    // one after each of the first half's first 12 MFMAs, so their LDS latency hides behind them
    bf8v a[2][4], b[2][8];
    auto rd = [&](int h) {
      const int slot = ((4 * h + (lane >> 4)) ^ (lane & 7)) * 16;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        a[h][i] = *reinterpret_cast<const bf8v*>(base + (wm * 64 + i * 16 + (lane & 15)) * 128 + slot);
#pragma unroll
      for (int j = 0; j < 8; ++j)
        b[h][j] = *reinterpret_cast<const bf8v*>(base + 32768 + (wn * 128 + j * 16 + (lane & 15)) * 128 + slot);
    };
    auto mm = [&](int h) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[h][i], b[h][j], acc[i][j], 0, 0, 0);
    };
    rd(0);
    __builtin_amdgcn_sched_barrier(0);
    // raised wave priority while this wave issues its MFMA bursts: the waves that are between
    // bursts (LDS reads, staging) yield issue slots to it
    __builtin_amdgcn_s_setprio(1);
    rd(1);
    mm(0);
#pragma unroll
    for (int q = 0; q < 12; ++q) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // one LDS read
    }
    __builtin_amdgcn_sched_group_barrier(0x008, 20, 0);
    __builtin_amdgcn_sched_barrier(0);
    mm(1);
    __builtin_amdgcn_s_setprio(0);
    // this wave's loads of stage kb + 1 have landed; the barrier makes every wave's visible and
    // ends the reads of stage kb
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
  }
#undef NQS_G2_DL
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = wm * 64 + i * 16 + 4 * (lane >> 4) + r;
        const int64_t col = wn * 128 + j * 16 + (lane & 15);
        if (m0 + row >= n0 + col) C[(om + row) * ldc + on + col] = acc[i][j][r];
      }
}

// C (lower part of the tail tiles) = sum over the chunks of the partials, chunk 0 first.
// grid (ntail, 256 rows), block 256 columns
__global__ __launch_bounds__(256) void k_gram_tail_sum(GramTail tail, int nt, float* __restrict__ C, int64_t ldc) {
  const int64_t i = blockIdx.x;
  int bm, bn;
  supertile_of(tail.t0 + i, bm, bn);
  if (bm >= nt || bn > bm) return;
  const int row = blockIdx.y, col = threadIdx.x;
  const int64_t gr = (int64_t)bm * G2_BM + row, gc = (int64_t)bn * G2_BM + col;
  if (gr < gc) return;
  const float* p = tail.part + i * tail.nsplit * (int64_t)(G2_BM * G2_BM) + row * G2_BM + col;
  float s = 0.f;
  for (int q = 0; q < tail.nsplit; ++q) s += p[(int64_t)q * G2_BM * G2_BM];
  C[gr * ldc + gc] = s;
}

// CUs of the current device (0 if unknown: no tail split then)
static int device_cus() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  if (cus[dev] == 0 && hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus[dev] = 0;
  return cus[dev];
}

int gram_bf16_batched(const GramBatch& gb, int nsys, int64_t n, int64_t ldc, hipStream_t st, float* scratch,
                      int64_t scratch_elems) {
  NQS_REQUIRE(nsys >= 1 && nsys <= kGramBatch, NQS_ERR_ARG, "batched Gram takes 1 to 32 systems");
  NQS_REQUIRE(gram_bf16_blocked(n, 64), NQS_ERR_UNSUPPORTED, "batched bf16 Gram needs 2W to be a multiple of 256");
  for (int s = 0; s < nsys; ++s) {
    NQS_REQUIRE(gb.K[s] % 64 == 0 && ((uintptr_t)gb.A[s] & 15) == 0, NQS_ERR_ARG,
                "bf16 Gram operand must be 16-byte aligned with K a multiple of 64");
  }
  const int64_t nt = n / G2_BM;
  const int64_t tiles = supertile_count(nt);
  const size_t lds = 2 * 2 * G2_BM * G2_BK * sizeof(__bf16);  // 128 KB
  static bool attr = false;
  if (!attr) {
    NQS_CHECK_HIP(hipFuncSetAttribute((const void*)k_syrk_bf16_256, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
    attr = true;
  }
  // One workgroup fills a CU (its LDS), so the tiles run in rounds of #CU.  With one system, a
  // last round that would leave most CUs idle is split along K instead (partials in `scratch`).
  GramTail tail{nullptr, 0, 0, 0, 1};
  int64_t main_tiles = tiles;
  const int cus = (nsys == 1 && scratch != nullptr) ? device_cus() : 0;
  if (cus > 0 && tiles > cus) {
    const int64_t r = tiles % cus, nk = gb.K[0] / G2_BK;
    int64_t s = r > 0 ? cus / r : 0;
    if (s > 16) s = 16;
    if (s > nk) s = nk;
    if (s >= 2) {
      const int64_t kchunk = (nk + s - 1) / s;
      s = (nk + kchunk - 1) / kchunk;  // every chunk non-empty
      if (r * s * G2_BM * G2_BM <= scratch_elems) {
        main_tiles = tiles - r;
        tail = GramTail{scratch, main_tiles, r, kchunk, (int)s};
      }
    }
  }
  hipLaunchKernelGGL(k_syrk_bf16_256, dim3((unsigned)(8 * ((main_tiles + 7) / 8)), (unsigned)nsys), dim3(512), lds, st,
                     gb, ldc, main_tiles, (int)nt, GramTail{nullptr, 0, 0, 0, 1});
  NQS_LAUNCH_CHECK();
  if (tail.part != nullptr) {
    hipLaunchKernelGGL(k_syrk_bf16_256, dim3((unsigned)(tail.ntail * tail.nsplit)), dim3(512), lds, st, gb, ldc,
                       tiles, (int)nt, tail);
    NQS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gram_tail_sum, dim3((unsigned)tail.ntail, G2_BM), dim3(G2_BM), 0, st, tail, (int)nt, gb.C[0],
                       ldc);
    NQS_LAUNCH_CHECK();
  }
  return NQS_OK;
}

int gram_bf16(const __bf16* A, int64_t lda, int64_t n, int64_t K, float* C, int64_t ldc, float* scratch,
              int64_t scratch_elems, hipStream_t st) {
  NQS_REQUIRE(n % GB_BM == 0, NQS_ERR_UNSUPPORTED, "bf16 Gram needs 2W to be a multiple of 128");
  NQS_REQUIRE(((uintptr_t)A & 15) == 0, NQS_ERR_ARG, "bf16 Gram operand must be 16-byte aligned");
  if (gram_bf16_blocked(n, K)) {  // A in the blocked layout written by k_center
    GramBatch gb;
    gb.A[0] = A;
    gb.K[0] = K;
    gb.C[0] = C;
    return gram_bf16_batched(gb, 1, n, ldc, st, scratch, scratch_elems);
  }
  NQS_REQUIRE(K % GB_BK == 0 && lda % 8 == 0, NQS_ERR_UNSUPPORTED, "bf16 Gram needs K % 64 == 0 and lda % 8 == 0");
  NQS_REQUIRE(((uintptr_t)A & 15) == 0, NQS_ERR_ARG, "bf16 Gram operand must be 16-byte aligned");
  const int64_t nt = n / GB_BM;
  const int64_t tiles = nt * (nt + 1) / 2;
  hipLaunchKernelGGL(k_syrk_bf16, dim3((unsigned)tiles), dim3(256), 0, st, A, lda, K, C, ldc);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// ------------------------------------------------------------------ T (lower) -> float64 + lam I
template <typename T>
__global__ void k_to_f64_shift(const T* __restrict__ Tin, int64_t ldt, int64_t n, double lam, double* __restrict__ A,
                               int64_t lda) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t r = blockIdx.y;
  if (c > r || c >= n) return;
  double v = (double)Tin[r * ldt + c];
  if (r == c) v += lam;
  A[r * lda + c] = v;
}

template <typename T>
int to_f64_shift(const T* Tin, int64_t ldt, int64_t n, double lam, double* A, int64_t lda, hipStream_t st) {
  hipLaunchKernelGGL(k_to_f64_shift<T>, dim3((unsigned)cdiv(n, 256), (unsigned)n), dim3(256), 0, st, Tin, ldt, n, lam,
                     A, lda);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// ------------------------------------------------------------------ deflation of the centring null space
// Centring makes the Re rows and the Im rows of Otil sum to zero, so T u = T v = 0 for
// u = 1_Re/sqrt(W), v = 1_Im/sqrt(W).  etil is orthogonal to u and v and Otil^T u = Otil^T v = 0,
// so the solution and dtheta do not depend on T's restriction to span{u, v}; rounding of a
// low-precision Gram, however, can push those two eigenvalues below -lam.  The lower triangle of
// T is replaced by  Pi T Pi + c (u u^T + v v^T),  Pi = I - u u^T - v v^T, c = mean diag T, which
// has the same solution in exact arithmetic and is positive definite.
// ts[i] = (T u)_i, ts[n + i] = (T v)_i, from the lower triangle: row part (c <= r) ...
__global__ __launch_bounds__(256) void k_defl_rows(const double* __restrict__ A, int64_t lda, int64_t W,
                                                   double* __restrict__ ts) {
  __shared__ double red[8];
  const int64_t r = blockIdx.x;
  const int64_t n = 2 * W;
  double su = 0, sv = 0;
  for (int64_t c = threadIdx.x; c <= r; c += blockDim.x) {
    const double a = A[r * lda + c];
    if (c < W) su += a; else sv += a;
  }
  su = wave_sum(su);
  sv = wave_sum(sv);
  if ((threadIdx.x & 63) == 0) {
    red[2 * (threadIdx.x >> 6)] = su;
    red[2 * (threadIdx.x >> 6) + 1] = sv;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
      a += red[2 * q];
      b += red[2 * q + 1];
    }
    const double s = 1.0 / sqrt((double)W);
    ts[r] = a * s;
    ts[n + r] = b * s;
  }
}
// ... plus the column part (r > c): one thread per column and a 256-row slab per blockIdx.y
// (coalesced over consecutive columns).  Each slab writes its partial sums (zeros where it holds
// no rows below the diagonal) to part[slab][2n]; k_defl_cols_sum adds them to ts in slab order, so
// the result does not depend on the order in which workgroups run.
__global__ void k_defl_cols(const double* __restrict__ A, int64_t lda, int64_t W, double* __restrict__ part) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = 2 * W;
  const int64_t r0 = (int64_t)blockIdx.y * 256;
  if (c >= n) return;
  double su = 0, sv = 0;
  const int64_t r1 = r0 + 256 < n ? r0 + 256 : n;
  for (int64_t r = (c + 1 > r0 ? c + 1 : r0); r < r1; ++r) {
    const double a = A[r * lda + c];
    if (r < W) su += a; else sv += a;
  }
  const double s = 1.0 / sqrt((double)W);
  part[blockIdx.y * 2 * n + c] = su * s;
  part[blockIdx.y * 2 * n + n + c] = sv * s;
}
__global__ void k_defl_cols_sum(const double* __restrict__ part, int64_t n_slabs, int64_t n, double* __restrict__ ts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * n) return;
  double t = 0.0;
  for (int64_t y = 0; y < n_slabs; ++y) t += part[y * 2 * n + i];
  ts[i] += t;
}
int64_t defl_ts_elems(int64_t n) { return 2 * n + cdiv(n, 256) * 2 * n; }
// scalars: alpha = u.Tu, beta = u.Tv, gamma = v.Tv, c = mean diagonal (one workgroup)
__global__ void k_defl_scalars(const double* __restrict__ A, int64_t lda, int64_t W, const double* __restrict__ ts,
                               double* __restrict__ sc) {
  __shared__ double red[4 * 16];
  const int64_t n = 2 * W;
  const double s = 1.0 / sqrt((double)W);
  double al = 0, be = 0, ga = 0, dg = 0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    if (i < W) {
      al += s * ts[i];
      be += s * ts[n + i];
    } else {
      ga += s * ts[n + i];
    }
    dg += A[i * lda + i];
  }
  al = wave_sum(al);
  be = wave_sum(be);
  ga = wave_sum(ga);
  dg = wave_sum(dg);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[4 * w] = al;
    red[4 * w + 1] = be;
    red[4 * w + 2] = ga;
    red[4 * w + 3] = dg;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t[4] = {0, 0, 0, 0};
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q)
      for (int k = 0; k < 4; ++k) t[k] += red[4 * q + k];
    sc[0] = t[0];
    sc[1] = t[1];
    sc[2] = t[2];
    sc[3] = t[3] / (double)n;
  }
}
__global__ void k_defl_apply(double* __restrict__ A, int64_t lda, int64_t W, double lam,
                             const double* __restrict__ ts, const double* __restrict__ sc) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t r = blockIdx.y;
  const int64_t n = 2 * W;
  if (c > r) return;
  const double s = 1.0 / sqrt((double)W);
  const double ur = r < W ? s : 0.0, vr = r < W ? 0.0 : s;
  const double uc = c < W ? s : 0.0, vc = c < W ? 0.0 : s;
  const double ar = ts[r], br = ts[n + r], ac = ts[c], bc = ts[n + c];
  const double al = sc[0], be = sc[1], ga = sc[2], cc = sc[3];
  double d = -(ur * ac + vr * bc) - (ar * uc + br * vc) + al * ur * uc + be * (ur * vc + vr * uc) + ga * vr * vc;
  d += cc * (ur * uc + vr * vc);
  if (r == c) d += lam;
  A[r * lda + c] += d;
}

int deflate_centring(double* A, int64_t lda, int64_t W, double lam, double* ts, double* sc, hipStream_t st) {
  const int64_t n = 2 * W;
  hipLaunchKernelGGL(k_defl_rows, dim3((unsigned)n), dim3(256), 0, st, A, lda, W, ts);
  NQS_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_defl_cols, dim3((unsigned)cdiv(n, 256), (unsigned)cdiv(n, 256)), dim3(256), 0, st, A, lda, W,
                     ts + 2 * n);
  NQS_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_defl_cols_sum, dim3((unsigned)cdiv(2 * n, 256)), dim3(256), 0, st, ts + 2 * n, cdiv(n, 256), n,
                     ts);
  NQS_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_defl_scalars, dim3(1), dim3(1024), 0, st, A, lda, W, ts, sc);
  NQS_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_defl_apply, dim3((unsigned)cdiv(n, 256), (unsigned)n), dim3(256), 0, st, A, lda, W, lam, ts, sc);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// ------------------------------------------------------------------ Cholesky (lower, float64)
// Right-looking over outer panels of CH_NBO = 512 columns (rank-512 trailing updates on the f64
// MFMA GEMM, lower tiles only); inside an outer panel, left-looking over CH_NB = 64 columns:
// factor the 64 x 64 diagonal block, solve the rows below it with its inverse, update the rest
// of the outer panel.
#define CH_NB 64
#define CH_NBO 512

// Factor the 64 x 64 diagonal block at (j0, j0) in place and store its inverse in Linv
// (potrf_diag64, one wave).  The factorisation launches this for its first block only; every later
// block is factored by the workgroup that finishes it inside the preceding update (GemmArgs::potrf_*).
__global__ __launch_bounds__(64) void k_potrf_diag(double* __restrict__ A, int64_t lda, int64_t j0,
                                                   double* __restrict__ Linv, int* __restrict__ info, int64_t sA,
                                                   int64_t sL) {
  __shared__ double Ls[CH_NB][CH_NB + 1];
  __shared__ double rdiag[CH_NB];
  A += blockIdx.z * sA;
  potrf_diag64(A + j0 * lda + j0, lda, j0, Linv + blockIdx.z * sL, info + blockIdx.z, Ls, rdiag);
}

// The factor's columns J0 .. J0 + nbo - 1 below their diagonal blocks, from the outer panel buffer
// P ([n - J0][CH_NBO], row r - J0, column c - J0) into A: element (r, c) with r >= (start of c's
// 64-column block) + 64.  One thread per element of rows J0 + 64 .. n - 1.
__global__ void k_copy_outer_panel(const double* __restrict__ P, int64_t n, int64_t J0, int nbo,
                                   double* __restrict__ A, int64_t lda, int64_t sP, int64_t sA) {
  P += blockIdx.z * sP;
  A += blockIdx.z * sA;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t rows = n - J0 - CH_NB;
  if (q >= rows * nbo) return;
  const int64_t rr = CH_NB + q / nbo;  // row - J0
  const int cc = (int)(q % nbo);       // column - J0
  if (rr < (cc & ~(CH_NB - 1)) + CH_NB) return;
  A[(J0 + rr) * lda + J0 + cc] = P[rr * CH_NBO + cc];
}

int64_t cholesky_panel_elems(int64_t n) { return n * CH_NBO; }

static GemmArgs<double> g64(const double* A, const double* B, double* C, int64_t M, int64_t N, int64_t K, int64_t lda,
                            int64_t ldb, int64_t ldc, double alpha, double beta, int tri) {
  GemmArgs<double> g;
  g.A = A;
  g.B = B;
  g.C = C;
  g.M = M;
  g.N = N;
  g.K = K;
  g.lda = lda;
  g.ldb = ldb;
  g.ldc = ldc;
  g.strideA = g.strideB = g.strideC = 0;
  g.alpha = alpha;
  g.beta = beta;
  g.bias = nullptr;
  g.bias_period = 1;
  g.bias_len = 0;
  g.tri = tri;
  g.mirror = 0;
  g.split = 0;
  return g;
}

// A (n x n, lower) -> L with A = L L^T, in place.  Linv: (n/nb) diagonal-block inverses,
// panel: n x nb scratch.
static GemmArgs<double> strided(GemmArgs<double> g, int64_t sA, int64_t sB, int64_t sC) {
  g.strideA = sA;
  g.strideB = sB;
  g.strideC = sC;
  return g;
}

// The side stream of the Cholesky lookahead, one per device (default priority), and the two
// events that order it against the caller's stream.  Calls take the mutex while they create
// or enqueue, so concurrent callers never interleave their records and waits.
struct CholSide {
  hipStream_t s = nullptr;
  hipEvent_t ready = nullptr, done = nullptr;
};
static std::mutex g_chol_mu;
constexpr int kMaxDevices = 64;
static CholSide g_chol_side[kMaxDevices];

static int chol_side(CholSide** out) {
  int dev = 0;
  NQS_CHECK_HIP(hipGetDevice(&dev));
  NQS_REQUIRE(dev >= 0 && dev < kMaxDevices, NQS_ERR_ARG, "device index out of range");
  CholSide& c = g_chol_side[dev];
  if (!c.s) {
    NQS_CHECK_HIP(hipStreamCreateWithFlags(&c.s, hipStreamNonBlocking));
    NQS_CHECK_HIP(hipEventCreateWithFlags(&c.ready, hipEventDisableTiming));
    NQS_CHECK_HIP(hipEventCreateWithFlags(&c.done, hipEventDisableTiming));
  }
  *out = &c;
  return NQS_OK;
}

// nb independent factorizations of the same size (system z at A + z sA, Linv + z sL,
// panel + z sP, info + z) share every launch.
//
// Lookahead of one outer panel: after panel b is factored, its update of panel b + 1's columns
// runs on the caller's stream, and its update of every column right of that (the bulk of the
// flops) on the side stream, where it overlaps the factorization of panel b + 1 (disjoint
// columns).  The caller's stream waits for that update before it applies panel b + 1 to the
// columns it touched, and once more before returning.
int cholesky(double* A, int64_t lda, int64_t n, double* Linv, double* panel, int* info, hipStream_t st, int nb,
             int64_t sA, int64_t sL, int64_t sP) {
  NQS_REQUIRE(n % CH_NB == 0, NQS_ERR_UNSUPPORTED, "Cholesky size must be a multiple of 64");
  NQS_REQUIRE(nb >= 1 && nb <= 65535, NQS_ERR_ARG, "bad batch count");
  std::lock_guard<std::mutex> lock(g_chol_mu);
  CholSide* side = nullptr;
  NQS_TRY(chol_side(&side));
  bool pending = false;  // a side-stream update the caller's stream has not waited for
  // on every return, including an early error return, the caller's stream ends up ordered after
  // the side stream's last update, so no update is left writing into A behind the caller's back
  struct Join {
    hipStream_t st;
    CholSide* side;
    bool* pending;
    ~Join() {
      if (*pending) (void)hipStreamWaitEvent(st, side->done, 0);
    }
  } join{st, side, &pending};
  // the diagonal factor of block j (with its inverse) into the update that finishes the block
  auto with_potrf = [&](GemmArgs<double> g, int64_t j) {
    g.potrf_linv = Linv + (j / CH_NB) * CH_NB * CH_NB;
    g.potrf_info = info;
    g.potrf_j0 = j;
    g.potrf_sl = sL;
    return g;
  };
  // block 0 on its own; every later diagonal block is factored inside the product that finalises it
  // (the update of the rest of its outer panel, or the trailing update of the panel before), so the
  // dependency chain has no launch of its own for it
  hipLaunchKernelGGL(k_potrf_diag, dim3(1, 1, (unsigned)nb), dim3(64), 0, st, A, lda, (int64_t)0, Linv, info, sA, sL);
  NQS_LAUNCH_CHECK();
  for (int64_t J0 = 0; J0 < n; J0 += CH_NBO) {
    const int64_t nbo = n - J0 < CH_NBO ? n - J0 : CH_NBO;
    for (int64_t j0 = J0; j0 < J0 + nbo; j0 += CH_NB) {
      double* Li = Linv + (j0 / CH_NB) * CH_NB * CH_NB;
      const int64_t m = n - j0 - CH_NB;
      if (m == 0) break;
      // the block's factor column A21 Linv^T (NT GEMM with B = Linv, K = 64) into the outer panel
      // buffer, rows j0 + 64 .. n - 1 and columns j0 .. j0 + 63 relative to J0
      double* pb = panel + (j0 + CH_NB - J0) * CH_NBO + (j0 - J0);
      NQS_TRY(gemm_nt<double>(
          strided(g64(A + (j0 + CH_NB) * lda + j0, Li, pb, m, CH_NB, CH_NB, lda, CH_NB, CH_NBO, 1.0, 0.0, 0), sA, sL,
                  sP),
          nb, st));
      // rest of the outer panel: A[j0+64:, j0+64 : J0+nbo] -= pb pb[0:w]^T (lower part), then the
      // factor of block j0 + 64
      const int64_t w = J0 + nbo - (j0 + CH_NB);
      if (w > 0)
        NQS_TRY(gemm_nt<double>(with_potrf(strided(g64(pb, pb, A + (j0 + CH_NB) * lda + j0 + CH_NB, m, w, CH_NB,
                                                        CH_NBO, CH_NBO, lda, -1.0, 1.0, 1),
                                                    sP, sP, sA),
                                            j0 + CH_NB),
                                nb, st));
    }
    // the outer panel's factor columns into A, one launch per outer panel
    if (n - J0 > CH_NB) {
      hipLaunchKernelGGL(k_copy_outer_panel, dim3((unsigned)cdiv((n - J0 - CH_NB) * nbo, 256), 1, (unsigned)nb),
                         dim3(256), 0, st, panel, n, J0, (int)nbo, A, lda, sP, sA);
      NQS_LAUNCH_CHECK();
    }
    // trailing update with the whole outer panel, A22 -= Lp Lp^T (lower tiles), K = nbo: first the
    // next outer panel's columns (here), then the columns right of it (side stream)
    const int64_t J1 = J0 + nbo, M = n - J1;
    if (M <= 0) continue;
    const int64_t nb1 = M < CH_NBO ? M : CH_NBO, M2 = M - nb1;
    const double* Lp = A + J1 * lda + J0;
    if (pending) {
      NQS_CHECK_HIP(hipStreamWaitEvent(st, side->done, 0));
      pending = false;
    }
    NQS_TRY(gemm_nt<double>(
        with_potrf(strided(g64(Lp, Lp, A + J1 * lda + J1, M, nb1, nbo, lda, lda, lda, -1.0, 1.0, 1), sA, sA, sA), J1),
        nb, st));
    if (M2 > 0) {
      const double* Lq = Lp + nb1 * lda;
      NQS_CHECK_HIP(hipEventRecord(side->ready, st));
      NQS_CHECK_HIP(hipStreamWaitEvent(side->s, side->ready, 0));
      NQS_TRY(gemm_nt<double>(
          strided(g64(Lq, Lq, A + (J1 + nb1) * lda + J1 + nb1, M2, M2, nbo, lda, lda, lda, -1.0, 1.0, 1), sA, sA, sA),
          nb, side->s));
      NQS_CHECK_HIP(hipEventRecord(side->done, side->s));
      pending = true;
    }
  }
  if (pending) {
    pending = false;
    NQS_CHECK_HIP(hipStreamWaitEvent(st, side->done, 0));
  }
  return NQS_OK;
}

// ------------------------------------------------------------------ triangular solves with the factor
// L y = b (forward) or L^T x = y (backward) in one launch of n / 64 workgroups.  Workgroup t (the
// order in which workgroups start: a ticket from an atomic counter) solves the 64-row block
// b = t (forward) or nb - 1 - t (backward): it subtracts the contributions of the blocks solved
// before it, each as soon as that block's values appear, applies the diagonal block's inverse and
// writes its block.  The output vector starts filled with kTsPending, a signalling-NaN bit pattern
// that no arithmetic produces (a NaN from arithmetic is quiet), and every value is written once
// with an agent-scope atomic store, so a reader polls the values themselves (agent-scope atomic
// loads) until none is pending: no flags, no fences.  A workgroup waits only for blocks of lower
// tickets, which started before it and in turn wait only for lower tickets, so the chain advances
// whatever the number of resident workgroups.  Thread (lane, wave w) holds 16 entries of every
// block it applies (row lane of the block, columns 16 w .. 16 w + 15; transposed for the backward
// solve) and accumulates its partial dot product over the blocks in publishing order; the four
// waves' partials are added in a fixed order, so the result does not depend on timing.  The next
// off-diagonal block is loaded while the current one's values are awaited.  A wait longer than
// kTsSpinLimit polls gives up and its values count as NaN, so a broken chain ends the launch with a
// non-finite solution instead of a hang.
constexpr int kTsSpinLimit = 1 << 22;
constexpr uint64_t kTsPending = 0x7FF40000000DEDEDull;

// per system: n doubles (the forward solve's output) and two ticket counters
int64_t chol_sync_elems(int64_t n) { return 2 * n + 4; }

// out[0, n) = kTsPending and, if cnt, cnt[0] = cnt[1] = 0, for each of nb systems (strides in elements)
__global__ void k_ts_reset(uint64_t* __restrict__ out, int64_t n, int64_t so, int32_t* __restrict__ cnt,
                           int64_t sc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[blockIdx.z * so + i] = kTsPending;
  if (cnt && i < 2) cnt[blockIdx.z * sc + i] = 0;
}

// this lane's value (entry `lane & 15` of a 16-value slice) once all 64 lanes see published values
__device__ __forceinline__ double ts_poll(const double* p, bool& ok) {
  uint64_t* q = reinterpret_cast<uint64_t*>(const_cast<double*>(p));
  uint64_t v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (ok) {
    int spins = 0;
    while (!__all(v != kTsPending)) {
      if (++spins > kTsSpinLimit) {
        ok = false;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  return v == kTsPending ? __builtin_nan("") : __builtin_bit_cast(double, v);
}

template <int TRANS>
__device__ __forceinline__ void ts_block(const double* __restrict__ A, int64_t lda, int b, int j, int wave, int lane,
                                         double (&m)[16]) {
  if (TRANS == 0) {
    const double* p = A + (int64_t)(64 * b + lane) * lda + 64 * j + 16 * wave;
#pragma unroll
    for (int k = 0; k < 16; ++k) m[k] = p[k];
  } else {
    const double* p = A + (int64_t)(64 * j + 16 * wave) * lda + 64 * b + lane;
#pragma unroll
    for (int k = 0; k < 16; ++k) m[k] = p[(int64_t)k * lda];
  }
}

// rhs -> out (both n per system, at strides sr, so); cnt: this direction's ticket counter
template <int TRANS>
__global__ __launch_bounds__(256) void k_trsv_chain(const double* __restrict__ A, int64_t lda, int nb,
                                                    const double* __restrict__ Linv, const double* rhs, int64_t sr,
                                                    double* out, int64_t so, int32_t* cnt, int64_t sc, int64_t sA,
                                                    int64_t sL) {
  A += blockIdx.z * sA;
  Linv += blockIdx.z * sL;
  rhs += blockIdx.z * sr;
  out += blockIdx.z * so;
  cnt += blockIdx.z * sc;
  __shared__ int tk;
  __shared__ int bad;
  __shared__ double part[4][64];
  __shared__ double accs[64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) {
    tk = atomicAdd(cnt, 1);
    bad = 0;
  }
  __syncthreads();
  const int t = tk;
  const int b = TRANS ? nb - 1 - t : t;
  // this thread's 16 entries of the diagonal block's inverse, as applied
  const double* Li = Linv + (int64_t)b * CH_NB * CH_NB;
  double li[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) li[k] = TRANS ? Li[(16 * wave + k) * CH_NB + lane] : Li[lane * CH_NB + 16 * wave + k];
  double acc = 0.0;
  bool ok = true;
  // step s: await block j(s)'s 16 values of this wave's slice (one per lane of a 16-lane group,
  // broadcast by readlane), issue the loads of block j(s + 1) into the other buffer, then
  // accumulate block j(s) from the buffer loaded one step earlier
  auto step = [&](int s, const double (&cur)[16], double (&nxt)[16]) {
    const int j = TRANS ? nb - 1 - s : s;
    const double xl = ts_poll(out + 64 * j + 16 * wave + (lane & 15), ok);
    // unconditional (the last step reloads block j, in bounds), so that these loads stay in
    // flight on every path
    ts_block<TRANS>(A, lda, b, s + 1 < t ? (TRANS ? j - 1 : j + 1) : j, wave, lane, nxt);
#pragma unroll
    for (int k = 0; k < 16; ++k) acc += cur[k] * readlane_f64(xl, k);
  };
  double ma[16], mb[16];
  if (t > 0) ts_block<TRANS>(A, lda, b, TRANS ? nb - 1 : 0, wave, lane, ma);
  for (int s = 0; s < t; s += 2) {
    step(s, ma, mb);
    if (s + 1 < t) step(s + 1, mb, ma);
  }
  if (!ok && lane == 0) bad = 1;  // benign race: every writer stores 1
  part[wave][lane] = acc;
  __syncthreads();
  if (wave == 0) accs[lane] = rhs[64 * b + lane] - ((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]));
  __syncthreads();
  double s2 = 0.0;
#pragma unroll
  for (int k = 0; k < 16; ++k) s2 += li[k] * accs[16 * wave + k];
  part[wave][lane] = s2;
  __syncthreads();
  if (wave == 0) {
    double v = (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
    if (bad || v != v) v = __builtin_nan("");  // a quiet NaN, never the pending pattern
    __hip_atomic_store(reinterpret_cast<uint64_t*>(out + 64 * b + lane), __builtin_bit_cast(uint64_t, v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

int chol_solve(const double* L, int64_t lda, int64_t n, const double* Linv, double* x, int32_t* sync, hipStream_t st,
               int nb, int64_t sA, int64_t sL, int64_t sx) {
  NQS_REQUIRE(n % CH_NB == 0, NQS_ERR_UNSUPPORTED, "triangular solve size must be a multiple of 64");
  NQS_REQUIRE(nb >= 1 && nb <= 65535, NQS_ERR_ARG, "bad batch count");
  NQS_REQUIRE(sync && ((uintptr_t)sync & 7) == 0, NQS_ERR_ARG, "triangular solve sync buffer must be 8-byte aligned");
  // per system: y (n doubles) then the two ticket counters; x is the forward's right-hand side,
  // then (refilled with kTsPending once the forward is done) the backward's output
  const unsigned z = (unsigned)nb;
  const int nbk = (int)(n / CH_NB);
  const int64_t ss = chol_sync_elems(n);  // int32 elements per system
  double* y = reinterpret_cast<double*>(sync);
  int32_t* cnt = sync + 2 * n;
  const dim3 fg((unsigned)cdiv(n, 256), 1, z);
  hipLaunchKernelGGL(k_ts_reset, fg, dim3(256), 0, st, reinterpret_cast<uint64_t*>(y), n, ss / 2, cnt, ss);
  NQS_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_trsv_chain<0>, dim3((unsigned)nbk, 1, z), dim3(256), 0, st, L, lda, nbk, Linv, x, sx, y, ss / 2,
                     cnt, ss, sA, sL);
  NQS_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_ts_reset, fg, dim3(256), 0, st, reinterpret_cast<uint64_t*>(x), n, sx, nullptr, 0);
  NQS_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_trsv_chain<1>, dim3((unsigned)nbk, 1, z), dim3(256), 0, st, L, lda, nbk, Linv, y, ss / 2, x, sx,
                     cnt + 1, ss, sA, sL);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// a fill kernel rather than a memset, so that a caller may capture the calls into a graph
__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
int fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, st, p, n, v);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// ------------------------------------------------------------------ GEMVs over O and T
__global__ void k_axpy(int64_t n, double a, const double* __restrict__ x, const double* __restrict__ y,
                       double b, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a * x[i] + b * y[i];
}
int axpby(int64_t n, double a, const double* x, double b, const double* y, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_axpy, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, st, n, a, x, y, b, out);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

// ------------------------------------------------------------------ median / MAD clip (minsr.py:142-152)
// One workgroup; values sorted in LDS by a bitonic network (W <= 8192 per pass of 16384 slots).
__device__ void bitonic_sort(double* v, int n2) {
  for (int k = 2; k <= n2; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < n2; i += blockDim.x) {
        const int l = i ^ j;
        if (l > i) {
          const bool up = (i & k) == 0;
          const double a = v[i], b = v[l];
          if ((a > b) == up) {
            v[i] = b;
            v[l] = a;
          }
        }
      }
      __syncthreads();
    }
}

__global__ __launch_bounds__(1024) void k_clip_stats(const double* __restrict__ e, int W, double clip_k,
                                                     double* __restrict__ e_clip, double* __restrict__ stats) {
  // every operation rounded on its own, as numpy evaluates the reference's expressions
#pragma clang fp contract(off)
  extern __shared__ double v[];
  int n2 = 1;
  while (n2 < W) n2 <<= 1;
  for (int i = threadIdx.x; i < n2; i += blockDim.x) v[i] = (i < W) ? e[2 * i] : INFINITY;
  __syncthreads();
  bitonic_sort(v, n2);
  const double med = (W & 1) ? v[W / 2] : 0.5 * (v[W / 2 - 1] + v[W / 2]);
  __syncthreads();
  for (int i = threadIdx.x; i < n2; i += blockDim.x) v[i] = (i < W) ? fabs(e[2 * i] - med) : INFINITY;
  __syncthreads();
  bitonic_sort(v, n2);
  const double mad = ((W & 1) ? v[W / 2] : 0.5 * (v[W / 2 - 1] + v[W / 2])) + 1e-12;
  // rounded as numpy does (k * MAD first, then the sum), no fused multiply-add
  const double kmad = clip_k * mad;
  const double lo = med - kmad, hi = med + kmad;
  for (int i = threadIdx.x; i < W; i += blockDim.x) {
    const double x = e[2 * i];
    e_clip[2 * i] = x < lo ? lo : (x > hi ? hi : x);
    e_clip[2 * i + 1] = e[2 * i + 1];
  }
  // mean and variance: per-thread strided sums, then the waves' sums in wave order (a fixed
  // order, so the statistics do not depend on timing)
  __shared__ double red[16];
  double s = 0;
  for (int i = threadIdx.x; i < W; i += blockDim.x) s += e[2 * i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  double tot = 0;
  for (int q = 0; q < (int)(blockDim.x >> 6); ++q) tot += red[q];
  const double mean = tot / W;
  __syncthreads();
  double s2 = 0;
  for (int i = threadIdx.x; i < W; i += blockDim.x) s2 += (e[2 * i] - mean) * (e[2 * i] - mean);
  s2 = wave_sum(s2);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double var = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) var += red[q];
    stats[0] = mean;
    stats[1] = sqrt(var / W) / sqrt((double)W);
    stats[2] = med;
    stats[3] = mad;
  }
}

int clip_stats(const double* e, int64_t W, double clip_k, double* e_clip, double* stats, hipStream_t st) {
  NQS_REQUIRE(W >= 1 && W <= 16384, NQS_ERR_UNSUPPORTED, "clip supports 1 <= W <= 16384");
  int n2 = 1;
  while (n2 < W) n2 <<= 1;
  static bool attr = false;
  if (!attr) {
    NQS_CHECK_HIP(hipFuncSetAttribute((const void*)k_clip_stats, hipFuncAttributeMaxDynamicSharedMemorySize, 131072));
    attr = true;
  }
  hipLaunchKernelGGL(k_clip_stats, dim3(1), dim3(1024), sizeof(double) * n2, st, e, (int)W, clip_k, e_clip, stats);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

int make_etil(const double* e, int64_t W, double* et, hipStream_t st) {
  hipLaunchKernelGGL(k_etil, dim3(1), dim3(1024), 0, st, e, W, et);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

template int minsr_center<double>(double*, int64_t, int64_t, int64_t, double*, __bf16*, int, int64_t, double*,
                                  hipStream_t);
template int minsr_center<float>(float*, int64_t, int64_t, int64_t, double*, __bf16*, int, int64_t, double*,
                                 hipStream_t);
template int gemv_cols_c<double>(const double*, int64_t, int64_t, int64_t, const double*, const double*, double*,
                                 double*, double*, int64_t, hipStream_t);
template int gemv_cols_c<float>(const float*, int64_t, int64_t, int64_t, const double*, const double*, double*,
                                double*, double*, int64_t, hipStream_t);
template int gemv_rows_c<double>(const double*, int64_t, int64_t, int64_t, const double*, const double*, double*,
                                 double*, hipStream_t);
template int gemv_rows_c<float>(const float*, int64_t, int64_t, int64_t, const double*, const double*, double*,
                                double*, hipStream_t);
template int to_f64_shift<double>(const double*, int64_t, int64_t, double, double*, int64_t, hipStream_t);
template int to_f64_shift<float>(const float*, int64_t, int64_t, double, double*, int64_t, hipStream_t);

}  // namespace nqs
