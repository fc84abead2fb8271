// Fused value-only Psiformer forward for gfx950 (mixed precision, bf16x3 dense layers).
//
// The value-only jet (one channel: the sampler's proposals and the log psi apply) runs the whole
// backbone in one launch: embedding, every pre-norm block (LayerNorm, q/k/v, multi-head
// attention, output projection and residual, LayerNorm, Dense_0, tanh-GELU, Dense_1 and
// residual), the final LayerNorm and the complex backflow projection (src/nqs/psiformer.py:67-81,
// 128-136).  Only the Fourier features come in and only the backflow activations go out; nothing
// in between touches HBM.  The layer-by-layer path of the same forward moves every activation
// through HBM between its kernels (about 50 KB per token row per forward).
//
// Work split.  One workgroup of 4 waves takes 32 token rows (two configurations of 16 tokens).
// Wave w owns columns 32 w .. 32 w + 31 of the residual stream x for all 32 rows, as one
// v_mfma_f32_32x32x16_bf16 accumulator block kept in registers for the whole forward, so the
// residual adds need no data movement.  LDS holds the bf16 hi/lo planes of the current dense
// layer's input (LayerNorm output or attention output, K = 128) and one more region that is in
// turn the LayerNorm input (float), q | k | v (float) and the GELU output planes (K = 256):
// 66 KB, two workgroups per CU.  Weights come from global memory (L2) as pre-split bf16 hi/lo
// operand fragments that the packer writes once per parameter update.
//
// Arithmetic.  Every dense block accumulates exactly as k_gemm_nt_bf does (16-wide k chunks,
// lo*hi, hi*lo, hi*hi per chunk, the same number of chunks, the same epilogue: acc + bias, or
// acc + (x + bias) for the residual layers); the LayerNorm value row is ln_value and the
// attention value channel attn_value, the functions the layer-by-layer kernels use; GELU is
// gelu_derivs.  The fused forward therefore reproduces the layer-by-layer forward's float32
// values.
#include "nqs_common.h"
#include "kernels.h"
#include "attn.h"
#include "model.h"
#include <cstdlib>

namespace nqs {

constexpr int FU_D = 128, FU_N = 16, FU_NH = 8, FU_HD = 16, FU_FF = 256;
constexpr int FU_R = 32;              // token rows per workgroup
constexpr int FU_HLD = FU_D + 8;      // bf16 row pitch of the K = 128 operand planes
constexpr int FU_GLD = FU_FF + 8;     // bf16 row pitch of the K = 256 operand planes
constexpr int FU_QLD = 3 * FU_D + 4;  // float row pitch of q | k | v
constexpr int FU_MAXL = 8;            // layers
constexpr int FU_HBYTES = 2 * FU_R * FU_HLD * 2;
constexpr int FU_QBYTES = FU_R * FU_QLD * 4;
constexpr int FU_LDS = FU_HBYTES + FU_QBYTES;
static_assert(2 * FU_R * FU_GLD * 2 <= FU_QBYTES && FU_R * FU_D * 4 <= FU_QBYTES, "LDS region sizes");

// A dense layer's weights (N_out x K, K-contiguous, K padded with zeros to a multiple of 16) as
// bf16 hi/lo operand fragments of v_mfma_f32_32x32x16_bf16, written once by the packer:
// frag[nb][kc][plane][lane][8] holds, for column block nb and 16-wide k chunk kc, lane l's eight
// values of column 32 nb + l % 32 at k = 16 kc + 8 (l / 32) + e (plane 0 hi, plane 1 lo), so a
// wave reads one operand as a contiguous 1 KB and does no conversion work.
struct FusedLayer {
  const float *ln0_s, *ln0_b, *qkv_b, *o_b, *ln1_s, *ln1_b, *b0, *b1;
  const u4v *qkv, *o, *w0, *w1;  // fragments
};
struct FusedFwd {
  const float* X0;  // rows x kpad features
  int64_t rows;
  int kpad;
  const u4v* emb;  // fragments, K padded to 64
  const float* emb_b;
  FusedLayer L[FU_MAXL];
  int depth;
  const float *lnf_s, *lnf_b, *bf_b;
  const u4v* bf;  // fragments
  int nbf;        // backflow outputs per row (2 K N)
  float* out;     // rows x nbf
};

// acc = A W[32 nb .. 32 nb + 32)^T over the first NCH of the nkc 16-wide k chunks: A is the 32-row
// bf16 hi/lo operand in LDS (pitch lda), W the layer's fragments.  The same products in the same
// order as k_gemm_nt_bf (lo*hi, hi*lo, hi*hi per chunk; zero chunks past K included).  The
// fragments of up to eight chunks are loaded before the first product (the scheduling barrier keeps
// the compiler from sinking the loads next to their uses), so a block waits for L2 once per eight
// chunks instead of once per chunk.
template <int NCH>
__device__ __forceinline__ f16v fu_block(const __bf16* Ah, const __bf16* Al, int lda, const u4v* __restrict__ frag,
                                         int nkc, int nb, int lane) {
  constexpr int H = NCH > 8 ? 8 : NCH;
  f16v acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  const u4v* fb = frag + (int64_t)nb * nkc * 128 + lane;
  const int ko = 8 * (lane >> 5);
  const __bf16* ah_p = Ah + (lane & 31) * lda + ko;
  const __bf16* al_p = Al + (lane & 31) * lda + ko;
#pragma unroll
  for (int c0 = 0; c0 < NCH; c0 += H) {
    u4v bq[2 * H];
#pragma unroll
    for (int c = 0; c < H; ++c) {
      bq[2 * c] = fb[(c0 + c) * 128];
      bq[2 * c + 1] = fb[(c0 + c) * 128 + 64];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int c = 0; c < H; ++c) {
      const bf8v bh = __builtin_bit_cast(bf8v, bq[2 * c]);
      const bf8v bl = __builtin_bit_cast(bf8v, bq[2 * c + 1]);
      const bf8v ah = *reinterpret_cast<const bf8v*>(ah_p + 16 * (c0 + c));
      const bf8v al = *reinterpret_cast<const bf8v*>(al_p + 16 * (c0 + c));
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
    }
  }
  return acc;
}

// The weight stream of the dense layers: every unit a wave consumes is eight 16-wide k chunks of
// one 32-column block (16 fragment registers).  fu_load8 issues a unit's loads, fu_mma8 runs its
// products; the kernel issues the next unit's loads as soon as the current unit's products are
// issued, so the L2 latency of the weights overlaps the epilogue and whatever LayerNorm, attention
// or barrier comes before the next dense block.
__device__ __forceinline__ void fu_load8(u4v (&b)[16], const u4v* __restrict__ frag, int nkc, int nb, int c0,
                                         int lane) {
  const u4v* fb = frag + ((int64_t)nb * nkc + c0) * 128 + lane;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    b[2 * c] = fb[c * 128];
    b[2 * c + 1] = fb[c * 128 + 64];
  }
  __builtin_amdgcn_sched_barrier(0);
}

// (the LDS operand reads run one chunk ahead of their products; the scheduling barrier per chunk
// keeps the compiler from hoisting all of them, which would cost 64 more registers)
__device__ __forceinline__ void fu_mma8(f16v& acc, const u4v (&b)[16], const __bf16* Ah, const __bf16* Al, int lda,
                                        int c0, int lane) {
  const int ko = 8 * (lane >> 5);
  const __bf16* ah_p = Ah + (lane & 31) * lda + ko + 16 * c0;
  const __bf16* al_p = Al + (lane & 31) * lda + ko + 16 * c0;
  bf8v ahn = *reinterpret_cast<const bf8v*>(ah_p), aln = *reinterpret_cast<const bf8v*>(al_p);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const bf8v ah = ahn, al = aln;
    if (c < 7) {
      ahn = *reinterpret_cast<const bf8v*>(ah_p + 16 * (c + 1));
      aln = *reinterpret_cast<const bf8v*>(al_p + 16 * (c + 1));
    }
    const bf8v bh = __builtin_bit_cast(bf8v, b[2 * c]);
    const bf8v bl = __builtin_bit_cast(bf8v, b[2 * c + 1]);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
  }
}

__device__ __forceinline__ f16v fu_zero() {
  f16v z;
#pragma unroll
  for (int r = 0; r < 16; ++r) z[r] = 0.f;
  return z;
}

// Packer: src (N x K floats, row stride ld) -> fragments with nkc = Kp / 16 chunks (k >= K zero).
__global__ void k_fragsplit(const float* __restrict__ src, int N, int K, int ld, int nkc, u4v* __restrict__ dst) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // one (nb, kc, lane)
  if (id >= (int64_t)(N / 32) * nkc * 64) return;
  const int lane = (int)(id % 64);
  const int64_t kc = (id / 64) % nkc, nb = id / 64 / nkc;
  const float* p = src + (nb * 32 + (lane & 31)) * (int64_t)ld;
  const int k0 = 16 * (int)kc + 8 * (lane >> 5);
  float v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = k0 + e < K ? p[k0 + e] : 0.f;
  uint32_t h[4], l[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) split_bf16x2(v[2 * e], v[2 * e + 1], h[e], l[e]);
  u4v* q = dst + ((nb * nkc + kc) * 2) * 64 + lane;
  q[0] = (u4v){h[0], h[1], h[2], h[3]};
  q[64] = (u4v){l[0], l[1], l[2], l[3]};
}

// accumulator slot r of a lane: row 8 (r / 4) + 4 (lane / 32) + r % 4, column lane % 32 of the block
__device__ __forceinline__ int fu_row(int lane, int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// one float into the bf16 hi/lo planes at element offset o
__device__ __forceinline__ void fu_put(__bf16* Ph, __bf16* Pl, int o, float v) {
  uint32_t h, l;
  split_bf16x2(v, 0.f, h, l);
  reinterpret_cast<uint16_t*>(Ph)[o] = (uint16_t)(h & 0xffffu);
  reinterpret_cast<uint16_t*>(Pl)[o] = (uint16_t)(l & 0xffffu);
}

// eight consecutive floats from LDS (16-byte aligned)
__device__ __forceinline__ void fu_ld8(const float* p, float (&v)[8]) {
  const f4v a = *reinterpret_cast<const f4v*>(p), b = *reinterpret_cast<const f4v*>(p + 4);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    v[e] = a[e];
    v[4 + e] = b[e];
  }
}

// eight consecutive floats into the bf16 hi/lo planes at element offset o (16-byte aligned)
__device__ __forceinline__ void fu_put8(__bf16* Ph, __bf16* Pl, int o, const float (&v)[8]) {
  uint32_t h[4], l[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) split_bf16x2(v[2 * e], v[2 * e + 1], h[e], l[e]);
  *reinterpret_cast<u4v*>(Ph + o) = (u4v){h[0], h[1], h[2], h[3]};
  *reinterpret_cast<u4v*>(Pl + o) = (u4v){l[0], l[1], l[2], l[3]};
}

// LayerNorm of the residual stream (wave w's block xr) into the K = 128 operand planes: the
// blocks go through the float scratch Xs, then wave w normalises rows 8 w .. 8 w + 7, four at a
// time on 16-lane groups (ln_value's layout: lane j holds features 8 j .. 8 j + 7), as
// k_layernorm_jet does.
__device__ __forceinline__ void fu_layernorm(const f16v& xr, int wave, int lane, float* Xs, const float* __restrict__ s,
                                             const float* __restrict__ b, __bf16* Hh, __bf16* Hl) {
  __syncthreads();  // every wave is done with the region Xs aliases
#pragma unroll
  for (int r = 0; r < 16; ++r) Xs[fu_row(lane, r) * FU_D + 32 * wave + (lane & 31)] = xr[r];
  __syncthreads();
  const int j = lane & 15;
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const int row = (FU_R / 4) * wave + 4 * it + (lane >> 4);
    float xv[8], sc[8], y[8], mu, rs;
    fu_ld8(Xs + row * FU_D + 8 * j, xv);
    ln_value<float, 8>(xv, j, FU_D, s, b, sc, y, mu, rs);
    fu_put8(Hh, Hl, row * FU_HLD + 8 * j, y);
  }
  __syncthreads();
}

__global__ __launch_bounds__(256, 2) void k_fused_value(FusedFwd a) {
  extern __shared__ __attribute__((aligned(16))) char fsm[];
  __bf16* Hh = reinterpret_cast<__bf16*>(fsm);
  __bf16* Hl = Hh + FU_R * FU_HLD;
  char* rq = fsm + FU_HBYTES;
  float* Q = reinterpret_cast<float*>(rq);      // q | k | v rows, and the LayerNorm scratch
  __bf16* Gh = reinterpret_cast<__bf16*>(rq);   // GELU output planes
  __bf16* Gl = Gh + FU_R * FU_GLD;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t row0 = (int64_t)blockIdx.x * FU_R;
  const int cw = 32 * wave + (lane & 31);  // this lane's column of the wave's block

  // ---- features -> operand planes (k >= kpad zero up to 64)
  {
    const int r = tid >> 3, k0 = 8 * (tid & 7);
    const int64_t gr = row0 + r < a.rows ? row0 + r : a.rows - 1;
    f4v v0 = (f4v)0.f, v1 = (f4v)0.f;
    if (k0 < a.kpad) {
      v0 = *reinterpret_cast<const f4v*>(a.X0 + gr * a.kpad + k0);
      v1 = *reinterpret_cast<const f4v*>(a.X0 + gr * a.kpad + k0 + 4);
    }
    uint32_t h_[4], l_[4];
    split_bf16x2(v0[0], v0[1], h_[0], l_[0]);
    split_bf16x2(v0[2], v0[3], h_[1], l_[1]);
    split_bf16x2(v1[0], v1[1], h_[2], l_[2]);
    split_bf16x2(v1[2], v1[3], h_[3], l_[3]);
    *reinterpret_cast<u4v*>(Hh + r * FU_HLD + k0) = (u4v){h_[0], h_[1], h_[2], h_[3]};
    *reinterpret_cast<u4v*>(Hl + r * FU_HLD + k0) = (u4v){l_[0], l_[1], l_[2], l_[3]};
  }
  __syncthreads();
  f16v xr;
  {
    const f16v acc = a.kpad <= 32 ? fu_block<2>(Hh, Hl, FU_HLD, a.emb, 4, wave, lane)
                                  : fu_block<4>(Hh, Hl, FU_HLD, a.emb, 4, wave, lane);
    const float bv = a.emb_b[cw];
#pragma unroll
    for (int r = 0; r < 16; ++r) xr[r] = acc[r] + bv;
  }
  const float sc_att = 1.f / sqrtf((float)FU_HD);
  u4v bq[16];  // the weight unit in flight: loaded as soon as the previous unit's products are issued
  if (a.depth > 0)
    fu_load8(bq, a.L[0].qkv, 8, wave, 0, lane);
  else
    fu_load8(bq, a.bf, 8, wave, 0, lane);
  for (int l = 0; l < a.depth; ++l) {
    const FusedLayer& L = a.L[l];
    fu_layernorm(xr, wave, lane, Q, L.ln0_s, L.ln0_b, Hh, Hl);
    // q, k, v: wave w takes column blocks w, w + 4, w + 8 of the 384
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      f16v acc = fu_zero();
      fu_mma8(acc, bq, Hh, Hl, FU_HLD, 0, lane);
      if (j < 2)
        fu_load8(bq, L.qkv, 8, wave + 4 * (j + 1), 0, lane);
      else
        fu_load8(bq, L.o, 8, wave, 0, lane);
      const int col0 = 32 * (wave + 4 * j);
      const float bv = L.qkv_b[col0 + (lane & 31)];
#pragma unroll
      for (int r = 0; r < 16; ++r) Q[fu_row(lane, r) * FU_QLD + col0 + (lane & 31)] = acc[r] + bv;
    }
    __syncthreads();
    // attention: wave w takes (configuration, head) pairs 4 w .. 4 w + 3 of the 2 x 8
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int p = 4 * wave + t, cfg = p >> 3, h = p & 7;
      AttnRegs<float, 1> x0;
      attn_load<float, 1>(Q + cfg * FU_N * FU_QLD, lane, FU_N, FU_HD, FU_D, h, FU_QLD, x0);
      float P[4];
      f4v o[1];
      attn_value<float, 1>(x0, lane, FU_N, sc_att, P, o);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int tok = Mma16<float>::row(lane, r);
        fu_put(Hh, Hl, (cfg * FU_N + tok) * FU_HLD + h * FU_HD + (lane & 15), o[0][r]);
      }
    }
    __syncthreads();
    {
      f16v acc = fu_zero();
      fu_mma8(acc, bq, Hh, Hl, FU_HLD, 0, lane);
      fu_load8(bq, L.w0, 8, wave, 0, lane);
      const float bv = L.o_b[cw];
#pragma unroll
      for (int r = 0; r < 16; ++r) xr[r] = acc[r] + (xr[r] + bv);
    }
    fu_layernorm(xr, wave, lane, Q, L.ln1_s, L.ln1_b, Hh, Hl);
    // Dense_0 + GELU: wave w takes column blocks w and w + 4 of the 256 hidden units
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      f16v acc = fu_zero();
      fu_mma8(acc, bq, Hh, Hl, FU_HLD, 0, lane);
      if (j == 0)
        fu_load8(bq, L.w0, 8, wave + 4, 0, lane);
      else
        fu_load8(bq, L.w1, 16, wave, 0, lane);
      const int col0 = 32 * (wave + 4 * j);
      const float bv = L.b0[col0 + (lane & 31)];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const float u = acc[r] + bv;
        float g0, g1, g2;
        gelu_derivs<float>(u, g0, g1, g2);
        fu_put(Gh, Gl, fu_row(lane, r) * FU_GLD + col0 + (lane & 31), g0);
      }
    }
    __syncthreads();
    {
      // Dense_1, K = 256 as two units
      f16v acc = fu_zero();
      fu_mma8(acc, bq, Gh, Gl, FU_GLD, 0, lane);
      fu_load8(bq, L.w1, 16, wave, 8, lane);
      fu_mma8(acc, bq, Gh, Gl, FU_GLD, 8, lane);
      if (l + 1 < a.depth)
        fu_load8(bq, a.L[l + 1].qkv, 8, wave, 0, lane);
      else
        fu_load8(bq, a.bf, 8, wave, 0, lane);
      const float bv = L.b1[cw];
#pragma unroll
      for (int r = 0; r < 16; ++r) xr[r] = acc[r] + (xr[r] + bv);
    }
  }
  fu_layernorm(xr, wave, lane, Q, a.lnf_s, a.lnf_b, Hh, Hl);
  // backflow: wave w takes column block w, and w + 4 when there are eight
  const int nu = a.nbf / 128;
  for (int j = 0; j < nu; ++j) {
    f16v acc = fu_zero();
    fu_mma8(acc, bq, Hh, Hl, FU_HLD, 0, lane);
    if (j + 1 < nu) fu_load8(bq, a.bf, 8, wave + 4, 0, lane);
    const int col0 = 32 * (wave + 4 * j);
    const float bv = a.bf_b[col0 + (lane & 31)];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int64_t gr = row0 + fu_row(lane, r);
      if (gr < a.rows) a.out[gr * a.nbf + col0 + (lane & 31)] = acc[r] + bv;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Fused first-order jet forward: the spin-flip pass (src/nqs/local_energy.py:63-115 through
// psiformer.py:67-81), whose configurations carry three channels, the value and the two position
// tangents of the flipped electron (48 token rows, channel c, token t at row 16 c + t).
//
// One workgroup of 4 waves per configuration.  The dense layers run on the value-only kernel's
// v_mfma_f32_32x32x16_bf16 blocks and weight fragments, on two overlapping 32-row blocks: A = rows
// 0..31 (value, first tangent) and B = rows 16..47 (first tangent, second tangent).  The first
// tangent is computed twice from the same operands, so its two copies stay equal; every dense
// product is the layer-by-layer GEMM's (k_gemm_nt_bf: 16-wide k chunks, lo*hi, hi*lo, hi*hi).  In
// a block, accumulator slot r < 8 of a lane and slot r + 8 are the same token and column 16 rows
// apart, so the GELU jet of a tangent row finds its value row in the same lane.  Wave w owns
// columns 32 w .. 32 w + 31 of the residual stream (both blocks, 32 registers).  Jet rules: LayerNorm
// ln_value / ln_tangent, attention attn_value / attn_tangent, GELU as in the first-order Dense_1
// prologue of the layer-by-layer path (value rows gelu(u), tangent rows gelu'(u of the token's value
// row) du); biases enter the value rows only.
//
// LDS (76.8 KB, two workgroups per CU): the K = 128 operand planes (LayerNorm outputs), and one
// region that is in turn the LayerNorm scratch, [attention output planes | q, k, v of two heads],
// and the K = 256 GELU output planes.  q, k and v are formed two heads at a time (four quarters);
// in each quarter wave w runs head 2 p + w / 2 for tangent channel 1 + w % 2, both waves of a head
// forming its value channel.
constexpr int FJ_C = 3, FJ_R = FJ_C * FU_N;
constexpr int FJ_QLD = 3 * 32 + 4;  // float pitch of q | k | v of two heads
constexpr int FJ_HBYTES = 2 * FJ_R * FU_HLD * 2;
constexpr int FJ_GBYTES = 2 * FJ_R * FU_GLD * 2;
constexpr int FJ_LDS = FJ_HBYTES + FJ_GBYTES;
static_assert(FJ_HBYTES + FJ_R * FJ_QLD * 4 <= FJ_GBYTES && FJ_R * FU_D * 4 <= FJ_GBYTES, "LDS region sizes");

// the weight unit in flight: four 16-wide k chunks (from c0) of one 32-column block
__device__ __forceinline__ void fj_load4(u4v (&b)[8], const u4v* __restrict__ frag, int nkc, int nb, int c0,
                                         int lane) {
  const u4v* fb = frag + ((int64_t)nb * nkc + c0) * 128 + lane;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    b[2 * c] = fb[c * 128];
    b[2 * c + 1] = fb[c * 128 + 64];
  }
  __builtin_amdgcn_sched_barrier(0);
}

// acc[q] += the 32-row block starting at row 16 q of the operand planes (Ah, Al) times the unit,
// over the planes' chunks c0 .. c0 + 3
template <int NB>
__device__ __forceinline__ void fj_mma4(f16v* acc, const u4v (&b)[8], const __bf16* Ah, const __bf16* Al, int lda,
                                        int c0, int lane) {
  const int ko = 8 * (lane >> 5) + 16 * c0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const bf8v bh = __builtin_bit_cast(bf8v, b[2 * c]);
    const bf8v bl = __builtin_bit_cast(bf8v, b[2 * c + 1]);
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int o = (16 * q + (lane & 31)) * lda + ko + 16 * c;
      const bf8v ah = *reinterpret_cast<const bf8v*>(Ah + o);
      const bf8v al = *reinterpret_cast<const bf8v*>(Al + o);
      acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc[q], 0, 0, 0);
      acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc[q], 0, 0, 0);
      acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc[q], 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// LayerNorm jet of the residual stream (wave w's blocks xr) into the K = 128 operand planes: the
// blocks go through the float scratch Xs; wave w then normalises tokens 4 w .. 4 w + 3 on its four
// 16-lane groups, each as the value row and its two tangent rows (ln_value / ln_tangent layout).
__device__ __forceinline__ void fj_layernorm(const f16v (&xr)[2], int wave, int lane, float* Xs,
                                             const float* __restrict__ s, const float* __restrict__ b, __bf16* Hh,
                                             __bf16* Hl) {
  __syncthreads();  // every wave is done with the region Xs aliases
  const int cw = 32 * wave + (lane & 31);
#pragma unroll
  for (int r = 0; r < 16; ++r) Xs[fu_row(lane, r) * FU_D + cw] = xr[0][r];
#pragma unroll
  for (int r = 8; r < 16; ++r) Xs[(16 + fu_row(lane, r)) * FU_D + cw] = xr[1][r];
  __syncthreads();
  const int j = lane & 15, t = 4 * wave + (lane >> 4);
  float xv[8], sc[8], y[8], mu, rs;
  fu_ld8(Xs + t * FU_D + 8 * j, xv);
  ln_value<float, 8>(xv, j, FU_D, s, b, sc, y, mu, rs);
  fu_put8(Hh, Hl, t * FU_HLD + 8 * j, y);
#pragma unroll
  for (int c = 1; c < FJ_C; ++c) {
    const int row = FU_N * c + t;
    float dx[8], dy[8];
    fu_ld8(Xs + row * FU_D + 8 * j, dx);
    ln_tangent<float, 8>(xv, dx, sc, FU_D, mu, rs, dy);
    fu_put8(Hh, Hl, row * FU_HLD + 8 * j, dy);
  }
  __syncthreads();
}

// residual epilogue of blocks A and B: x = acc + (x + bias) on the value rows (block A slots < 8),
// acc + x elsewhere, as the layer-by-layer GEMM's epilogue with the bias on value rows only
__device__ __forceinline__ void fj_residual(f16v (&xr)[2], const f16v (&acc)[2], float bv) {
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    xr[0][r] = acc[0][r] + (xr[0][r] + (r < 8 ? bv : 0.f));
    xr[1][r] = acc[1][r] + xr[1][r];
  }
}

__global__ __launch_bounds__(256, 2) void k_fused_jet1(FusedFwd a) {
  extern __shared__ __attribute__((aligned(16))) char fsm[];
  __bf16* Hh = reinterpret_cast<__bf16*>(fsm);
  __bf16* Hl = Hh + FJ_R * FU_HLD;
  char* rg = fsm + FJ_HBYTES;
  float* Xs = reinterpret_cast<float*>(rg);     // LayerNorm scratch
  __bf16* Oh = reinterpret_cast<__bf16*>(rg);   // attention output planes
  __bf16* Ol = Oh + FJ_R * FU_HLD;
  float* Qq = reinterpret_cast<float*>(rg + FJ_HBYTES);  // q | k | v of two heads
  __bf16* Gh = reinterpret_cast<__bf16*>(rg);   // GELU output planes
  __bf16* Gl = Gh + FJ_R * FU_GLD;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t row0 = (int64_t)blockIdx.x * FJ_R;
  const int cl = lane & 31;  // this lane's column within a 32-column block
  const f16v z16 = fu_zero();

  // ---- features -> operand planes (k >= kpad zero up to 64)
  for (int it = tid; it < FJ_R * 8; it += 256) {
    const int r = it >> 3, k0 = 8 * (it & 7);
    f4v v0 = (f4v)0.f, v1 = (f4v)0.f;
    if (k0 < a.kpad) {
      v0 = *reinterpret_cast<const f4v*>(a.X0 + (row0 + r) * a.kpad + k0);
      v1 = *reinterpret_cast<const f4v*>(a.X0 + (row0 + r) * a.kpad + k0 + 4);
    }
    uint32_t h_[4], l_[4];
    split_bf16x2(v0[0], v0[1], h_[0], l_[0]);
    split_bf16x2(v0[2], v0[3], h_[1], l_[1]);
    split_bf16x2(v1[0], v1[1], h_[2], l_[2]);
    split_bf16x2(v1[2], v1[3], h_[3], l_[3]);
    *reinterpret_cast<u4v*>(Hh + r * FU_HLD + k0) = (u4v){h_[0], h_[1], h_[2], h_[3]};
    *reinterpret_cast<u4v*>(Hl + r * FU_HLD + k0) = (u4v){l_[0], l_[1], l_[2], l_[3]};
  }
  __syncthreads();
  u4v bq[8];  // the weight unit in flight
  f16v xr[2];
  {
    // embedding: K = kpad <= 64, four chunks (zero past kpad)
    fj_load4(bq, a.emb, 4, wave, 0, lane);
    f16v acc[2] = {z16, z16};
    fj_mma4<2>(acc, bq, Hh, Hl, FU_HLD, 0, lane);
    const float bv = a.emb_b[32 * wave + cl];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      xr[0][r] = acc[0][r] + (r < 8 ? bv : 0.f);
      xr[1][r] = acc[1][r];
    }
  }
  const float sc_att = 1.f / sqrtf((float)FU_HD);
  // q, k, v of a quarter: six items (column block u = q, k, v of heads 2 p, 2 p + 1; block A or B),
  // wave w takes items w and w + 4
  auto qkv_nb = [](int p, int item) { return 4 * (item >> 1) + p; };
  if (a.depth > 0)
    fj_load4(bq, a.L[0].qkv, 8, qkv_nb(0, wave), 0, lane);
  else
    fj_load4(bq, a.bf, 8, wave, 0, lane);
  for (int l = 0; l < a.depth; ++l) {
    const FusedLayer& L = a.L[l];
    fj_layernorm(xr, wave, lane, Xs, L.ln0_s, L.ln0_b, Hh, Hl);
    for (int p = 0; p < 4; ++p) {
      if (p > 0) __syncthreads();  // the previous quarter's attention is done with Qq
      for (int it = wave; it < 6; it += 4) {
        const int nb = qkv_nb(p, it), blk = it & 1;
        f16v acc[1] = {z16};
        fj_mma4<1>(acc, bq, Hh + 16 * blk * FU_HLD, Hl + 16 * blk * FU_HLD, FU_HLD, 0, lane);
        fj_load4(bq, L.qkv, 8, nb, 4, lane);
        fj_mma4<1>(acc, bq, Hh + 16 * blk * FU_HLD, Hl + 16 * blk * FU_HLD, FU_HLD, 4, lane);
        // next unit: this quarter's next item, the next quarter's first, or the output projection
        if (it + 4 < 6)
          fj_load4(bq, L.qkv, 8, qkv_nb(p, it + 4), 0, lane);
        else if (p < 3)
          fj_load4(bq, L.qkv, 8, qkv_nb(p + 1, wave), 0, lane);
        else
          fj_load4(bq, L.o, 8, wave, 0, lane);
        const int qc = 32 * (it >> 1) + cl;
        const float bv = L.qkv_b[32 * nb + cl];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          if (blk == 1 && r < 8) continue;  // block B's first tangent rows: block A has them
          const float v = acc[0][r] + ((blk == 0 && r < 8) ? bv : 0.f);
          Qq[(16 * blk + fu_row(lane, r)) * FJ_QLD + qc] = v;
        }
      }
      __syncthreads();
      {
        const int hl = wave >> 1, ch = 1 + (wave & 1);
        const int oc = 16 * (2 * p + hl) + (lane & 15);
        AttnRegs<float, 1> x0, xc;
        attn_load<float, 1>(Qq, lane, FU_N, FU_HD, 32, hl, FJ_QLD, x0);
        float P[4];
        f4v o[1];
        attn_value<float, 1>(x0, lane, FU_N, sc_att, P, o);
        if ((wave & 1) == 0) {
#pragma unroll
          for (int r = 0; r < 4; ++r) fu_put(Oh, Ol, Mma16<float>::row(lane, r) * FU_HLD + oc, o[0][r]);
        }
        attn_load<float, 1>(Qq + FU_N * ch * FJ_QLD, lane, FU_N, FU_HD, 32, hl, FJ_QLD, xc);
        attn_tangent<float, 1>(x0, xc, P, sc_att, o);
#pragma unroll
        for (int r = 0; r < 4; ++r) fu_put(Oh, Ol, (FU_N * ch + Mma16<float>::row(lane, r)) * FU_HLD + oc, o[0][r]);
      }
    }
    __syncthreads();
    {
      // output projection and residual
      f16v acc[2] = {z16, z16};
      fj_mma4<2>(acc, bq, Oh, Ol, FU_HLD, 0, lane);
      fj_load4(bq, L.o, 8, wave, 4, lane);
      fj_mma4<2>(acc, bq, Oh, Ol, FU_HLD, 4, lane);
      fj_load4(bq, L.w0, 8, wave, 0, lane);
      fj_residual(xr, acc, L.o_b[32 * wave + cl]);
    }
    fj_layernorm(xr, wave, lane, Xs, L.ln1_s, L.ln1_b, Hh, Hl);
    // Dense_0 and the GELU jet: wave w takes hidden column blocks w and w + 4
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int nb = wave + 4 * j;
      f16v acc[2] = {z16, z16};
      fj_mma4<2>(acc, bq, Hh, Hl, FU_HLD, 0, lane);
      fj_load4(bq, L.w0, 8, nb, 4, lane);
      fj_mma4<2>(acc, bq, Hh, Hl, FU_HLD, 4, lane);
      if (j == 0)
        fj_load4(bq, L.w0, 8, wave + 4, 0, lane);
      else
        fj_load4(bq, L.w1, 16, wave, 0, lane);
      const int gc = 32 * nb + cl;
      const float bv = L.b0[gc];
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int row = fu_row(lane, r);
        const float u = acc[0][r] + bv;
        float g0, g1, g2;
        gelu_derivs<float>(u, g0, g1, g2);
        fu_put(Gh, Gl, row * FU_GLD + gc, g0);
        fu_put(Gh, Gl, (FU_N + row) * FU_GLD + gc, g1 * acc[0][r + 8]);
        fu_put(Gh, Gl, (2 * FU_N + row) * FU_GLD + gc, g1 * acc[1][r + 8]);
      }
    }
    __syncthreads();
    {
      // Dense_1 (K = 256: four units) and residual
      f16v acc[2] = {z16, z16};
#pragma unroll
      for (int c0 = 0; c0 < 16; c0 += 4) {
        fj_mma4<2>(acc, bq, Gh, Gl, FU_GLD, c0, lane);
        if (c0 < 12)
          fj_load4(bq, L.w1, 16, wave, c0 + 4, lane);
        else if (l + 1 < a.depth)
          fj_load4(bq, a.L[l + 1].qkv, 8, qkv_nb(0, wave), 0, lane);
        else
          fj_load4(bq, a.bf, 8, wave, 0, lane);
      }
      fj_residual(xr, acc, L.b1[32 * wave + cl]);
    }
  }
  fj_layernorm(xr, wave, lane, Xs, a.lnf_s, a.lnf_b, Hh, Hl);
  // backflow: column blocks wave, wave + 4, ... of the nbf outputs
  const int ncb = a.nbf / 32;
  for (int nb = wave; nb < ncb; nb += 4) {
    f16v acc[2] = {z16, z16};
    fj_mma4<2>(acc, bq, Hh, Hl, FU_HLD, 0, lane);
    fj_load4(bq, a.bf, 8, nb, 4, lane);
    fj_mma4<2>(acc, bq, Hh, Hl, FU_HLD, 4, lane);
    if (nb + 4 < ncb) fj_load4(bq, a.bf, 8, nb + 4, 0, lane);
    const float bv = a.bf_b[32 * nb + cl];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int64_t gr = row0 + fu_row(lane, r);
      a.out[gr * a.nbf + 32 * nb + cl] = acc[0][r] + (r < 8 ? bv : 0.f);
    }
#pragma unroll
    for (int r = 8; r < 16; ++r) a.out[(row0 + 16 + fu_row(lane, r)) * a.nbf + 32 * nb + cl] = acc[1][r];
  }
}

bool fused_model_ok(const Model& md) {
  if (md.split != 3 || md.N != FU_N || md.d != FU_D || md.nh != FU_NH || md.ff * md.d != FU_FF) return false;
  if (md.depth > FU_MAXL || md.kpad > 64 || md.kpad % 8 != 0) return false;
  return 2 * md.KN == 128 || 2 * md.KN == 256;  // one or two backflow units per wave
}

bool fused_value_ok(const Model& md, const JetGeom& jg) {
  if (!md.fused || md.tsize != 4 || jg.C != 1 || jg.nt != 0 || jg.has_lap) return false;
  const char* e = getenv("NQS_LAYERWISE_FORWARD");
  return !(e && e[0] == '1');
}

bool fused_jet1_ok(const Model& md, const JetGeom& jg) {
  if (!md.fused || md.tsize != 4 || jg.C != FJ_C || jg.nt != 2 || jg.has_lap) return false;
  const char* e = getenv("NQS_LAYERWISE_FORWARD");
  return !(e && e[0] == '1');
}

int pack_fused_weights(const Model& md, float* pk, hipStream_t st) {
  auto split = [&](int64_t src, int N, int K, int ld, int Kp, int64_t dst) -> int {
    const int nkc = Kp / 16;
    const int64_t n = (int64_t)(N / 32) * nkc * 64;
    hipLaunchKernelGGL(k_fragsplit, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, st, pk + src, N, K, ld, nkc,
                       reinterpret_cast<u4v*>(pk + dst));
    NQS_LAUNCH_CHECK();
    return NQS_OK;
  };
  const int d = md.d, ffd = md.ff * md.d;
  NQS_TRY(split(md.embT, d, md.kpad, md.kpad, 64, md.fz_emb));
  for (int l = 0; l < md.depth; ++l) {
    const LayerPack& P = md.layers[l];
    NQS_TRY(split(P.qkvT, 3 * d, d, d, d, P.fz_qkv));
    NQS_TRY(split(P.oT, d, d, d, d, P.fz_o));
    NQS_TRY(split(P.w0T, ffd, d, d, d, P.fz_w0));
    NQS_TRY(split(P.w1T, d, ffd, ffd, ffd, P.fz_w1));
  }
  return split(md.bfT, 2 * md.KN, d, d, d, md.fz_bf);
}

static FusedFwd fused_args(const Model& md, const float* pk, const float* X0, int64_t rows, float* out) {
  auto fr = [&](int64_t off) { return reinterpret_cast<const u4v*>(pk + off); };
  FusedFwd a;
  a.X0 = X0;
  a.rows = rows;
  a.kpad = md.kpad;
  a.emb = fr(md.fz_emb);
  a.emb_b = pk + md.emb_b;
  a.depth = md.depth;
  for (int l = 0; l < md.depth; ++l) {
    const LayerPack& P = md.layers[l];
    a.L[l] = FusedLayer{pk + P.ln0_s, pk + P.ln0_b, pk + P.qkv_b, pk + P.o_b, pk + P.ln1_s, pk + P.ln1_b,
                        pk + P.b0,    pk + P.b1,    fr(P.fz_qkv), fr(P.fz_o), fr(P.fz_w0), fr(P.fz_w1)};
  }
  a.lnf_s = pk + md.lnf_s;
  a.lnf_b = pk + md.lnf_b;
  a.bf = fr(md.fz_bf);
  a.bf_b = pk + md.bf_b;
  a.nbf = 2 * md.KN;
  a.out = out;
  return a;
}

int launch_fused_value(const Model& md, const float* pk, const float* X0, int64_t rows, float* out, hipStream_t st) {
  if (rows <= 0) return NQS_OK;
  NQS_REQUIRE(md.fused, NQS_ERR_UNSUPPORTED, "fused forward: architecture not supported");
  NQS_REQUIRE(rows % FU_N == 0, NQS_ERR_ARG, "fused forward: rows must be whole configurations");
  const FusedFwd a = fused_args(md, pk, X0, rows, out);
  static bool attr = false;
  if (!attr) {
    NQS_CHECK_HIP(hipFuncSetAttribute((const void*)k_fused_value, hipFuncAttributeMaxDynamicSharedMemorySize, FU_LDS));
    attr = true;
  }
  hipLaunchKernelGGL(k_fused_value, dim3((unsigned)cdiv(rows, FU_R)), dim3(256), FU_LDS, st, a);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

int launch_fused_jet1(const Model& md, const float* pk, const float* X0, int64_t n_cfg, float* out, hipStream_t st) {
  if (n_cfg <= 0) return NQS_OK;
  NQS_REQUIRE(md.fused, NQS_ERR_UNSUPPORTED, "fused forward: architecture not supported");
  NQS_REQUIRE(n_cfg <= 0x7fffffffLL, NQS_ERR_ARG, "fused jet forward: too many configurations per launch");
  const FusedFwd a = fused_args(md, pk, X0, n_cfg * FJ_R, out);
  static bool attr = false;
  if (!attr) {
    NQS_CHECK_HIP(hipFuncSetAttribute((const void*)k_fused_jet1, hipFuncAttributeMaxDynamicSharedMemorySize, FJ_LDS));
    attr = true;
  }
  hipLaunchKernelGGL(k_fused_jet1, dim3((unsigned)n_cfg), dim3(256), FJ_LDS, st, a);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

}  // namespace nqs
