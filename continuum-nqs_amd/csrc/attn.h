// Register-resident attention building blocks shared by the attention jet kernels and the fused
// value-only forward: per (configuration, head) one wave holds q, k and v of up to 16 tokens in the
// operand layout of v_mfma_*_16x16x4, read from any row-major [token][q | k | v] buffer (global
// memory or LDS) with row stride rs.
#pragma once
#include "nqs_common.h"
#include "kernels.h"

namespace nqs {

// Register-resident variant on 16x16x4 MFMA for N <= 16 tokens and head width hd <= 16 NQ
// (NQ = 1, 2).
// One wave per (cfg, head), no LDS.  Tokens are padded to 16 and features to 16 NQ with zeros
// (padded keys get score -inf, so their softmax weight is exactly 0).
//
// Scores are formed transposed, S^T = K Q^T, so a lane owns query a = lane%16 and the four keys
// b = row(lane, r) of its accumulator slots; a softmax row is then 4 slots of 16 lanes' worth
// reduced across the lane groups lane/16 (shuffles by 16 and 32).  The feature index of k-slot
// g in MFMA step s is q = 16 (s/4) + 4 g + s%4, so each lane reads 4 consecutive features of
// its token.  For P V the k index (the key) of slot g in step s is b = row(lane, s), i.e. the
// lane's own softmax slot s, so P enters the MFMA straight from registers and V is read as
// V[row(lane, s)][16 t + lane%16].
template <typename T, int NQ>
struct AttnRegs {
  T q[4 * NQ], k[4 * NQ], v[4 * NQ];
};

template <typename T, int NQ>
__device__ __forceinline__ void attn_load(const T* __restrict__ base, int lane, int N, int hd, int d, int h,
                                          int64_t rs, AttnRegs<T, NQ>& x) {
  const int li = lane & 15, g = lane >> 4;
  const int tq = li < N ? li : N - 1;
  if constexpr (sizeof(T) == 4) {
    if (hd == 16 * NQ) {
      // full-width head: a lane's 4 consecutive q (k) features are one 16-byte load
      typedef T v4 __attribute__((ext_vector_type(4)));
#pragma unroll
      for (int t = 0; t < NQ; ++t) {
        const v4 vq = *reinterpret_cast<const v4*>(base + tq * rs + h * hd + 16 * t + 4 * g);
        const v4 vk = *reinterpret_cast<const v4*>(base + tq * rs + d + h * hd + 16 * t + 4 * g);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          x.q[4 * t + e] = li < N ? vq[e] : (T)0;
          x.k[4 * t + e] = li < N ? vk[e] : (T)0;
        }
      }
#pragma unroll
      for (int s = 0; s < 4 * NQ; ++s) {
        const int b = Mma16<T>::row(lane, s & 3), qv = 16 * (s >> 2) + li;
        const int bc = b < N ? b : N - 1;
        const T vv = base[bc * rs + 2 * d + h * hd + qv];
        x.v[s] = b < N ? vv : (T)0;
      }
      return;
    }
  }
#pragma unroll
  for (int s = 0; s < 4 * NQ; ++s) {
    const int q = 16 * (s >> 2) + 4 * g + (s & 3);
    const int qc = q < hd ? q : hd - 1;
    const bool ok = li < N && q < hd;
    const T vq = base[tq * rs + h * hd + qc];
    const T vk = base[tq * rs + d + h * hd + qc];
    x.q[s] = ok ? vq : (T)0;
    x.k[s] = ok ? vk : (T)0;
    const int b = Mma16<T>::row(lane, s & 3), qv = 16 * (s >> 2) + li;
    const int bc = b < N ? b : N - 1, qvc = qv < hd ? qv : hd - 1;
    const T vv = base[bc * rs + 2 * d + h * hd + qvc];
    x.v[s] = (b < N && qv < hd) ? vv : (T)0;
  }
}

template <typename T>
__device__ __forceinline__ T xsum16(T v) {
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 32);
  return v;
}

// Value channel of one (configuration, head): scores S^T = K Q^T scaled by sc, the softmax weights P
// of the lane's query over its four key slots (keys past N weigh exactly 0), and O = P V in
// accumulator layout (query row(lane, r), feature 16 t + lane % 16 in slot r of o[t]).  No FMA
// contraction, so every caller rounds alike.
template <typename T, int NQ>
__device__ __forceinline__ void attn_value(const AttnRegs<T, NQ>& x0, int lane, int N, T sc, T (&P)[4],
                                           typename Mma16<T>::acc_t (&o)[NQ]) {
#pragma clang fp contract(off)
  typedef typename Mma16<T>::acc_t acc_t;
  const acc_t zacc = (acc_t)(T)0;
  acc_t s0 = zacc;
#pragma unroll
  for (int s = 0; s < 4 * NQ; ++s) s0 = Mma16<T>::mma(x0.k[s], x0.q[s], s0);
  T sv[4], mx;
#pragma unroll
  for (int r = 0; r < 4; ++r) sv[r] = Mma16<T>::row(lane, r) < N ? s0[r] * sc : -INFINITY;
  mx = sv[0];
#pragma unroll
  for (int r = 1; r < 4; ++r) mx = sv[r] > mx ? sv[r] : mx;
  T m2 = __shfl_xor(mx, 16);
  mx = m2 > mx ? m2 : mx;
  m2 = __shfl_xor(mx, 32);
  mx = m2 > mx ? m2 : mx;
  T z = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    P[r] = exp(sv[r] - mx);
    z += P[r];
  }
  const T iz = (T)1 / xsum16(z);
#pragma unroll
  for (int r = 0; r < 4; ++r) P[r] *= iz;
#pragma unroll
  for (int t = 0; t < NQ; ++t) {
    o[t] = zacc;
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) o[t] = Mma16<T>::mma(P[kb], x0.v[4 * t + kb], o[t]);
  }
}

// A first-order tangent channel of the same (configuration, head): with the value channel's
// registers x0 and softmax weights P, and the tangent's q, k, v in xc,
// dS^T = (K dQ^T + dK Q^T) sc, dP = P (dS - sum_b P dS), dO = dP V + P dV (same layout as
// attn_value's O).  No FMA contraction.
template <typename T, int NQ>
__device__ __forceinline__ void attn_tangent(const AttnRegs<T, NQ>& x0, const AttnRegs<T, NQ>& xc, const T (&P)[4],
                                             T sc, typename Mma16<T>::acc_t (&o)[NQ]) {
#pragma clang fp contract(off)
  typedef typename Mma16<T>::acc_t acc_t;
  const acc_t zacc = (acc_t)(T)0;
  acc_t sa = zacc;
#pragma unroll
  for (int s = 0; s < 4 * NQ; ++s) {
    sa = Mma16<T>::mma(x0.k[s], xc.q[s], sa);
    sa = Mma16<T>::mma(xc.k[s], x0.q[s], sa);
  }
  T m = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) m += P[r] * (sa[r] * sc);
  m = xsum16(m);
  T Pc[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) Pc[r] = P[r] * (sa[r] * sc - m);
#pragma unroll
  for (int t = 0; t < NQ; ++t) {
    o[t] = zacc;
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      o[t] = Mma16<T>::mma(Pc[kb], x0.v[4 * t + kb], o[t]);
      o[t] = Mma16<T>::mma(P[kb], xc.v[4 * t + kb], o[t]);
    }
  }
}

}  // namespace nqs
