// Launchers of the backward kernels (bwd.hip, det.hip).
#pragma once
#include "nqs_common.h"
#include "kernels.h"

namespace nqs {

template <typename T>
struct WgradArgs {
  const T* X;   // value activations [cfg][t][.], row stride ldx, first `in` columns used
  int64_t ldx;
  int in;
  const T* G;   // cotangents [cfg][c][t][.], row stride ldg, `out` columns
  int64_t ldg;
  int out;
  int N;
  int64_t n_cfg, w0, W_total;
  T* O;         // O matrix: row c * W_total + w, row stride ldo
  int64_t ldo;
  int64_t seg_base[3];  // theta offset of each output segment
  int seg_w;            // width of a segment; a leaf is (in, seg_w) row-major
};
template <typename T>
int launch_wgrad(const WgradArgs<T>& a, hipStream_t st);
template <typename T>
int launch_bgrad(const WgradArgs<T>& a, hipStream_t st);

template <typename T>
struct LnBwdArgs {
  const T* x;      // LayerNorm input (value) [cfg][t][d]
  const T* ybar;   // output cotangent [cfg][2][t][d]
  T* xbar;         // input cotangent [cfg][2][t][d]
  const T* scale;
  int d, N, accumulate;
  int64_t n_cfg, w0, W_total;
  T* O;
  int64_t ldo, off_scale, off_bias;
};
template <typename T>
int launch_ln_bwd(const LnBwdArgs<T>& a, hipStream_t st);
template <typename T>
int launch_gelu_bwd(int64_t n_cfg, int64_t per_cfg, const T* u, const T* gbar, T* ubar, hipStream_t st);
template <typename T>
int launch_attn_bwd(int64_t n_cfg, int N, int d, int nh, const T* qkv, const T* obar, T* qkvbar, hipStream_t st);

// Determinant-head backward: log psi, the backflow cotangent ybar [cfg][2][t][2KN] and the O
// entries of the determinant-head leaves (det_logw, dW, Jastrow).
struct DetBwdOffsets {
  int64_t dlw, dW, jas_same, jas_opp, jas_loga;  // theta offsets, -1 when absent
};
template <typename T>
int launch_det_bwd(int64_t n_cfg, const DetParams& p, const double* R, const int32_t* S, int64_t w0,
                   const T* bf, T* ybar, double* out_lp, T* O, int64_t ldo, int64_t W_total,
                   const DetBwdOffsets& off, hipStream_t st);

}  // namespace nqs
