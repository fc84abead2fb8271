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

// ---- vector-Jacobian product (walkers summed, no O rows).  With the cotangent folded into
// channel 0 of the backflow cotangent (vjp_fold: channel 0 := g_re ch0 + g_im ch1, channel 1 := 0),
// every backbone leaf's gradient is a sum over all rows (cfg, t) of channel 0:
//   S[i][j] = sum_rows X[row][i] G[row][j]   (X = NULL: X = 1, the bias)
// computed per slice of configurations into part[slice][i][j] (vjp_wsum) and added to g in slice
// order (vjp_wsum_reduce), so the result does not depend on the order in which workgroups run.
template <typename T>
int launch_vjp_fold(int64_t n_cfg, int N, int width, int64_t w0, const double* cot, T* ybar, hipStream_t st);
template <typename T>
int launch_vjp_wsum(const WgradArgs<T>& a, T* part, int64_t n_slices, double* g, hipStream_t st);
int64_t vjp_slices(int64_t n_cfg);
// Small per-walker gradient rows (LayerNorm scale / bias, determinant-head leaves) written by the
// O-row kernels into a compact buffer osm (row c * n_cfg + cfg) are summed over the walkers into g.
// Columns [start[s], start[s + 1]) go to g[dst[s] + q - start[s]] (start[n] = the column count);
// with cot != NULL row c of walker w0 + cfg is weighted by cot[2 (w0 + cfg) + c], else only channel
// 0 counts.
struct OsmMap {
  int n;
  int64_t start[5];
  int64_t dst[4];
};
template <typename T>
int launch_vjp_osm_reduce(const T* osm, int64_t ldsm, int64_t n_cfg, const OsmMap& map, const double* cot,
                          int64_t w0, double* g, hipStream_t st);

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
