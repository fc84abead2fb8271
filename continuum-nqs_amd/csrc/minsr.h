// Launchers of the MinSR kernels (minsr.hip).
#pragma once
#include "nqs_common.h"

namespace nqs {

// column means of O per half over the half's first mean_rows rows into mean [2][P]; then, if
// inplace, O -= mean per half and, if Ob, the centred bf16 copy (blocked layout when
// gram_bf16_blocked).  part: colsum_part_elems doubles.
int64_t colsum_part_elems(int64_t ldo);
template <typename T>
int minsr_center(T* O, int64_t ldo, int64_t W, int64_t P, double* mean, __bf16* Ob, int inplace, int64_t mean_rows,
                 double* part, hipStream_t st);
// Walkers whose mean shifts O before the bf16 down-cast.  The deflation projects the Gram of the
// shifted rows onto the exactly centred one (Pi T' Pi = T for any shift common to a half), and the
// solve projects its residual the same way, so the shift only has to keep the down-cast's rounding
// relative to the walker spread: a 512-walker mean differs from the full mean by about 4% of it.
constexpr int64_t kBf16MeanRows = 512;
// x <- Pi x: each half of x (2W) minus its mean
int proj_halves(double* x, int64_t W, hipStream_t st);
// out = (O - 1 mean^T)^T y  and  v = (O - 1 mean^T) u  (mean null: O is already centred)
template <typename T>
int gemv_cols_c(const T* O, int64_t ldo, int64_t W, int64_t P, const double* y, const double* mean, double* part,
                double* ysum, double* out, int64_t nout, hipStream_t st);
template <typename T>
int gemv_rows_c(const T* O, int64_t ldo, int64_t W, int64_t P, const double* u, const double* mean, double* mdot,
                double* v, hipStream_t st);
int make_etil(const double* e, int64_t W, double* et, hipStream_t st);
// true when minsr_center writes the bf16 copy in the blocked layout the 256-tile Gram reads
bool gram_bf16_blocked(int64_t n, int64_t K);
// scratch (scratch_elems floats, may be null) holds the K-split partials of a last partial round
// of tiles; it is free again when the call returns (on the stream)
int gram_bf16(const __bf16* A, int64_t lda, int64_t n, int64_t K, float* C, int64_t ldc, float* scratch,
              int64_t scratch_elems, hipStream_t st);
// Grams of up to kGramBatch same-size systems in one launch (operands in the blocked layout)
constexpr int kGramBatch = 32;
struct GramBatch {
  const __bf16* A[kGramBatch];
  int64_t K[kGramBatch];
  float* C[kGramBatch];
};
int gram_bf16_batched(const GramBatch& gb, int nsys, int64_t n, int64_t ldc, hipStream_t st, float* scratch = nullptr,
                      int64_t scratch_elems = 0);
template <typename T>
int to_f64_shift(const T* Tin, int64_t ldt, int64_t n, double lam, double* A, int64_t lda, hipStream_t st);
// ts must hold defl_ts_elems(2W) doubles: the two projections T u, T v (2 x 2W) then the per-slab
// partial column sums.
int64_t defl_ts_elems(int64_t n);
int deflate_centring(double* A, int64_t lda, int64_t W, double lam, double* ts, double* sc, hipStream_t st);
// nb same-size systems share each launch: system z uses A + z sA, Linv + z sL, panel + z sP,
// info + z (x + z sx in the solve).  panel: cholesky_panel_elems(n) doubles per system.
int64_t cholesky_panel_elems(int64_t n);
int cholesky(double* A, int64_t lda, int64_t n, double* Linv, double* panel, int* info, hipStream_t st, int nb = 1,
             int64_t sA = 0, int64_t sL = 0, int64_t sP = 0);
// L L^T x = b in place (x holds b): one launch per direction (k_trsv_chain).  sync: scratch of
// chol_sync_elems(n) int32 per system (at that stride), 8-byte aligned.
int chol_solve(const double* L, int64_t lda, int64_t n, const double* Linv, double* x, int32_t* sync, hipStream_t st,
               int nb = 1, int64_t sA = 0, int64_t sL = 0, int64_t sx = 0);
int64_t chol_sync_elems(int64_t n);
int fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st);
int axpby(int64_t n, double a, const double* x, double b, const double* y, double* out, hipStream_t st);
int clip_stats(const double* e, int64_t W, double clip_k, double* e_clip, double* stats, hipStream_t st);

}  // namespace nqs
