// Launchers of the MinSR kernels (minsr.hip).
#pragma once
#include "nqs_common.h"

namespace nqs {

template <typename T>
int minsr_center(T* O, int64_t ldo, int64_t W, int64_t P, double* mean, __bf16* Ob, hipStream_t st);
int make_etil(const double* e, int64_t W, double* et, hipStream_t st);
int gram_bf16(const __bf16* A, int64_t lda, int64_t n, int64_t K, float* C, int64_t ldc, hipStream_t st);
template <typename T>
int to_f64_shift(const T* Tin, int64_t ldt, int64_t n, double lam, double* A, int64_t lda, hipStream_t st);
int deflate_centring(double* A, int64_t lda, int64_t W, double lam, double* ts, double* sc, hipStream_t st);
int cholesky(double* A, int64_t lda, int64_t n, double* Linv, double* panel, int* info, hipStream_t st);
int chol_solve(const double* L, int64_t lda, int64_t n, const double* Linv, double* x, hipStream_t st);
template <typename T>
int gemv_rows(const T* A, int64_t lda, int64_t nrow, int64_t ncol, const double* x, double* y, double alpha,
              double beta, int sym_lower, hipStream_t st);
template <typename T>
int gemv_cols(const T* A, int64_t lda, int64_t nrow, int64_t ncol, const double* x, double* y, hipStream_t st);
int axpby(int64_t n, double a, const double* x, double b, const double* y, double* out, hipStream_t st);
int clip_stats(const double* e, int64_t W, double clip_k, double* e_clip, double* stats, hipStream_t st);

}  // namespace nqs
