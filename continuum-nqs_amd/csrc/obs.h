// Launchers of the observable kernels (observables.hip).
#pragma once
#include "nqs_common.h"

namespace nqs {

int launch_spin_est(int64_t W, int N, const int32_t* S, const double* lp, const double* lpf, double* out,
                    hipStream_t st);
int launch_spin_fourier(int64_t W, int N, const double* R, const double* sp, const double* q, int nq, double* out,
                        hipStream_t st);
int launch_spin_texture(int64_t n_el, const double* R, const double* sp, const double ai[4], int ngrid, double* m,
                        double* rho, hipStream_t st);
int launch_ratio_stats(int64_t W, const double* la, const double* lb, const double* ph, int shift, double* out,
                       hipStream_t st);

}  // namespace nqs
