// Launchers of the per-walker physics kernels (physics.hip).
#pragma once
#include "nqs_common.h"

namespace nqs {

struct EAssembleArgs {
  int64_t W;
  int N;
  double m_star, lambda_r, h_z;
  int n_field;
  const double* field;  // n_field x 8
  double field_scale;   // 0: raw Fourier sum; > 0: normalised texture times field_scale
  int n_pin;
  const double* pin;    // n_pin x 8: cx, cy, width, v0, hx, hy, hz, 0
  double A[4];          // cell (row-major), for the pinning images
  const double* R;      // (W, N, 2)
  const int32_t* S;     // (W, N)
  const double* lp;     // (W) complex: log psi
  const double* grad;   // (W, 2N) complex
  const double* lap;    // (W) complex
  const double* lpf;    // (W N) complex: log psi of the flipped configurations (NULL: none, the
                        // Rashba and in-plane field terms are zero)
  const double* tanf;   // (W N, 2) complex: d_{x_i}, d_{y_i} log psi of flipped configuration i
  double* out;          // (W) complex
};
int launch_eloc_assemble(const EAssembleArgs& a, hipStream_t st);

int launch_mcmc_propose(int64_t W, int N, const double* R, const double* xi, double step, const double* A,
                        const double* Ainv, double* Rp, hipStream_t st);
int launch_mcmc_accept_pos(int64_t W, int N, double* R, const double* Rp, double* logp, const double* lp_new,
                           const double* u, int32_t* acc, hipStream_t st);
int launch_mcmc_flip_propose(int64_t W, int N, const int32_t* S, const int32_t* fi, int32_t* Sp, hipStream_t st);
int launch_mcmc_accept_spin(int64_t W, int N, int32_t* S, const int32_t* fi, double* logp, const double* lp_new,
                            const double* u, int32_t* acc, hipStream_t st);

}  // namespace nqs
