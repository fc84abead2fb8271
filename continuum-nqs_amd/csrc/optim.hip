// First-order VMC training (src/nqs/train.py:28-65): the energy gradient of the surrogate loss
// and the Adam update of optax.adam, on the flat real parameter vector.
//
// Gradient: surrogate(theta) = 2 mean_w Re[conj(e_w - <e>) log psi_theta(w)], so
//   g_k = (2/W) sum_w [Re(e_w - <e>) Re O_wk + Im(e_w - <e>) Im O_wk]
// = O^T v over the 2W rows [Re O; Im O] with v = (2/W) [Re(e - <e>); Im(e - <e>)]: the column
// GEMV of the MinSR solve, so O is read once.
//
// Adam: optax's scale_by_adam followed by scale(-lr), per element of the real ravel.  A complex
// leaf (stored as [re..., im...]) is one complex element for Adam: its first moment is complex and
// its second moment is |g|^2, shared by both parts.  The complex gradient given to Adam is
// g_re - i g_im in mode 0 (jax.grad of a real loss with respect to a complex parameter, which is
// what the reference feeds optax) or g_re + i g_im in mode 1 (a descent step in both parts).
#include "nqs_common.h"
#include "minsr.h"

namespace nqs {

__global__ void k_adam(double* __restrict__ th, const double* __restrict__ g, double* __restrict__ mu,
                       double* __restrict__ nu, const int64_t* __restrict__ partner, int64_t P, double bc1, double bc2,
                       double lr, double b1, double b2, double eps, double eps_root, int mode) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const int64_t pj = partner ? partner[i] : -1;
  if (pj == -2) return;  // imaginary part: updated with its real part
  if (pj < 0) {
    const double gi = g[i];
    const double m = (1.0 - b1) * gi + b1 * mu[i];
    const double v = (1.0 - b2) * (gi * gi) + b2 * nu[i];
    mu[i] = m;
    nu[i] = v;
    const double u = (m / bc1) / (sqrt(v / bc2 + eps_root) + eps);
    th[i] = th[i] + (-lr) * u;
    return;
  }
  const double gr = g[i], gm = mode == 0 ? -g[pj] : g[pj];
  const double mr = (1.0 - b1) * gr + b1 * mu[i];
  const double mi = (1.0 - b1) * gm + b1 * mu[pj];
  const double v = (1.0 - b2) * (gr * gr + gm * gm) + b2 * nu[i];
  mu[i] = mr;
  mu[pj] = mi;
  nu[i] = v;
  nu[pj] = v;
  const double den = sqrt(v / bc2 + eps_root) + eps;
  th[i] = th[i] + (-lr) * ((mr / bc1) / den);
  th[pj] = th[pj] + (-lr) * ((mi / bc1) / den);
}

}  // namespace nqs

using namespace nqs;

extern "C" {

int64_t nqs_vmc_gradient_workspace_bytes(const nqs_minsr_desc* d) {
  if (!d || d->n_walkers <= 0 || d->n_params <= 0 || d->ldo < d->n_params || d->ldo % 4) return -1;
  return round_up(2 * d->n_walkers * 8, 256) + round_up(colsum_part_elems(d->ldo) * 8, 256) + 256;
}

int nqs_vmc_gradient(const nqs_minsr_desc* d, const void* O, const double* e, double* g, void* ws, int64_t ws_bytes,
                     void* stream) {
  const int64_t need = nqs_vmc_gradient_workspace_bytes(d);
  NQS_REQUIRE(need >= 0, NQS_ERR_ARG, "bad descriptor (walkers, params > 0, ldo >= params, ldo % 4 == 0)");
  NQS_REQUIRE(O && e && g && ws, NQS_ERR_ARG, "null pointer");
  NQS_REQUIRE(ws_bytes >= need, NQS_ERR_WORKSPACE, "workspace too small");
  NQS_REQUIRE(d->o_dtype == NQS_FP64 || d->o_dtype == NQS_MIXED, NQS_ERR_ARG, "o_dtype must be fp64 or mixed");
  hipStream_t st = (hipStream_t)stream;
  const int64_t W = d->n_walkers, P = d->n_params;
  char* p = (char*)ws;
  double* v = (double*)p;
  p += round_up(2 * W * 8, 256);
  double* part = (double*)p;
  p += round_up(colsum_part_elems(d->ldo) * 8, 256);
  double* ysum = (double*)p;
  NQS_TRY(make_etil(e, W, v, st));
  NQS_TRY(axpby(2 * W, 2.0 / (double)W, v, 0.0, v, v, st));
  if (d->o_dtype == NQS_FP64)
    return gemv_cols_c<double>((const double*)O, d->ldo, W, P, v, nullptr, part, ysum, g, P, st);
  return gemv_cols_c<float>((const float*)O, d->ldo, W, P, v, nullptr, part, ysum, g, P, st);
}

int nqs_adam_update(double* theta, const double* g, double* mu, double* nu, const int64_t* partner, int64_t P,
                    int64_t count, double lr, double b1, double b2, double eps, double eps_root, int32_t complex_mode,
                    void* stream) {
  NQS_REQUIRE(theta && g && mu && nu && P > 0, NQS_ERR_ARG, "null pointer or empty parameter vector");
  NQS_REQUIRE(count >= 1, NQS_ERR_ARG, "count is the 1-based step number");
  NQS_REQUIRE(complex_mode == 0 || complex_mode == 1, NQS_ERR_ARG, "complex_mode must be 0 (jax) or 1 (descent)");
  NQS_REQUIRE(b1 >= 0 && b1 < 1 && b2 >= 0 && b2 < 1, NQS_ERR_ARG, "decay rates must be in [0, 1)");
  const double bc1 = 1.0 - pow(b1, (double)count), bc2 = 1.0 - pow(b2, (double)count);
  hipLaunchKernelGGL(k_adam, dim3((unsigned)cdiv(P, 256)), dim3(256), 0, (hipStream_t)stream, theta, g, mu, nu, partner,
                     P, bc1, bc2, lr, b1, b2, eps, eps_root, (int)complex_mode);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}

}  // extern "C"
