// extern "C" entry points of libnqs_hip.so (declared in include/nqs_hip.h).
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>
#include "model.h"
#include "physics.h"
#include "minsr.h"
#include "obs.h"

namespace nqs {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
int pack_params(const Model& md, const double* theta, const double* gs, const double* wenv, void* packed,
                hipStream_t st);


// Rows of the jet activations processed per launch sequence: bounds the workspace while
// keeping every launch far above the 256-CU occupancy threshold.
static const int64_t kRowBudget = 1LL << 20;

// Walkers per chunk: the W walkers in equal chunks of at most the row budget, allowed 2% over it
// when that saves a chunk (8192 walkers of the N = 16 spin-flip pass: six chunks of 1366 instead of
// six of 1365 and one of 2).
static int64_t chunk_walkers(const Model& md, int64_t W, int cfg_per_walker, int C) {
  const int64_t per = (int64_t)cfg_per_walker * C * md.N;
  W = std::max<int64_t>(W, 1);
  const int64_t c = std::min<int64_t>(std::max<int64_t>(1, kRowBudget / per), W);
  const int64_t slack = std::max<int64_t>(c, (kRowBudget + kRowBudget / 50) / per);
  const int64_t n = (W + slack - 1) / slack;
  return (W + n - 1) / n;
}

static int64_t al(int64_t b) { return round_up(b, 256); }

struct ElocPlan {
  int64_t c1, c2;  // walkers per chunk in the Laplacian pass / the spin-flip pass
  int64_t fwd, extra;
};
static ElocPlan eloc_plan(const Model& md, int64_t W) {
  ElocPlan p;
  const int N = md.N;
  p.c1 = chunk_walkers(md, W, 1, 2 * N + 2);
  p.c2 = chunk_walkers(md, W, N, 3);
  p.fwd = std::max(fwd_bytes(md, p.c1, 2 * N + 2, 2 * N), fwd_bytes(md, p.c2 * N, 3, 2));
  p.extra = al(W * 16) + al(W * 2LL * N * 16) + al(W * 16) + al(W * N * 16) + al(W * N * 2 * 16);
  return p;
}

template <typename T>
static int fwd_logpsi(const Model& md, const void* packed, const double* R, const int32_t* S, int64_t W, double* out,
                      char* ws, hipStream_t st, int64_t cw) {
  char* p = ws;
  FwdBufs<T> b = carve_fwd<T>(md, p, cw, 1, 0);
  for (int64_t w0 = 0; w0 < W; w0 += cw) {
    JetGeom jg{std::min(cw, W - w0), md.N, 1, 0, 0, 0};
    FwdBufs<T> bb = b;
    bb.lp = out + 2 * w0;
    NQS_TRY(run_forward<T>(md, packed, jg, R, S, w0, bb, st));
  }
  return NQS_OK;
}

template <typename T>
static int eloc_impl(const Model& md, const nqs_hamiltonian_desc* h, const void* packed, const double* R,
                     const int32_t* S, int64_t W, double* eloc, double* lp_out, double* grad_out, double* lap_out,
                     char* ws, hipStream_t st) {
  const int N = md.N;
  const ElocPlan pl = eloc_plan(md, W);
  char* p = ws;
  char* fwd_base = p;
  p += al(pl.fwd);
  double* lp = (double*)p;
  p += al(W * 16);
  double* grad = (double*)p;
  p += al(W * 2LL * N * 16);
  double* lap = (double*)p;
  p += al(W * 16);
  double* lpf = (double*)p;
  p += al(W * N * 16);
  double* tanf = (double*)p;
  p += al(W * N * 2 * 16);
  if (lp_out) lp = lp_out;
  if (grad_out) grad = grad_out;
  if (lap_out) lap = lap_out;
  // pass 1: value, 2N position tangents, Laplacian
  {
    char* q = fwd_base;
    FwdBufs<T> b = carve_fwd<T>(md, q, pl.c1, 2 * N + 2, 2 * N);
    for (int64_t w0 = 0; w0 < W; w0 += pl.c1) {
      const int64_t n = std::min(pl.c1, W - w0);
      JetGeom jg{n, N, 2 * N + 2, 2 * N, 1, 1};
      FwdBufs<T> bb = b;
      bb.lp = lp + 2 * w0;
      bb.tan = grad + 2 * w0 * 2 * N;
      bb.lap = lap + 2 * w0;
      NQS_TRY(run_forward<T>(md, packed, jg, R, S, w0, bb, st));
    }
  }
  // pass 2: every single-spin flip, value + the two tangents of the flipped electron.  Only the
  // Rashba term and the in-plane field components use the flipped configurations: without them
  // (lambda_R = 0, no ext field, no pinning spin fields) their contribution is exactly zero and
  // the pass is skipped.
  const bool flips = h->lambda_r != 0.0 || h->n_field > 0 || h->n_pin > 0;
  if (flips) {
    char* q = fwd_base;
    FwdBufs<T> b = carve_fwd<T>(md, q, pl.c2 * N, 3, 2);
    for (int64_t w0 = 0; w0 < W; w0 += pl.c2) {
      const int64_t n = std::min(pl.c2, W - w0);
      JetGeom jg{n * N, N, 3, 2, 0, 2};
      FwdBufs<T> bb = b;
      bb.lp = lpf + 2 * w0 * N;
      bb.tan = tanf + 2 * w0 * N * 2;
      NQS_TRY(run_forward<T>(md, packed, jg, R, S, w0, bb, st));
    }
  }
  EAssembleArgs a;
  a.W = W;
  a.N = N;
  a.m_star = h->m_star;
  a.lambda_r = h->lambda_r;
  a.h_z = h->h_z;
  a.n_field = h->n_field;
  a.field = h->field_dev;
  a.field_scale = h->field_scale;
  a.n_pin = h->n_pin;
  a.pin = h->pin_dev;
  for (int q = 0; q < 4; ++q) a.A[q] = h->cell[q];
  a.R = R;
  a.S = S;
  a.lp = lp;
  a.grad = grad;
  a.lap = lap;
  a.lpf = flips ? lpf : nullptr;
  a.tanf = flips ? tanf : nullptr;
  a.out = eloc;
  return launch_eloc_assemble(a, st);
}

// Spin estimators: log psi of every walker, then of every single-spin-flipped configuration
// (value only, cfg = w N + i), then the per-electron estimators from the ratios.
static int64_t spin_fwd_bytes(const Model& md, int64_t W) {
  const int64_t c1 = chunk_walkers(md, W, 1, 1), c2 = chunk_walkers(md, W, md.N, 1);
  return std::max(fwd_bytes(md, c1, 1, 0), fwd_bytes(md, c2 * md.N, 1, 0));
}

template <typename T>
static int spin_impl(const Model& md, const void* packed, const double* R, const int32_t* S, int64_t W, double* out,
                     char* ws, hipStream_t st) {
  const int N = md.N;
  const int64_t c1 = chunk_walkers(md, W, 1, 1), c2 = chunk_walkers(md, W, N, 1);
  char* p = ws;
  char* fwd_base = p;
  p += al(spin_fwd_bytes(md, W));
  double* lp = (double*)p;
  p += al(W * 16);
  double* lpf = (double*)p;
  {
    char* q = fwd_base;
    FwdBufs<T> b = carve_fwd<T>(md, q, c1, 1, 0);
    for (int64_t w0 = 0; w0 < W; w0 += c1) {
      JetGeom jg{std::min(c1, W - w0), N, 1, 0, 0, 0};
      FwdBufs<T> bb = b;
      bb.lp = lp + 2 * w0;
      NQS_TRY(run_forward<T>(md, packed, jg, R, S, w0, bb, st));
    }
  }
  {
    char* q = fwd_base;
    FwdBufs<T> b = carve_fwd<T>(md, q, c2 * N, 1, 0);
    for (int64_t w0 = 0; w0 < W; w0 += c2) {
      JetGeom jg{std::min(c2, W - w0) * N, N, 1, 0, 0, 2};
      FwdBufs<T> bb = b;
      bb.lp = lpf + 2 * w0 * N;
      NQS_TRY(run_forward<T>(md, packed, jg, R, S, w0, bb, st));
    }
  }
  return launch_spin_est(W, N, S, lp, lpf, out, st);
}

static int64_t grad_chunk(const Model& md, int64_t W) {
  const int64_t per = std::max<int64_t>(1, grad_bytes(md, 1));
  const int64_t c = std::max<int64_t>(1, (8LL << 30) / per);
  return std::min<int64_t>(c, std::max<int64_t>(W, 1));
}

struct McmcPlan {
  int64_t cw, fwd;
};
static McmcPlan mcmc_plan(const Model& md, int64_t W) {
  McmcPlan p;
  p.cw = chunk_walkers(md, W, 1, 1);
  p.fwd = fwd_bytes(md, p.cw, 1, 0);
  return p;
}

template <typename T>
static int mcmc_impl(const Model& md, const void* packed, double* R, int32_t* S, double* logp, int64_t W, int ns,
                     int nf, double step, const double* xi, const double* u_pos, const int32_t* fi,
                     const double* u_flip, int32_t* acc_pos, int32_t* acc_spin, char* ws, hipStream_t st) {
  const int N = md.N;
  const McmcPlan pl = mcmc_plan(md, W);
  char* p = ws;
  char* fwd_base = p;
  p += al(pl.fwd);
  double* Rp = (double*)p;
  p += al(W * N * 2 * 8);
  int32_t* Sp = (int32_t*)p;
  p += al(W * N * 4);
  double* lpn = (double*)p;
  p += al(W * 16);
  char* q = fwd_base;
  FwdBufs<T> b = carve_fwd<T>(md, q, pl.cw, 1, 0);
  auto forward = [&](const double* RR, const int32_t* SS) -> int {
    for (int64_t w0 = 0; w0 < W; w0 += pl.cw) {
      JetGeom jg{std::min(pl.cw, W - w0), N, 1, 0, 0, 0};
      FwdBufs<T> bb = b;
      bb.lp = lpn + 2 * w0;
      NQS_TRY(run_forward<T>(md, packed, jg, RR, SS, w0, bb, st));
    }
    return NQS_OK;
  };
  for (int s = 0; s < ns; ++s) {
    NQS_TRY(launch_mcmc_propose(W, N, R, xi + (int64_t)s * W * N * 2, step, md.A, md.Ainv, Rp, st));
    NQS_TRY(forward(Rp, S));
    NQS_TRY(launch_mcmc_accept_pos(W, N, R, Rp, logp, lpn, u_pos + (int64_t)s * W, acc_pos, st));
    for (int f = 0; f < nf; ++f) {
      const int32_t* fif = fi + ((int64_t)s * nf + f) * W;
      NQS_TRY(launch_mcmc_flip_propose(W, N, S, fif, Sp, st));
      NQS_TRY(forward(R, Sp));
      NQS_TRY(launch_mcmc_accept_spin(W, N, S, fif, logp, lpn, u_flip + ((int64_t)s * nf + f) * W, acc_spin, st));
    }
  }
  return NQS_OK;
}

struct MinsrWs {
  double* mean;
  __bf16* Ob;
  float* T32;
  double* A64;
  double* Linv;
  double* panel;
  double *et, *y, *r, *u, *v, *ts, *sc;
  double* part;  // partial column sums (colsum_part_elems)
  int32_t* tsync;  // the chained triangular solve's counters and flags (chol_sync_elems)
  double* red;   // ysum[2], mdot[2]
  int64_t bytes;
};
static MinsrWs minsr_ws(const nqs_minsr_desc* d, char* base) {
  MinsrWs w;
  const int64_t n = 2 * d->n_walkers;
  char* p = base;
  auto grab = [&](int64_t b) {
    char* r = p;
    p += al(b);
    return r;
  };
  w.mean = (double*)grab(2 * d->n_params * 8);
  w.Ob = d->gram == NQS_GRAM_BF16 ? (__bf16*)grab(n * d->ldo * 2) : nullptr;
  w.T32 = d->gram != NQS_GRAM_FP64 ? (float*)grab(n * n * 4) : nullptr;
  w.A64 = (double*)grab(n * n * 8);
  w.Linv = (double*)grab(n * 64 * 8);
  w.panel = (double*)grab(cholesky_panel_elems(n) * 8);
  w.et = (double*)grab(n * 8);
  w.y = (double*)grab(n * 8);
  w.r = (double*)grab(n * 8);
  w.u = (double*)grab(d->ldo * 8);
  w.v = (double*)grab(n * 8);
  w.ts = (double*)grab(defl_ts_elems(n) * 8);
  w.sc = (double*)grab(8 * 8);
  w.part = (double*)grab(colsum_part_elems(d->ldo) * 8);
  w.red = (double*)grab(4 * 8);
  w.tsync = (int32_t*)grab(chol_sync_elems(n) * 4);
  w.bytes = p - base;
  return w;
}

// Optional stage timing of the MinSR calls: hipEvents recorded on each call's stream: 0 start,
// 1 centred, 2 Gram shifted and deflated, 3 factorised, 6 solve started, 4 solved and refined,
// 5 pushed through.
static bool g_minsr_profile = false;
static hipEvent_t g_minsr_ev[7];
static bool g_minsr_ev_ok = false;
static int minsr_mark(int i, hipStream_t st) {
  if (!g_minsr_profile) return NQS_OK;
  if (!g_minsr_ev_ok) {
    for (int k = 0; k < 7; ++k) NQS_CHECK_HIP(hipEventCreate(&g_minsr_ev[k]));
    g_minsr_ev_ok = true;
  }
  NQS_CHECK_HIP(hipEventRecord(g_minsr_ev[i], st));
  return NQS_OK;
}

// The MinSR direction in three stages on one workspace: the centred Gram shifted by lam and
// deflated (minsr_gram), its Cholesky factor (minsr_factor, the only stage that writes info), and
// the solve for the local energies with refinement and the push-through (minsr_solve).
template <typename T>
static int minsr_gram(const nqs_minsr_desc* d, T* O, char* ws, hipStream_t st) {
  const int64_t W = d->n_walkers, n = 2 * W, P = d->n_params, ldo = d->ldo;
  MinsrWs w = minsr_ws(d, ws);
  NQS_TRY(minsr_mark(0, st));
  // the bf16 Gram reads only the bf16 copy of O shifted by the mean of the first kBf16MeanRows
  // walkers, so O stays as given and the GEMVs of the solve apply the shift on the fly (the
  // deflation and the residual's projection make the result that of exact centring); the float64 /
  // float32 Gram reads O, which is centred in place with the mean over all walkers
  const bool keep = d->gram == NQS_GRAM_BF16;
  NQS_TRY(minsr_center<T>(O, ldo, W, P, w.mean, w.Ob, keep ? 0 : 1, keep ? std::min(W, kBf16MeanRows) : W, w.part,
                          st));
  NQS_TRY(minsr_mark(1, st));
  if (d->gram == NQS_GRAM_BF16) {
    // A64 is written only after the Gram, so it lends its space to the Gram's partials
    NQS_TRY(gram_bf16(w.Ob, ldo, n, ldo, w.T32, n, (float*)w.A64, 2 * n * n, st));
    NQS_TRY(to_f64_shift<float>(w.T32, n, n, 0.0, w.A64, n, st));
  } else {
    GemmArgs<T> g;
    g.A = O;
    g.B = O;
    g.M = n;
    g.N = n;
    g.K = P;
    g.lda = g.ldb = ldo;
    g.strideA = g.strideB = g.strideC = 0;
    g.alpha = (T)1;
    g.beta = (T)0;
    g.bias = nullptr;
    g.bias_period = 1;
    g.bias_len = 0;
    g.tri = 1;
    g.mirror = 0;
    g.split = 0;
    if (sizeof(T) == 8) {
      g.C = (T*)w.A64;
      g.ldc = n;
      NQS_TRY(gemm_nt<T>(g, 1, st));
    } else {
      g.C = (T*)w.T32;
      g.ldc = n;
      NQS_TRY(gemm_nt<T>(g, 1, st));
      NQS_TRY(to_f64_shift<float>(w.T32, n, n, 0.0, w.A64, n, st));
    }
  }
  NQS_TRY(deflate_centring(w.A64, n, W, d->lam, w.ts, w.sc, st));
  NQS_TRY(minsr_mark(2, st));
  return NQS_OK;
}

static int minsr_factor(const nqs_minsr_desc* d, int32_t* info, char* ws, hipStream_t st) {
  const int64_t n = 2 * d->n_walkers;
  MinsrWs w = minsr_ws(d, ws);
  NQS_TRY(fill_i32(info, 1, 0, st));
  NQS_TRY(cholesky(w.A64, n, n, w.Linv, w.panel, info, st));
  NQS_TRY(minsr_mark(3, st));
  return NQS_OK;
}

template <typename T>
static int minsr_solve(const nqs_minsr_desc* d, T* O, const double* eps, double* dth, char* ws, hipStream_t st) {
  const int64_t W = d->n_walkers, n = 2 * W, P = d->n_params, ldo = d->ldo;
  MinsrWs w = minsr_ws(d, ws);
  const double* mc = d->gram == NQS_GRAM_BF16 ? w.mean : nullptr;
  NQS_TRY(minsr_mark(6, st));
  NQS_TRY(make_etil(eps, W, w.et, st));
  NQS_CHECK_HIP(hipMemcpyAsync(w.y, w.et, n * 8, hipMemcpyDeviceToDevice, st));
  NQS_TRY(chol_solve(w.A64, n, n, w.Linv, w.y, w.tsync, st));
  for (int it = 0; it < d->refine; ++it) {
    // r = etil - Otil (Otil^T y) - lam y ;  y += (L L^T)^-1 r
    NQS_TRY(gemv_cols_c<T>(O, ldo, W, P, w.y, mc, w.part, w.red, w.u, ldo, st));
    NQS_TRY(gemv_rows_c<T>(O, ldo, W, P, w.u, mc, w.red + 2, w.v, st));
    NQS_TRY(axpby(n, 1.0, w.et, -1.0, w.v, w.r, st));
    NQS_TRY(axpby(n, 1.0, w.r, -d->lam, w.y, w.r, st));
    // with the shifted O (bf16 Gram), v carries the shift's constant per half: r <- Pi r
    if (mc) NQS_TRY(proj_halves(w.r, W, st));
    NQS_TRY(chol_solve(w.A64, n, n, w.Linv, w.r, w.tsync, st));
    NQS_TRY(axpby(n, 1.0, w.y, 1.0, w.r, w.y, st));
  }
  NQS_TRY(minsr_mark(4, st));
  NQS_TRY(gemv_cols_c<T>(O, ldo, W, P, w.y, mc, w.part, w.red, dth, P, st));
  NQS_TRY(minsr_mark(5, st));
  return NQS_OK;
}

template <typename T>
static int minsr_impl(const nqs_minsr_desc* d, T* O, const double* eps, double* dth, int32_t* info, char* ws,
                      hipStream_t st) {
  NQS_TRY(minsr_gram<T>(d, O, ws, st));
  NQS_TRY(minsr_factor(d, info, ws, st));
  return minsr_solve<T>(d, O, eps, dth, ws, st);
}

// ------------------------------------------------------------------ MinSR of many same-size systems
// Systems share n = 2W.  Per-system buffers (column means, bf16 operand, GEMV scratch) are carved
// one after another; the n x n matrices, inverse blocks, panels and right-hand sides of all
// systems sit at uniform strides so that the Cholesky and the triangular solves take every
// system in each launch, and the bf16 Grams of all systems are one launch.
struct MinsrSysWs {
  double *mean, *et, *u, *v, *ts, *sc, *part, *red;
  __bf16* Ob;
};
struct MinsrBatchWs {
  double *A64, *Linv, *panel, *Y, *Rr;
  float* T32;
  int32_t* tsync;
  std::vector<MinsrSysWs> sys;
  int64_t bytes;
};
static MinsrBatchWs minsr_batch_ws(const nqs_minsr_desc* d, int nsys, char* base) {
  MinsrBatchWs w;
  const int64_t n = 2 * d[0].n_walkers;
  char* p = base;
  auto grab = [&](int64_t b) {
    char* r = p;
    p += al(b);
    return r;
  };
  w.A64 = (double*)grab(nsys * n * n * 8);
  w.Linv = (double*)grab(nsys * n * 64 * 8);
  w.panel = (double*)grab(nsys * cholesky_panel_elems(n) * 8);
  w.Y = (double*)grab(nsys * n * 8);
  w.Rr = (double*)grab(nsys * n * 8);
  w.T32 = d[0].gram != NQS_GRAM_FP64 ? (float*)grab(nsys * n * n * 4) : nullptr;
  w.tsync = (int32_t*)grab(nsys * chol_sync_elems(n) * 4);
  for (int s = 0; s < nsys; ++s) {
    MinsrSysWs q;
    q.mean = (double*)grab(2 * d[s].n_params * 8);
    q.Ob = d[s].gram == NQS_GRAM_BF16 ? (__bf16*)grab(n * d[s].ldo * 2) : nullptr;
    q.et = (double*)grab(n * 8);
    q.u = (double*)grab(d[s].ldo * 8);
    q.v = (double*)grab(n * 8);
    q.ts = (double*)grab(defl_ts_elems(n) * 8);
    q.sc = (double*)grab(8 * 8);
    q.part = (double*)grab(colsum_part_elems(d[s].ldo) * 8);
    q.red = (double*)grab(4 * 8);
    w.sys.push_back(q);
  }
  w.bytes = p - base;
  return w;
}

template <typename T>
static int minsr_batched_impl(const nqs_minsr_desc* d, int nsys, void* const* Ov, const double* const* eps,
                              double* const* dth, int32_t* info, char* ws, hipStream_t st) {
  const int64_t W = d[0].n_walkers, n = 2 * W;
  MinsrBatchWs w = minsr_batch_ws(d, nsys, ws);
  NQS_TRY(fill_i32(info, nsys, 0, st));
  const bool keep = d[0].gram == NQS_GRAM_BF16;
  for (int s = 0; s < nsys; ++s) {
    const MinsrSysWs& q = w.sys[s];
    NQS_TRY(minsr_center<T>((T*)Ov[s], d[s].ldo, W, d[s].n_params, q.mean, q.Ob, keep ? 0 : 1,
                            keep ? std::min(W, kBf16MeanRows) : W, q.part, st));
    NQS_TRY(make_etil(eps[s], W, q.et, st));
  }
  if (keep) {
    for (int s0 = 0; s0 < nsys; s0 += kGramBatch) {
      const int cnt = std::min(kGramBatch, nsys - s0);
      GramBatch gb;
      for (int k = 0; k < cnt; ++k) {
        gb.A[k] = w.sys[s0 + k].Ob;
        gb.K[k] = d[s0 + k].ldo;
        gb.C[k] = w.T32 + (s0 + k) * n * n;
      }
      NQS_TRY(gram_bf16_batched(gb, cnt, n, n, st));
    }
  }
  for (int s = 0; s < nsys; ++s) {
    double* A = w.A64 + s * n * n;
    if (keep) {
      NQS_TRY(to_f64_shift<float>(w.T32 + s * n * n, n, n, 0.0, A, n, st));
    } else {
      GemmArgs<T> g;
      g.A = g.B = (const T*)Ov[s];
      g.M = g.N = n;
      g.K = d[s].n_params;
      g.lda = g.ldb = d[s].ldo;
      g.strideA = g.strideB = g.strideC = 0;
      g.alpha = (T)1;
      g.beta = (T)0;
      g.bias = nullptr;
      g.bias_period = 1;
      g.bias_len = 0;
      g.tri = 1;
      g.mirror = 0;
      g.split = 0;
      g.ldc = n;
      if (sizeof(T) == 8) {
        g.C = (T*)A;
        NQS_TRY(gemm_nt<T>(g, 1, st));
      } else {
        g.C = (T*)(w.T32 + s * n * n);
        NQS_TRY(gemm_nt<T>(g, 1, st));
        NQS_TRY(to_f64_shift<float>(w.T32 + s * n * n, n, n, 0.0, A, n, st));
      }
    }
    NQS_TRY(deflate_centring(A, n, W, d[s].lam, w.sys[s].ts, w.sys[s].sc, st));
  }
  NQS_TRY(cholesky(w.A64, n, n, w.Linv, w.panel, info, st, nsys, n * n, n * 64, cholesky_panel_elems(n)));
  for (int s = 0; s < nsys; ++s)
    NQS_CHECK_HIP(hipMemcpyAsync(w.Y + s * n, w.sys[s].et, n * 8, hipMemcpyDeviceToDevice, st));
  NQS_TRY(chol_solve(w.A64, n, n, w.Linv, w.Y, w.tsync, st, nsys, n * n, n * 64, n));
  for (int it = 0; it < d[0].refine; ++it) {
    for (int s = 0; s < nsys; ++s) {
      const MinsrSysWs& q = w.sys[s];
      const double* mc = keep ? q.mean : nullptr;
      const T* O = (const T*)Ov[s];
      double* y = w.Y + s * n;
      double* r = w.Rr + s * n;
      NQS_TRY(gemv_cols_c<T>(O, d[s].ldo, W, d[s].n_params, y, mc, q.part, q.red, q.u, d[s].ldo, st));
      NQS_TRY(gemv_rows_c<T>(O, d[s].ldo, W, d[s].n_params, q.u, mc, q.red + 2, q.v, st));
      NQS_TRY(axpby(n, 1.0, q.et, -1.0, q.v, r, st));
      NQS_TRY(axpby(n, 1.0, r, -d[s].lam, y, r, st));
      if (mc) NQS_TRY(proj_halves(r, W, st));
    }
    NQS_TRY(chol_solve(w.A64, n, n, w.Linv, w.Rr, w.tsync, st, nsys, n * n, n * 64, n));
    NQS_TRY(axpby(nsys * n, 1.0, w.Y, 1.0, w.Rr, w.Y, st));
  }
  for (int s = 0; s < nsys; ++s) {
    const MinsrSysWs& q = w.sys[s];
    NQS_TRY(gemv_cols_c<T>((const T*)Ov[s], d[s].ldo, W, d[s].n_params, w.Y + s * n, keep ? q.mean : nullptr, q.part,
                           q.red, dth[s], d[s].n_params, st));
  }
  return NQS_OK;
}

static int check_minsr_desc(const nqs_minsr_desc* d) {
  NQS_REQUIRE(d->n_walkers > 0 && d->n_params > 0 && d->ldo >= d->n_params, NQS_ERR_ARG, "bad MinSR descriptor");
  NQS_REQUIRE(d->ldo % 64 == 0, NQS_ERR_ARG, "ldo must be a multiple of 64");
  NQS_REQUIRE((2 * d->n_walkers) % 64 == 0, NQS_ERR_UNSUPPORTED, "2W must be a multiple of 64");
  NQS_REQUIRE(d->o_dtype == NQS_FP64 || d->o_dtype == NQS_MIXED, NQS_ERR_ARG, "bad O dtype");
  NQS_REQUIRE(!(d->gram == NQS_GRAM_FP64 && d->o_dtype != NQS_FP64), NQS_ERR_ARG, "a float64 Gram needs a float64 O");
  NQS_REQUIRE(!(d->gram == NQS_GRAM_FP32 && d->o_dtype != NQS_MIXED), NQS_ERR_ARG, "a float32 Gram needs a float32 O");
  NQS_REQUIRE(d->gram >= 0 && d->gram <= 2 && d->refine >= 0, NQS_ERR_ARG, "bad Gram mode");
  return NQS_OK;
}

}  // namespace nqs

using namespace nqs;

extern "C" {

int nqs_version(void) { return 1; }
int64_t nqs_desc_size(int32_t which) {
  switch (which) {
    case 0: return (int64_t)sizeof(nqs_model_desc);
    case 1: return (int64_t)sizeof(nqs_hamiltonian_desc);
    case 2: return (int64_t)sizeof(nqs_ewald_desc);
    case 3: return (int64_t)sizeof(nqs_minsr_desc);
    default: return -1;
  }
}
const char* nqs_last_error(void) { return g_last_error.c_str(); }

int64_t nqs_param_count(const nqs_model_desc* m) {
  Model md;
  if (build_model(m, md) != NQS_OK) return -1;
  return md.P;
}

int nqs_param_leaf(const nqs_model_desc* m, int32_t idx, char* name, int32_t name_len, int64_t shape[3],
                   int32_t* ndim, int64_t* offset, int32_t* is_complex) {
  Model md;
  NQS_TRY(build_model(m, md));
  NQS_REQUIRE(idx >= 0 && idx < (int32_t)md.leaves.size(), NQS_ERR_ARG, "leaf index out of range");
  const Leaf& l = md.leaves[idx];
  if (name && name_len > 0) {
    std::strncpy(name, l.name.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  for (int i = 0; i < 3; ++i) shape[i] = l.shape[i];
  *ndim = l.ndim;
  *offset = l.offset;
  *is_complex = l.cplx ? 1 : 0;
  return NQS_OK;
}

int64_t nqs_packed_bytes(const nqs_model_desc* m) {
  Model md;
  if (build_model(m, md) != NQS_OK) return -1;
  return md.packed_bytes;
}

int nqs_pack_params(const nqs_model_desc* m, const double* theta, const double* gs, const double* wenv, void* packed,
                    void* stream) {
  Model md;
  NQS_TRY(build_model(m, md));
  NQS_REQUIRE(theta && gs && wenv && packed, NQS_ERR_ARG, "null pointer");
  return pack_params(md, theta, gs, wenv, packed, (hipStream_t)stream);
}

int64_t nqs_workspace_bytes(const nqs_model_desc* m, int32_t op, int64_t W) {
  Model md;
  if (build_model(m, md) != NQS_OK || W < 0) return -1;
  switch (op) {
    case NQS_OP_LOGPSI:
      return fwd_bytes(md, chunk_walkers(md, W, 1, 1), 1, 0);
    case NQS_OP_ELOC: {
      ElocPlan p = eloc_plan(md, W);
      return al(p.fwd) + p.extra;
    }
    case NQS_OP_GRAD:
      return grad_bytes(md, grad_chunk(md, W));
    case NQS_OP_VJP: {
      const int64_t c = grad_chunk(md, W);
      return grad_bytes(md, c) + vjp_bytes(md, c);
    }
    case NQS_OP_MCMC: {
      McmcPlan p = mcmc_plan(md, W);
      return al(p.fwd) + al(W * md.N * 16) + al(W * md.N * 4) + al(W * 16);
    }
    case NQS_OP_SPIN:
      return al(spin_fwd_bytes(md, W)) + al(W * 16) + al(W * md.N * 16);
    default:
      set_error("unknown op");
      return -1;
  }
}

static int check_ws(const nqs_model_desc* m, int op, int64_t W, int64_t ws_bytes) {
  const int64_t need = nqs_workspace_bytes(m, op, W);
  NQS_REQUIRE(need >= 0, NQS_ERR_ARG, "bad descriptor");
  NQS_REQUIRE(ws_bytes >= need, NQS_ERR_WORKSPACE,
              "workspace too small: need " + std::to_string(need) + " bytes, got " + std::to_string(ws_bytes));
  return NQS_OK;
}

int nqs_psiformer_fwd(const nqs_model_desc* m, const void* packed, const double* R, const int32_t* S, int64_t W,
                      double* logpsi, void* ws, int64_t ws_bytes, void* stream) {
  Model md;
  NQS_TRY(build_model(m, md));
  NQS_REQUIRE(packed && R && S && logpsi && W >= 0, NQS_ERR_ARG, "null pointer or negative walker count");
  NQS_TRY(check_ws(m, NQS_OP_LOGPSI, W, ws_bytes));
  const int64_t cw = chunk_walkers(md, W, 1, 1);
  if (md.tsize == 8) return fwd_logpsi<double>(md, packed, R, S, W, logpsi, (char*)ws, (hipStream_t)stream, cw);
  return fwd_logpsi<float>(md, packed, R, S, W, logpsi, (char*)ws, (hipStream_t)stream, cw);
}

// ---- allocation-free log psi apply.  The activations live in `scratch`, a device array in this
// library's code object (each GPU holds its own copy until the process ends).  Host code never
// passes the array's name to a kernel: its device address is looked up once per device with
// hipGetSymbolAddress (that device current) and kept in g_scratch.  Calls that overlap on one GPU
// share the array and race.
static constexpr int64_t kScratchBytes = 64LL << 20;
static constexpr int kMaxDevices = 64;
static __device__ __align__(256) char scratch[kScratchBytes];
static std::mutex g_scratch_mu;
static char* g_scratch[kMaxDevices] = {};

static int scratch_address(char** out) {
  int dev = -1;
  const hipError_t e = hipGetDevice(&dev);
  NQS_REQUIRE(e == hipSuccess, NQS_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  NQS_REQUIRE(dev >= 0 && dev < kMaxDevices, NQS_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    NQS_REQUIRE(r == hipSuccess && p, NQS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(r));
    g_scratch[dev] = (char*)p;
  }
  *out = g_scratch[dev];
  return NQS_OK;
}

int nqs_psiformer_fwd_noalloc(const nqs_model_desc* m, const void* packed, const double* R, const int32_t* S,
                              int64_t W, double* logpsi, void* stream) {
  Model md;
  NQS_TRY(build_model(m, md));
  NQS_REQUIRE(packed && R && S && logpsi && W >= 0, NQS_ERR_ARG, "null pointer or negative walker count");
  // the largest walker chunk whose activations fit the scratch array
  int64_t lo = 0, hi = chunk_walkers(md, W, 1, 1);
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) / 2;
    if (fwd_bytes(md, mid, 1, 0) <= kScratchBytes) lo = mid;
    else hi = mid - 1;
  }
  NQS_REQUIRE(lo >= 1, NQS_ERR_WORKSPACE,
              "nqs_psiformer_fwd_noalloc: one walker's activations need " + std::to_string(fwd_bytes(md, 1, 1, 0)) +
                  " bytes, beyond the library's " + std::to_string(kScratchBytes) + "-byte scratch");
  if (W == 0) return NQS_OK;
  char* ws = nullptr;
  NQS_TRY(scratch_address(&ws));
  if (md.tsize == 8) return fwd_logpsi<double>(md, packed, R, S, W, logpsi, ws, (hipStream_t)stream, lo);
  return fwd_logpsi<float>(md, packed, R, S, W, logpsi, ws, (hipStream_t)stream, lo);
}

int nqs_eloc(const nqs_model_desc* m, const nqs_hamiltonian_desc* h, const void* packed, const double* R,
             const int32_t* S, int64_t W, double* eloc, double* lp, double* grad, double* lap, void* ws,
             int64_t ws_bytes, void* stream) {
  Model md;
  NQS_TRY(build_model(m, md));
  NQS_REQUIRE(h && packed && R && S && eloc && W >= 0, NQS_ERR_ARG, "null pointer or negative walker count");
  NQS_REQUIRE(h->n_field == 0 || h->field_dev, NQS_ERR_ARG, "ext field terms without data");
  NQS_REQUIRE(h->n_pin >= 0 && (h->n_pin == 0 || h->pin_dev), NQS_ERR_ARG, "pinning centres without data");
  NQS_REQUIRE(h->field_scale >= 0.0 && h->field_scale < HUGE_VAL, NQS_ERR_ARG, "field_scale must be finite and >= 0");
  NQS_TRY(check_ws(m, NQS_OP_ELOC, W, ws_bytes));
  if (md.tsize == 8)
    return eloc_impl<double>(md, h, packed, R, S, W, eloc, lp, grad, lap, (char*)ws, (hipStream_t)stream);
  return eloc_impl<float>(md, h, packed, R, S, W, eloc, lp, grad, lap, (char*)ws, (hipStream_t)stream);
}

int nqs_spin_estimators(const nqs_model_desc* m, const void* packed, const double* R, const int32_t* S, int64_t W,
                        double* spins, void* ws, int64_t ws_bytes, void* stream) {
  Model md;
  NQS_TRY(build_model(m, md));
  NQS_REQUIRE(packed && R && S && spins && W >= 0, NQS_ERR_ARG, "null pointer or negative walker count");
  NQS_TRY(check_ws(m, NQS_OP_SPIN, W, ws_bytes));
  if (md.tsize == 8) return spin_impl<double>(md, packed, R, S, W, spins, (char*)ws, (hipStream_t)stream);
  return spin_impl<float>(md, packed, R, S, W, spins, (char*)ws, (hipStream_t)stream);
}

// g = 0 by a kernel (no memset node, so the call may be captured into a graph)
__global__ void k_zero_f64(double* __restrict__ g, int64_t n) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) g[q] = 0.0;
}

int nqs_psiformer_vjp(const nqs_model_desc* m, const void* packed, const double* R, const int32_t* S, int64_t W,
                      const double* cot, double* g, double* logpsi, void* ws, int64_t ws_bytes, void* stream) {
  Model md;
  NQS_TRY(build_model(m, md));
  NQS_REQUIRE(packed && R && S && cot && g && W >= 0, NQS_ERR_ARG, "null pointer or negative walker count");
  NQS_TRY(check_ws(m, NQS_OP_VJP, W, ws_bytes));
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)cdiv(std::max<int64_t>(md.P, 1), 256)), dim3(256), 0, st, g, md.P);
  NQS_LAUNCH_CHECK();
  const int64_t cw = grad_chunk(md, W);
  for (int64_t w0 = 0; w0 < W; w0 += cw) {
    const int64_t n = std::min(cw, W - w0);
    double* lp = logpsi ? logpsi + 2 * w0 : nullptr;
    if (md.tsize == 8)
      NQS_TRY(run_backward<double>(md, packed, R, S, w0, n, W, nullptr, 0, lp, (char*)ws, st, cot, g));
    else
      NQS_TRY(run_backward<float>(md, packed, R, S, w0, n, W, nullptr, 0, lp, (char*)ws, st, cot, g));
  }
  return NQS_OK;
}

int nqs_psiformer_grad(const nqs_model_desc* m, const void* packed, const double* R, const int32_t* S, int64_t W,
                       void* O, int64_t ldo, double* logpsi, void* ws, int64_t ws_bytes, void* stream) {
  Model md;
  NQS_TRY(build_model(m, md));
  NQS_REQUIRE(packed && R && S && O && W >= 0, NQS_ERR_ARG, "null pointer or negative walker count");
  NQS_REQUIRE(ldo >= md.P, NQS_ERR_ARG, "ldo smaller than the parameter count");
  NQS_TRY(check_ws(m, NQS_OP_GRAD, W, ws_bytes));
  const int64_t cw = grad_chunk(md, W);
  for (int64_t w0 = 0; w0 < W; w0 += cw) {
    const int64_t n = std::min(cw, W - w0);
    double* lp = logpsi ? logpsi + 2 * w0 : nullptr;
    if (md.tsize == 8)
      NQS_TRY(run_backward<double>(md, packed, R, S, w0, n, W, (double*)O, ldo, lp, (char*)ws, (hipStream_t)stream));
    else
      NQS_TRY(run_backward<float>(md, packed, R, S, w0, n, W, (float*)O, ldo, lp, (char*)ws, (hipStream_t)stream));
  }
  return NQS_OK;
}

int64_t nqs_minsr_workspace_bytes(const nqs_minsr_desc* d) {
  if (!d || d->n_walkers <= 0 || d->n_params <= 0 || d->ldo < d->n_params) {
    set_error("bad MinSR descriptor");
    return -1;
  }
  return minsr_ws(d, nullptr).bytes;
}

// the checks every single-system MinSR call makes on its descriptor and workspace
static int minsr_check(const nqs_minsr_desc* d, int64_t ws_bytes) {
  const int64_t need = nqs_minsr_workspace_bytes(d);
  NQS_REQUIRE(need >= 0, NQS_ERR_ARG, "bad MinSR descriptor");
  NQS_REQUIRE(ws_bytes >= need, NQS_ERR_WORKSPACE, "MinSR workspace too small: need " + std::to_string(need));
  NQS_REQUIRE(d->ldo % 64 == 0, NQS_ERR_ARG, "ldo must be a multiple of 64");
  NQS_REQUIRE((2 * d->n_walkers) % 64 == 0, NQS_ERR_UNSUPPORTED, "2W must be a multiple of 64");
  NQS_REQUIRE(d->o_dtype == NQS_FP64 || d->o_dtype == NQS_MIXED, NQS_ERR_ARG, "bad O dtype");
  NQS_REQUIRE(!(d->gram == NQS_GRAM_FP64 && d->o_dtype != NQS_FP64), NQS_ERR_ARG, "a float64 Gram needs a float64 O");
  NQS_REQUIRE(!(d->gram == NQS_GRAM_FP32 && d->o_dtype != NQS_MIXED), NQS_ERR_ARG, "a float32 Gram needs a float32 O");
  NQS_REQUIRE(d->gram >= 0 && d->gram <= 2 && d->refine >= 0, NQS_ERR_ARG, "bad Gram mode");
  return NQS_OK;
}

int nqs_minsr_gram(const nqs_minsr_desc* d, void* O, void* ws, int64_t ws_bytes, void* stream) {
  NQS_REQUIRE(d && O && ws, NQS_ERR_ARG, "null pointer");
  NQS_TRY(minsr_check(d, ws_bytes));
  hipStream_t st = (hipStream_t)stream;
  if (d->o_dtype == NQS_FP64) return minsr_gram<double>(d, (double*)O, (char*)ws, st);
  return minsr_gram<float>(d, (float*)O, (char*)ws, st);
}

int nqs_minsr_factor(const nqs_minsr_desc* d, int32_t* info, void* ws, int64_t ws_bytes, void* stream) {
  NQS_REQUIRE(d && info && ws, NQS_ERR_ARG, "null pointer");
  NQS_TRY(minsr_check(d, ws_bytes));
  return minsr_factor(d, info, (char*)ws, (hipStream_t)stream);
}

int nqs_minsr_solve(const nqs_minsr_desc* d, void* O, const double* eps, double* dth, void* ws, int64_t ws_bytes,
                    void* stream) {
  NQS_REQUIRE(d && O && eps && dth && ws, NQS_ERR_ARG, "null pointer");
  NQS_TRY(minsr_check(d, ws_bytes));
  hipStream_t st = (hipStream_t)stream;
  if (d->o_dtype == NQS_FP64) return minsr_solve<double>(d, (double*)O, eps, dth, (char*)ws, st);
  return minsr_solve<float>(d, (float*)O, eps, dth, (char*)ws, st);
}

int nqs_minsr_direction(const nqs_minsr_desc* d, void* O, const double* eps, double* dth, int32_t* info, void* ws,
                        int64_t ws_bytes, void* stream) {
  NQS_REQUIRE(d && O && eps && dth && info && ws, NQS_ERR_ARG, "null pointer");
  NQS_TRY(minsr_check(d, ws_bytes));
  hipStream_t st = (hipStream_t)stream;
  if (d->o_dtype == NQS_FP64) return minsr_impl<double>(d, (double*)O, eps, dth, info, (char*)ws, st);
  return minsr_impl<float>(d, (float*)O, eps, dth, info, (char*)ws, st);
}

int64_t nqs_minsr_batched_workspace_bytes(const nqs_minsr_desc* d, int32_t n_sys) {
  if (!d || n_sys < 1) {
    set_error("need at least one MinSR descriptor");
    return -1;
  }
  for (int s = 0; s < n_sys; ++s) {
    if (check_minsr_desc(d + s) != NQS_OK) return -1;
    if (d[s].n_walkers != d[0].n_walkers || d[s].gram != d[0].gram || d[s].o_dtype != d[0].o_dtype ||
        d[s].refine != d[0].refine) {
      set_error("batched MinSR systems must share walkers, Gram mode, O dtype and refinement count");
      return -1;
    }
  }
  if (d[0].gram == NQS_GRAM_BF16 && (2 * d[0].n_walkers) % 256 != 0) {
    set_error("the batched bf16 Gram needs 2W to be a multiple of 256");
    return -1;
  }
  return minsr_batch_ws(d, n_sys, nullptr).bytes;
}

int nqs_minsr_direction_batched(const nqs_minsr_desc* d, int32_t n_sys, void* const* O, const double* const* eps,
                                double* const* dth, int32_t* info, void* ws, int64_t ws_bytes, void* stream) {
  NQS_REQUIRE(d && O && eps && dth && info && ws && n_sys >= 1, NQS_ERR_ARG, "null pointer or no systems");
  for (int s = 0; s < n_sys; ++s) NQS_REQUIRE(O[s] && eps[s] && dth[s], NQS_ERR_ARG, "null system pointer");
  const int64_t need = nqs_minsr_batched_workspace_bytes(d, n_sys);
  NQS_REQUIRE(need >= 0, NQS_ERR_ARG, g_last_error);
  NQS_REQUIRE(ws_bytes >= need, NQS_ERR_WORKSPACE, "batched MinSR workspace too small: need " + std::to_string(need));
  hipStream_t st = (hipStream_t)stream;
  if (d[0].o_dtype == NQS_FP64) return minsr_batched_impl<double>(d, n_sys, O, eps, dth, info, (char*)ws, st);
  return minsr_batched_impl<float>(d, n_sys, O, eps, dth, info, (char*)ws, st);
}

int nqs_minsr_profile(int32_t enable, float* ms_out) {
  if (enable >= 0) g_minsr_profile = enable != 0;
  if (ms_out) {
    NQS_REQUIRE(g_minsr_ev_ok, NQS_ERR_ARG, "no profiled MinSR call yet");
    NQS_CHECK_HIP(hipEventSynchronize(g_minsr_ev[5]));
    // centre 0-1, Gram 1-2, Cholesky 2-3, solve + refinement 6-4, push-through 4-5
    const int from[5] = {0, 1, 2, 6, 4}, to[5] = {1, 2, 3, 4, 5};
    for (int k = 0; k < 5; ++k)
      NQS_CHECK_HIP(hipEventElapsedTime(&ms_out[k], g_minsr_ev[from[k]], g_minsr_ev[to[k]]));
  }
  return NQS_OK;
}

int nqs_clip_energies(const double* e, int64_t W, double clip_k, double* e_clip, double* stats, void* stream) {
  NQS_REQUIRE(e && e_clip && stats, NQS_ERR_ARG, "null pointer");
  return clip_stats(e, W, clip_k, e_clip, stats, (hipStream_t)stream);
}

int nqs_ewald(const nqs_ewald_desc* e, const double* R, int64_t W, int32_t N, double* out, void* stream) {
  NQS_REQUIRE(e && R && out && W >= 0, NQS_ERR_ARG, "null pointer or negative walker count");
  NQS_REQUIRE(e->n_img >= 0 && e->n_g >= 0 && (e->n_img == 0 || e->img_dev) && (e->n_g == 0 || (e->g_dev && e->rg_dev)),
              NQS_ERR_ARG, "bad Ewald tables");
  return launch_ewald(*e, R, W, N, out, (hipStream_t)stream);
}

int nqs_mcmc_sweeps(const nqs_model_desc* m, const void* packed, double* R, int32_t* S, double* logp, int64_t W,
                    int32_t n_sweeps, int32_t n_flip, double step, const double* xi, const double* u_pos,
                    const int32_t* flip_i, const double* u_flip, int32_t* acc_pos, int32_t* acc_spin, void* ws,
                    int64_t ws_bytes, void* stream) {
  Model md;
  NQS_TRY(build_model(m, md));
  NQS_REQUIRE(packed && R && S && logp && W >= 0 && n_sweeps >= 0 && n_flip >= 0, NQS_ERR_ARG, "bad arguments");
  NQS_REQUIRE(n_sweeps == 0 || (xi && u_pos), NQS_ERR_ARG, "missing position streams");
  NQS_REQUIRE(n_sweeps == 0 || n_flip == 0 || (flip_i && u_flip), NQS_ERR_ARG, "missing spin streams");
  NQS_TRY(check_ws(m, NQS_OP_MCMC, W, ws_bytes));
  if (md.tsize == 8)
    return mcmc_impl<double>(md, packed, R, S, logp, W, n_sweeps, n_flip, step, xi, u_pos, flip_i, u_flip, acc_pos,
                             acc_spin, (char*)ws, (hipStream_t)stream);
  return mcmc_impl<float>(md, packed, R, S, logp, W, n_sweeps, n_flip, step, xi, u_pos, flip_i, u_flip, acc_pos,
                          acc_spin, (char*)ws, (hipStream_t)stream);
}

}  // extern "C"
