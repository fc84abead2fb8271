/* C ABI of the gfx950 VMC hot path (libnqs_hip.so).
 *
 * The reference (wenqin-chen/continuum-nqs) is pure Python/JAX: its "interface" for this path is
 * the set of closure factories that the Python host layer (continuum-nqs_amd/nqs/) re-exposes
 * with the same names.  Each entry point below is what one of those factories' returned
 * callables binds to (the ctypes stubs are in continuum-nqs_amd/nqs/_lib.py):
 *
 *   nqs_psiformer_fwd        <- make_logpsi(model, params)(R, S)         src/nqs/psiformer.py:192-196
 *   nqs_eloc                 <- make_local_energy(...)(R, S)             src/nqs/local_energy.py:40-117
 *   nqs_ewald                <- make_ewald[_lattice](...)[2](R)           src/nqs/ewald.py:26-73
 *   nqs_mcmc_sweeps          <- make_mcmc(...) sweep, scanned            src/nqs/sampler.py:25-91
 *   nqs_psiformer_grad       <- per_sample_O / jax.grad of log psi       src/nqs/minsr.py:56-65
 *   nqs_minsr_*              <- minsr_direction(O, eps, lam)             src/nqs/minsr.py:68-81
 *   nqs_param_*              <- real_ravel(params) ordering              src/nqs/minsr.py:30-53
 *
 * Conventions: every pointer argument named *_dev is device memory owned by the caller;
 * complex arrays are interleaved (re, im) float64; positions R are (W, N, 2) float64, spins S
 * are (W, N) int32 in {0, 1}.  Calls are asynchronous on `stream` (a hipStream_t, may be NULL
 * for the default stream), never allocate device memory, and use only the workspace the
 * caller passes (size from nqs_workspace_bytes).  Return value 0 = success, < 0 = error code;
 * nqs_last_error() gives a message for the calling thread.  No exception crosses the ABI.
 */
#ifndef NQS_HIP_H
#define NQS_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  NQS_OK = 0,
  NQS_ERR_ARG = -1,
  NQS_ERR_HIP = -2,
  NQS_ERR_WORKSPACE = -3,
  NQS_ERR_NUMERIC = -4,
  NQS_ERR_UNSUPPORTED = -5
};

/* precision of the backbone (features, attention, MLP, LayerNorm, backflow); the determinant,
 * envelope, Jastrow, Ewald and all linear algebra of MinSR's solve stay float64 in both. */
enum { NQS_FP64 = 0, NQS_MIXED = 1 };

/* how the backbone's dense layers use MFMA in an NQS_MIXED model (activations are float32 in
 * all three, accumulation is float32): BF16X3 splits each float32 operand into a bf16 hi/lo pair
 * and sums three bf16 products; FP32 uses float32 MFMA; BF16 rounds the operands to bf16 once.
 * Ignored for NQS_FP64 models (float64 MFMA). */
enum { NQS_GEMM_BF16X3 = 0, NQS_GEMM_FP32 = 1, NQS_GEMM_BF16 = 2 };

/* ops for nqs_workspace_bytes */
enum { NQS_OP_LOGPSI = 0, NQS_OP_ELOC = 1, NQS_OP_MCMC = 2, NQS_OP_GRAD = 3, NQS_OP_SPIN = 4, NQS_OP_VJP = 5 };

/* Architecture + cell of ContinuumSpinorPsiformer (src/nqs/psiformer.py:100-116). */
typedef struct nqs_model_desc {
  int32_t n_elec;        /* N electrons (= orbitals per determinant) */
  int32_t n_g;           /* plane waves of the spinor envelope */
  int32_t embed_dim;     /* d */
  int32_t depth;         /* transformer blocks */
  int32_t n_heads;
  int32_t ff_mult;
  int32_t n_freq;        /* Fourier harmonics per axis */
  int32_t n_dets;        /* K generalized spinor determinants */
  int32_t learnable_env; /* adds the complex leaf dW (N, n_g, 2) */
  int32_t use_jastrow;
  int32_t jastrow_kind;  /* 0 = rbf, 1 = cusp (both on the minimum-image distance, as the reference);
                          2 = cusp on the smooth periodic chord distance (rectangular cells; no
                          reference counterpart) */
  int32_t n_knots;
  int32_t precision;     /* NQS_FP64 or NQS_MIXED */
  int32_t gemm;          /* NQS_GEMM_* for NQS_MIXED */
  double cell[4];        /* A row-major; columns are the lattice vectors a1, a2 */
  double cell_inv[4];    /* A^-1 row-major, as the caller computes it (positions are wrapped
                            with exactly these numbers) */
  double cusp;
} nqs_model_desc;

/* Single-particle Hamiltonian of make_local_energy (src/nqs/local_energy.py:40-44).
 * The external field ext_field(r), coupled as -sum_i f(r_i).sigma_i (local_energy.py:84-111), is a
 * Fourier series m(r) = sum_t c_t cos(q_t.r) + s_t sin(q_t.r), term t = 8 doubles
 * (qx, qy, cx, cy, cz, sx, sy, sz) in device memory (a constant field is a term with q = 0);
 * f = m when field_scale = 0, and f = field_scale m / (|m| + 1e-12) when field_scale > 0 (the
 * normalised textures of src/nqs/pinning.py:20-100).  n_field = 0 means ext_field = None. */
typedef struct nqs_hamiltonian_desc {
  double m_star;
  double lambda_r;
  double h_z;
  int32_t n_field;
  int32_t n_pin;          /* Gaussian pinning centres (0 = none) */
  const double* field_dev;
  /* Gaussian pinning centres (no reference counterpart): centre p has the periodic profile
   * g_p(r) = sum_{n in {-1,0,1}^2} exp(-|u_p - A n|^2 / (2 width_p^2)), u_p = the minimum image
   * of r - c_p (fractional coordinates rounded to the nearest integer).  It adds the scalar
   * potential v0_p g_p(r_i) and the spin field h_p g_p(r_i) (coupled like ext_field) for every
   * electron.  pin_dev (n_pin, 8) = (cx, cy, width, v0, hx, hy, hz, 0) in device memory. */
  const double* pin_dev;
  double cell[4];         /* A row-major, for the pinning images (ignored when n_pin = 0) */
  double field_scale;     /* 0: f = m; > 0: f = field_scale m / (|m| + 1e-12) */
} nqs_hamiltonian_desc;

/* Precomputed Ewald tables of make_ewald_lattice (src/nqs/ewald.py:26-58). */
typedef struct nqs_ewald_desc {
  double alpha;
  double eta;
  double v_g0;
  double zeta_m;
  int32_t n_img;
  int32_t n_g;
  const double* img_dev; /* (n_img, 2) lattice images A n */
  const double* g_dev;   /* (n_g, 2) reciprocal vectors, G = 0 excluded */
  const double* rg_dev;  /* (n_g,) (2 pi / area) erfc(|G|/2 eta)/|G| */
} nqs_ewald_desc;

int nqs_version(void);
const char* nqs_last_error(void);
/* sizeof of the descriptor structs, for bindings to check their layout: which = 0 model,
 * 1 hamiltonian, 2 ewald, 3 minsr; -1 for another value.  (No reference counterpart: the
 * reference passes Python objects, not structs.) */
int64_t nqs_desc_size(int32_t which);

/* ---- parameters: flat float64 vector theta in real_ravel order (minsr.py:30-53) */
int64_t nqs_param_count(const nqs_model_desc* m);
/* Leaf idx of the flattened tree; returns NQS_ERR_ARG past the last leaf. */
int nqs_param_leaf(const nqs_model_desc* m, int32_t idx, char* name, int32_t name_len,
                   int64_t shape[3], int32_t* ndim, int64_t* offset, int32_t* is_complex);
int64_t nqs_packed_bytes(const nqs_model_desc* m);
/* theta -> kernel layout (transposed weights in the backbone precision, float64 envelope).
 * gs_dev (n_g, 2) float64; wenv_dev (N, n_g, 2) complex interleaved. */
int nqs_pack_params(const nqs_model_desc* m, const double* theta_dev, const double* gs_dev,
                    const double* wenv_dev, void* packed_dev, void* stream);

int64_t nqs_workspace_bytes(const nqs_model_desc* m, int32_t op, int64_t n_walkers);

/* ---- log psi apply: logpsi_dev (W) complex */
int nqs_psiformer_fwd(const nqs_model_desc* m, const void* packed_dev, const double* R_dev,
                      const int32_t* S_dev, int64_t n_walkers, double* logpsi_dev, void* ws_dev,
                      int64_t ws_bytes, void* stream);

/* ---- log psi apply that allocates nothing and takes no workspace.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and result as
 * nqs_psiformer_fwd without the workspace: the activations live in a 64 MiB scratch array inside
 * the library (one per GPU, held until the process ends), and walkers are processed in chunks that
 * fit it.  Returns NQS_ERR_WORKSPACE when one walker's activations exceed 64 MiB.  Calls that
 * overlap on one GPU (other streams or threads) share the scratch array and race. */
int nqs_psiformer_fwd_noalloc(const nqs_model_desc* m, const void* packed_dev, const double* R_dev,
                              const int32_t* S_dev, int64_t n_walkers, double* logpsi_dev, void* stream);

/* ---- local energy (kinetic by forward Laplacian + Rashba SOC + Zeeman + ext field).
 * eloc_dev (W) complex; logpsi_dev (W) complex or NULL; grad_dev (W, 2N) complex or NULL;
 * lap_dev (W) complex or NULL (the Laplacian of log psi). */
int nqs_eloc(const nqs_model_desc* m, const nqs_hamiltonian_desc* h, const void* packed_dev,
             const double* R_dev, const int32_t* S_dev, int64_t n_walkers, double* eloc_dev,
             double* logpsi_dev, double* grad_dev, double* lap_dev, void* ws_dev, int64_t ws_bytes,
             void* stream);

/* ---- per-walker gradient of log psi with respect to theta (the O matrix of MinSR)
 * Hand-written backward: a forward pass keeping only the block inputs, then the reverse sweep
 * recomputing each block's attention/MLP activations.  O_dev is (2W, ldo) with ldo >= P, in
 * float64 for NQS_FP64 models and float32 for NQS_MIXED: row w = Re O_w, row W + w = Im O_w,
 * column k = real_ravel index (complex leaves: re part then im part).  Columns >= P are not
 * written.  logpsi_dev (W) complex or NULL. */
int nqs_psiformer_grad(const nqs_model_desc* m, const void* packed_dev, const double* R_dev,
                       const int32_t* S_dev, int64_t n_walkers, void* O_dev, int64_t ldo,
                       double* logpsi_dev, void* ws_dev, int64_t ws_bytes, void* stream);

/* ---- vector-Jacobian product of log psi with respect to theta, summed over the walkers
 * (jax.vjp of the batched log psi; the backward of the differentiable log psi that stands in for
 * jax.grad in src/nqs/minsr.py:56-65 and src/nqs/train.py:48-58):
 *   g_dev (P) = sum_w [cot_w.re d Re log psi_w / d theta + cot_w.im d Im log psi_w / d theta]
 * cot_dev (W) complex interleaved.  The reverse sweep of nqs_psiformer_grad with the cotangent
 * folded in at the determinant head; weight, bias and LayerNorm gradients are summed over all
 * walkers and tokens in slices, the slices added in a fixed order (no O rows, no atomics).
 * logpsi_dev (W) complex or NULL.  Workspace: nqs_workspace_bytes(m, NQS_OP_VJP, W). */
int nqs_psiformer_vjp(const nqs_model_desc* m, const void* packed_dev, const double* R_dev,
                      const int32_t* S_dev, int64_t n_walkers, const double* cot_dev, double* g_dev,
                      double* logpsi_dev, void* ws_dev, int64_t ws_bytes, void* stream);

/* ---- Ewald Coulomb energy per walker: energy_dev (W) float64 */
int nqs_ewald(const nqs_ewald_desc* e, const double* R_dev, int64_t n_walkers, int32_t n_elec,
              double* energy_dev, void* stream);

/* ---- Metropolis sweeps over positions and spins with caller-supplied streams
 * xi (n_sweeps, W, N, 2) f64, u_pos (n_sweeps, W) f64, flip_i (n_sweeps, n_flip, W) i32,
 * u_flip (n_sweeps, n_flip, W) f64.  R, S, logp (W) = 2 Re log psi of the current state, are
 * updated in place; acc_pos (W), acc_spin (W) count accepted moves (int32, accumulated).  A flip
 * index outside [0, N) is a rejected proposal (no spin changes, no count). */
int nqs_mcmc_sweeps(const nqs_model_desc* m, const void* packed_dev, double* R_dev, int32_t* S_dev,
                    double* logp_dev, int64_t n_walkers, int32_t n_sweeps, int32_t n_flip,
                    double step, const double* xi_dev, const double* u_pos_dev,
                    const int32_t* flip_i_dev, const double* u_flip_dev, int32_t* acc_pos_dev,
                    int32_t* acc_spin_dev, void* ws_dev, int64_t ws_bytes, void* stream);

/* ---- MinSR natural-gradient direction (src/nqs/minsr.py:68-81)
 *   Otil = [Re dO; Im dO] (columns centred over walkers), etil = [Re ec; Im ec],
 *   dtheta = Otil^T (Otil Otil^T + lam I)^-1 etil
 * The Gram runs on MFMA in the chosen precision; the shifted Gram is factored by a float64
 * Cholesky; `refine` steps of iterative refinement use the matrix-free residual
 * etil - Otil (Otil^T y) - lam y from the float32/float64 O; the push-through reads O. */
enum { NQS_GRAM_FP64 = 0, NQS_GRAM_FP32 = 1, NQS_GRAM_BF16 = 2 };
typedef struct nqs_minsr_desc {
  int64_t n_walkers; /* W (2W must be a multiple of 128 for the bf16 Gram, of 64 otherwise) */
  int64_t n_params;  /* P */
  int64_t ldo;       /* row stride of O, >= P, multiple of 64; columns >= P must hold zeros */
  int32_t o_dtype;   /* NQS_FP64: O is float64; NQS_MIXED: O is float32 */
  int32_t gram;      /* NQS_GRAM_* (FP64 needs a float64 O) */
  int32_t refine;    /* iterative refinement steps */
  int32_t reserved;
  double lam;        /* diagonal shift added to the unnormalised Gram */
} nqs_minsr_desc;
int64_t nqs_minsr_workspace_bytes(const nqs_minsr_desc* d);
/* O_dev (2W, ldo): with the bf16 Gram it is read only (the centring is applied on the fly); with
 * the float64 / float32 Gram it is centred in place.  Pad columns P..ldo-1 are ignored.
 * eps_dev (W) complex local energies (already clipped);
 * dtheta_dev (P) float64; info_dev (int32, device): 0, or 1 + the first non-positive pivot. */
int nqs_minsr_direction(const nqs_minsr_desc* d, void* O_dev, const double* eps_dev, double* dtheta_dev,
                        int32_t* info_dev, void* ws_dev, int64_t ws_bytes, void* stream);
/* The same direction in three calls on one workspace (no reference counterpart: they let a caller
 * overlap the factorisation with the local energies, which the Gram does not need).
 * nqs_minsr_gram: centres O (in place for the float64 / float32 Gram, as nqs_minsr_direction does) and
 * forms the shifted, deflated Gram; nqs_minsr_factor: its Cholesky factor, info_dev as in
 * nqs_minsr_direction; nqs_minsr_solve: the solve for eps with refinement and the push-through into
 * dtheta.  The calls must come in this order with the same descriptor, O and workspace, each after
 * the previous one has completed on the device (same stream, or ordered by events); the workspace must
 * not be used otherwise in between.  gram + factor + solve computes exactly nqs_minsr_direction. */
int nqs_minsr_gram(const nqs_minsr_desc* d, void* O_dev, void* ws_dev, int64_t ws_bytes, void* stream);
int nqs_minsr_factor(const nqs_minsr_desc* d, int32_t* info_dev, void* ws_dev, int64_t ws_bytes, void* stream);
int nqs_minsr_solve(const nqs_minsr_desc* d, void* O_dev, const double* eps_dev, double* dtheta_dev, void* ws_dev,
                    int64_t ws_bytes, void* stream);
/* Stage timing of the MinSR calls (for measurement, no reference counterpart): enable >= 0
 * switches hipEvent recording on (1) or off (0) for later calls; ms_out (5 floats, host) receives
 * the last profiled direction's centre, Gram, Cholesky, solve+refinement, push-through times (each
 * stage timed on the stream it ran on). */
int nqs_minsr_profile(int32_t enable, float* ms_out);
/* MinSR directions of n_sys independent systems with the same walker count (the batched uneven
 * ensemble; no reference counterpart: the reference solves one system per call).  descs (n_sys)
 * must share n_walkers, gram, o_dtype and refine; n_params, ldo and lam may differ.  O_dev,
 * eps_dev, dtheta_dev are host arrays of n_sys device pointers, info_dev (n_sys) device int32.
 * The bf16 Grams of all systems are one launch (2W a multiple of 256); the Cholesky
 * factorizations and the triangular solves of all systems share each launch.  Per system the
 * result equals nqs_minsr_direction's. */
int64_t nqs_minsr_batched_workspace_bytes(const nqs_minsr_desc* descs, int32_t n_sys);
int nqs_minsr_direction_batched(const nqs_minsr_desc* descs, int32_t n_sys, void* const* O_dev,
                                const double* const* eps_dev, double* const* dtheta_dev, int32_t* info_dev,
                                void* ws_dev, int64_t ws_bytes, void* stream);
/* Launch-shape autotuner (no reference counterpart; the reference's XLA compiler has no such
 * knob): the first call of a tuned kernel (attention jet: waves per workgroup; determinant head:
 * workgroup size) for a shape class times each candidate once on the caller's stream and keeps the
 * fastest for the rest of the process.  Tuning synchronises the stream once per shape class and
 * is skipped while the stream is captured into a graph.  enable = 0 uses the default shapes;
 * reset forgets the choices; report writes one line per tuned shape class into buf (len bytes,
 * NUL-terminated) and returns the full length. */
int nqs_autotune_enable(int32_t enable);
int nqs_autotune_reset(void);
int64_t nqs_autotune_report(char* buf, int64_t len);

/* ---- first-order training (src/nqs/train.py:28-65)
 * Energy gradient of the surrogate loss 2 mean_w Re[conj(e_w - <e>) log psi(w)] (train.py:48-58):
 * g_dev (P) = (2/W) sum_w [Re(e_w - <e>) Re O_w + Im(e_w - <e>) Im O_w], from the O rows of
 * nqs_psiformer_grad (desc: n_walkers, n_params, ldo, o_dtype; the other fields are ignored) and the
 * local energies e_dev (W complex).  O is read, not modified. */
int64_t nqs_vmc_gradient_workspace_bytes(const nqs_minsr_desc* d);
int nqs_vmc_gradient(const nqs_minsr_desc* d, const void* O_dev, const double* e_dev, double* g_dev, void* ws_dev,
                     int64_t ws_bytes, void* stream);
/* Adam step of optax.adam(lr, b1, b2, eps, eps_root) (optax scale_by_adam, then -lr) on the flat
 * real ravel theta (P), in place with its moments mu, nu (P each, zero before the first step);
 * count = 1-based step number.  partner_dev (P, int64, NULL = all real): for the real part of a
 * complex leaf element, the index of its imaginary part; -2 at that imaginary part; -1 for real
 * parameters.  A complex element is one Adam element (complex first moment, second moment |g|^2
 * shared by both parts); complex_mode 0 feeds Adam g_re - i g_im (what jax.grad of a real loss
 * gives for a complex parameter, as the reference's optax step receives it), 1 feeds
 * g_re + i g_im (a descent step in both parts). */
int nqs_adam_update(double* theta_dev, const double* g_dev, double* mu_dev, double* nu_dev, const int64_t* partner_dev,
                    int64_t n_params, int64_t count, double lr, double b1, double b2, double eps, double eps_root,
                    int32_t complex_mode, void* stream);

/* ---- spin observables (src/nqs/observables.py:31-128)
 * Single-particle spin estimators of make_spin_estimators (observables.py:31-47) for every
 * electron of every walker: ratio_i = exp(log psi(S with s_i flipped) - log psi(S)) (non-finite
 * -> 0), sigma_x = ratio_i, sigma_y = -i (1 - 2 s_i) ratio_i, sigma_z = 1 - 2 s_i.  spins_dev:
 * (W, N, 3) complex (x, y, z).  Workspace: nqs_workspace_bytes(m, NQS_OP_SPIN, W). */
int nqs_spin_estimators(const nqs_model_desc* m, const void* packed_dev, const double* R_dev, const int32_t* S_dev,
                        int64_t n_walkers, double* spins_dev, void* ws_dev, int64_t ws_bytes, void* stream);
/* Spin Fourier sums over walkers (structure_factor, spin_order_parameter: observables.py:94-124)
 * from real spins (W, N, 3): per wave vector q (n_q, 2), out (n_q, 8) = m_x(q), m_y(q), m_z(q)
 * (complex, <sum_i sigma_a^i exp(-i q.r_i)>/N), S_zz(q), S_perp(q) (<|.|^2>/N). */
int nqs_spin_fourier(const double* R_dev, const double* spins_dev, int64_t n_walkers, int32_t n_elec,
                     const double* q_dev, int32_t n_q, double* out_dev, void* stream);
/* Texture sums of spin_texture (observables.py:56-76): bin (ix, iy) of each electron from its
 * fractional coordinates mod 1 (cell_inv = A^-1 row-major, host memory), m_dev (ngrid, ngrid, 3)
 * = sum of the real spins in the bin, rho_dev (ngrid, ngrid) = count (not yet divided by W). */
int nqs_spin_texture(const double* R_dev, const double* spins_dev, int64_t n_walkers, int32_t n_elec,
                     const double cell_inv[4], int32_t ngrid, double* m_dev, double* rho_dev, void* stream);

/* ---- ratio estimators (mc_overlap, src/nqs/manybody_chern.py:23-34; rotation_eigenvalue_vmc,
 * src/nqs/topology.py:114-122): r_w = phase_w exp(lb_w - la_w - shift) from two log psi arrays
 * (W complex) and an optional phase (W complex, NULL = 1); shift = max_w Re(lb - la) when
 * shift != 0, else 0.  out_dev[6] = mean Re r, mean Im r, mean |r|^2, effective-sample fraction
 * (sum|r|)^2 / sum|r|^2 / W, shift, std r (sqrt(mean |r - mean r|^2)). */
int nqs_ratio_stats(const double* la_dev, const double* lb_dev, const double* phase_dev, int64_t n_walkers,
                    int32_t shift, double* out_dev, void* stream);

/* Median/MAD clip of Re e (minsr.py:145-148), W <= 16384: e_clip_dev (W) complex (Im kept);
 * stats_dev[4] = mean Re e (unclipped), std(Re e)/sqrt(W), median, MAD. */
int nqs_clip_energies(const double* e_dev, int64_t n_walkers, double clip_k, double* e_clip_dev,
                      double* stats_dev, void* stream);

#ifdef __cplusplus
}
#endif
#endif
