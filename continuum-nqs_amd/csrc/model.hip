// Model plan: the real_ravel parameter layout (src/nqs/minsr.py:30-53 over the flax tree of
// src/nqs/psiformer.py:50-188), the packed kernel layout, the parameter packer and the jet
// forward driver that chains the backbone kernels.
#include <algorithm>
#include <cstring>
#include "model.h"
#include "bwd.h"

namespace nqs {

static void add_leaf(std::vector<std::pair<std::vector<std::string>, Leaf>>& v, std::vector<std::string> path,
                     std::initializer_list<int64_t> shape, bool cplx) {
  Leaf l;
  l.name.clear();
  for (size_t i = 0; i < path.size(); ++i) l.name += (i ? "/" : "") + path[i];
  l.ndim = (int)shape.size();
  int64_t sz = 1;
  int q = 0;
  for (int64_t s : shape) {
    l.shape[q++] = s;
    sz *= s;
  }
  for (; q < 3; ++q) l.shape[q] = 1;
  l.size = sz;
  l.cplx = cplx;
  l.offset = 0;
  v.push_back({path, l});
}

int build_model(const nqs_model_desc* m, Model& md) {
  NQS_REQUIRE(m != nullptr, NQS_ERR_ARG, "null model descriptor");
  md.m = *m;
  md.N = m->n_elec;
  md.d = m->embed_dim;
  md.depth = m->depth;
  md.nh = m->n_heads;
  md.ff = m->ff_mult;
  md.F = m->n_freq;
  md.K = m->n_dets;
  md.n_g = m->n_g;
  NQS_REQUIRE(md.N >= 1 && md.N <= 32, NQS_ERR_UNSUPPORTED, "n_elec must be in [1, 32]");
  NQS_REQUIRE(md.d >= 1 && md.nh >= 1 && md.d % md.nh == 0, NQS_ERR_ARG, "embed_dim must be a multiple of n_heads");
  NQS_REQUIRE(md.d <= 512 && md.ff * md.d <= 512, NQS_ERR_UNSUPPORTED, "embed_dim*ff_mult must be <= 512");
  NQS_REQUIRE(md.K >= 1 && md.n_g >= 1 && md.F >= 1 && md.depth >= 0 && md.ff >= 1, NQS_ERR_ARG, "bad architecture");
  NQS_REQUIRE(m->precision == NQS_FP64 || m->precision == NQS_MIXED, NQS_ERR_ARG, "bad precision");
  NQS_REQUIRE(m->precision == NQS_FP64 || (m->gemm >= NQS_GEMM_BF16X3 && m->gemm <= NQS_GEMM_BF16), NQS_ERR_ARG,
              "bad gemm mode");
  md.split = m->precision == NQS_FP64 ? 0 : (m->gemm == NQS_GEMM_BF16X3 ? 3 : m->gemm == NQS_GEMM_BF16 ? 1 : 0);
  md.kin = 4 * md.F + 2;
  md.kpad = (int)round_up(md.kin, 8);
  md.KN = md.K * md.N;
  for (int i = 0; i < 4; ++i) md.A[i] = m->cell[i];
  const double det = md.A[0] * md.A[3] - md.A[1] * md.A[2];
  NQS_REQUIRE(det != 0.0, NQS_ERR_ARG, "singular cell");
  for (int i = 0; i < 4; ++i) md.Ainv[i] = m->cell_inv[i];
  const double e00 = md.A[0] * md.Ainv[0] + md.A[1] * md.Ainv[2] - 1.0, e01 = md.A[0] * md.Ainv[1] + md.A[1] * md.Ainv[3];
  const double e10 = md.A[2] * md.Ainv[0] + md.A[3] * md.Ainv[2], e11 = md.A[2] * md.Ainv[1] + md.A[3] * md.Ainv[3] - 1.0;
  NQS_REQUIRE(fabs(e00) + fabs(e01) + fabs(e10) + fabs(e11) < 1e-10, NQS_ERR_ARG, "cell_inv is not the inverse of cell");
  NQS_REQUIRE(!m->use_jastrow || (m->jastrow_kind >= 0 && m->jastrow_kind <= 2), NQS_ERR_ARG, "bad jastrow_kind");
  NQS_REQUIRE(!(m->use_jastrow && m->jastrow_kind == 2) || (md.A[1] == 0.0 && md.A[2] == 0.0), NQS_ERR_UNSUPPORTED,
              "the periodic cusp Jastrow needs a rectangular cell");

  // ---- leaves of the flax tree (psiformer.py:50-188)
  const int d = md.d, ffd = md.ff * md.d, N = md.N, K = md.K;
  std::vector<std::pair<std::vector<std::string>, Leaf>> v;
  auto dense = [&](std::vector<std::string> p, int64_t in, int64_t out) {
    auto pb = p, pk = p;
    pb.push_back("bias");
    pk.push_back("kernel");
    add_leaf(v, pb, {out}, false);
    add_leaf(v, pk, {in, out}, false);
  };
  auto ln = [&](std::vector<std::string> p) {
    auto pb = p, ps = p;
    pb.push_back("bias");
    ps.push_back("scale");
    add_leaf(v, pb, {d}, false);
    add_leaf(v, ps, {d}, false);
  };
  ln({"params", "LayerNorm_0"});
  dense({"params", "embed"}, md.kin, d);
  add_leaf(v, {"params", "backflow", "bias"}, {md.KN}, true);
  add_leaf(v, {"params", "backflow", "kernel"}, {d, md.KN}, true);
  for (int l = 0; l < md.depth; ++l) {
    const std::string b = "TransformerBlock_" + std::to_string(l);
    dense({"params", b, "Dense_0"}, d, ffd);
    dense({"params", b, "Dense_1"}, ffd, d);
    ln({"params", b, "LayerNorm_0"});
    ln({"params", b, "LayerNorm_1"});
    for (const char* q : {"q", "k", "v", "o"}) dense({"params", b, "attn", q}, d, d);
  }
  if (m->learnable_env) add_leaf(v, {"params", "dW"}, {N, md.n_g, 2}, true);
  if (K > 1) add_leaf(v, {"params", "det_logw"}, {K}, true);
  if (m->use_jastrow) {
    if (m->jastrow_kind == 0) {
      add_leaf(v, {"params", "jas_c_same"}, {m->n_knots}, false);
      add_leaf(v, {"params", "jas_c_opp"}, {m->n_knots}, false);
    } else {
      add_leaf(v, {"params", "jastrow_loga"}, {}, false);
    }
  }
  std::stable_sort(v.begin(), v.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  md.leaves.clear();
  int64_t off = 0;
  for (auto& e : v) {
    e.second.offset = off;
    off += e.second.size * (e.second.cplx ? 2 : 1);
    md.leaves.push_back(e.second);
  }
  md.P = off;

  // ---- packed layout (T elements)
  int64_t o = 0;
  auto take = [&](int64_t n) {
    const int64_t r = o;
    o += round_up(n, 64);
    return r;
  };
  md.embT = take((int64_t)d * md.kpad);
  md.emb_b = take(d);
  md.emb = take((int64_t)md.kpad * d);
  md.layers.resize(md.depth);
  for (int l = 0; l < md.depth; ++l) {
    LayerPack& L = md.layers[l];
    L.ln0_s = take(d);
    L.ln0_b = take(d);
    L.qkvT = take(3LL * d * d);
    L.qkv_b = take(3LL * d);
    L.oT = take((int64_t)d * d);
    L.o_b = take(d);
    L.ln1_s = take(d);
    L.ln1_b = take(d);
    L.w0T = take((int64_t)ffd * d);
    L.b0 = take(ffd);
    L.w1T = take((int64_t)d * ffd);
    L.b1 = take(d);
    L.qkv = take(3LL * d * d);
    L.o = take((int64_t)d * d);
    L.w0 = take((int64_t)d * ffd);
    L.w1 = take((int64_t)ffd * d);
  }
  md.lnf_s = take(d);
  md.lnf_b = take(d);
  md.bfT = take(2LL * md.KN * d);
  md.bf_b = take(2LL * md.KN);
  md.bfW = take((int64_t)d * 2 * md.KN);
  md.fused = fused_model_ok(md);
  if (md.fused) {
    md.fz_emb = take((int64_t)d * 64);
    for (int l = 0; l < md.depth; ++l) {
      LayerPack& L = md.layers[l];
      L.fz_qkv = take(3LL * d * d);
      L.fz_o = take((int64_t)d * d);
      L.fz_w0 = take((int64_t)ffd * d);
      L.fz_w1 = take((int64_t)d * ffd);
    }
    md.fz_bf = take(2LL * md.KN * d);
  }
  md.n_backbone = o;
  md.tsize = (m->precision == NQS_FP64) ? 8 : 4;
  md.dbl_off = round_up(o * md.tsize, 256);
  int64_t q = 0;
  md.gs_off = q;
  q += round_up(2LL * md.n_g, 8);
  md.wenv_off = q;
  q += round_up(4LL * N * md.n_g, 8);
  md.dlw_off = q;
  q += round_up(2LL * K, 8);
  md.jas_off = q;
  q += round_up(2LL * std::max(1, m->n_knots), 8);
  md.n_dbl = q;
  md.packed_bytes = md.dbl_off + q * 8;
  return NQS_OK;
}

const Leaf* find_leaf(const Model& md, const std::string& name) {
  for (const Leaf& l : md.leaves)
    if (l.name == name) return &l;
  return nullptr;
}

// ---- packer kernels
template <typename T>
__global__ void k_copy2d(const double* __restrict__ src, int64_t rows, int64_t cols, int64_t src_rs, int64_t src_cs,
                         T* __restrict__ dst, int64_t dst_rs) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= rows * cols) return;
  const int64_t r = q / cols, c = q % cols;
  dst[r * dst_rs + c] = (T)src[r * src_rs + c * src_cs];
}

__global__ void k_pack_env(const double* __restrict__ wenv, const double* __restrict__ dre, const double* __restrict__ dim,
                           int64_t n, double* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  out[2 * q] = wenv[2 * q] + (dre ? dre[q] : 0.0);
  out[2 * q + 1] = wenv[2 * q + 1] + (dim ? dim[q] : 0.0);
}

template <typename T>
static int copy2d(const double* src, int64_t rows, int64_t cols, int64_t src_rs, int64_t src_cs, T* dst,
                  int64_t dst_rs, hipStream_t st) {
  const int64_t n = rows * cols;
  if (n == 0) return NQS_OK;
  hipLaunchKernelGGL(k_copy2d<T>, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, st, src, rows, cols, src_rs, src_cs,
                     dst, dst_rs);
  NQS_LAUNCH_CHECK();
  return NQS_OK;
}


template <typename T>
static int pack_backbone(const Model& md, const double* th, T* pk, hipStream_t st) {
  const int d = md.d, ffd = md.ff * md.d, KN = md.KN;
  auto L = [&](const std::string& n) -> const Leaf* { return find_leaf(md, n); };
  // Dense kernel (in, out) row-major -> transposed (out, ld) and plain (in, ld)
  auto dense_T = [&](const Leaf* k, T* dst, int64_t ld) {
    return copy2d<T>(th + k->offset, k->shape[1], k->shape[0], 1, k->shape[1], dst, ld, st);
  };
  auto dense_N = [&](const Leaf* k, T* dst, int64_t ld) {
    return copy2d<T>(th + k->offset, k->shape[0], k->shape[1], k->shape[1], 1, dst, ld, st);
  };
  auto vec = [&](const Leaf* b, T* dst) { return copy2d<T>(th + b->offset, 1, b->size, 0, 1, dst, 0, st); };
  NQS_TRY(dense_T(L("params/embed/kernel"), pk + md.embT, md.kpad));
  NQS_TRY(dense_N(L("params/embed/kernel"), pk + md.emb, d));
  NQS_TRY(vec(L("params/embed/bias"), pk + md.emb_b));
  for (int l = 0; l < md.depth; ++l) {
    const std::string b = "params/TransformerBlock_" + std::to_string(l) + "/";
    const LayerPack& P = md.layers[l];
    NQS_TRY(vec(L(b + "LayerNorm_0/scale"), pk + P.ln0_s));
    NQS_TRY(vec(L(b + "LayerNorm_0/bias"), pk + P.ln0_b));
    NQS_TRY(vec(L(b + "LayerNorm_1/scale"), pk + P.ln1_s));
    NQS_TRY(vec(L(b + "LayerNorm_1/bias"), pk + P.ln1_b));
    const char* qkv[3] = {"q", "k", "v"};
    for (int s = 0; s < 3; ++s) {
      NQS_TRY(dense_T(L(b + "attn/" + qkv[s] + "/kernel"), pk + P.qkvT + (int64_t)s * d * d, d));
      NQS_TRY(dense_N(L(b + "attn/" + qkv[s] + "/kernel"), pk + P.qkv + s * d, 3 * d));
      NQS_TRY(vec(L(b + "attn/" + qkv[s] + "/bias"), pk + P.qkv_b + s * d));
    }
    NQS_TRY(dense_T(L(b + "attn/o/kernel"), pk + P.oT, d));
    NQS_TRY(dense_N(L(b + "attn/o/kernel"), pk + P.o, d));
    NQS_TRY(vec(L(b + "attn/o/bias"), pk + P.o_b));
    NQS_TRY(dense_T(L(b + "Dense_0/kernel"), pk + P.w0T, d));
    NQS_TRY(dense_N(L(b + "Dense_0/kernel"), pk + P.w0, ffd));
    NQS_TRY(vec(L(b + "Dense_0/bias"), pk + P.b0));
    NQS_TRY(dense_T(L(b + "Dense_1/kernel"), pk + P.w1T, ffd));
    NQS_TRY(dense_N(L(b + "Dense_1/kernel"), pk + P.w1, d));
    NQS_TRY(vec(L(b + "Dense_1/bias"), pk + P.b1));
  }
  NQS_TRY(vec(L("params/LayerNorm_0/scale"), pk + md.lnf_s));
  NQS_TRY(vec(L("params/LayerNorm_0/bias"), pk + md.lnf_b));
  const Leaf* bk = L("params/backflow/kernel");
  const Leaf* bb = L("params/backflow/bias");
  // backflow kernel (d, KN) complex: re block then im block
  NQS_TRY(copy2d<T>(th + bk->offset, KN, d, 1, KN, pk + md.bfT, d, st));
  NQS_TRY(copy2d<T>(th + bk->offset + bk->size, KN, d, 1, KN, pk + md.bfT + (int64_t)KN * d, d, st));
  NQS_TRY(copy2d<T>(th + bk->offset, d, KN, KN, 1, pk + md.bfW, 2 * KN, st));
  NQS_TRY(copy2d<T>(th + bk->offset + bk->size, d, KN, KN, 1, pk + md.bfW + KN, 2 * KN, st));
  NQS_TRY(copy2d<T>(th + bb->offset, 1, 2 * KN, 0, 1, pk + md.bf_b, 0, st));
  return NQS_OK;
}

int pack_params(const Model& md, const double* theta, const double* gs, const double* wenv, void* packed,
                hipStream_t st) {
  NQS_CHECK_HIP(hipMemsetAsync(packed, 0, md.packed_bytes, st));
  if (md.tsize == 8)
    NQS_TRY(pack_backbone<double>(md, theta, (double*)packed, st));
  else
    NQS_TRY(pack_backbone<float>(md, theta, (float*)packed, st));
  if (md.fused) NQS_TRY(pack_fused_weights(md, (float*)packed, st));
  double* dp = (double*)((char*)packed + md.dbl_off);
  NQS_CHECK_HIP(hipMemcpyAsync(dp + md.gs_off, gs, sizeof(double) * 2 * md.n_g, hipMemcpyDeviceToDevice, st));
  const Leaf* dw = find_leaf(md, "params/dW");
  const int64_t nenv = (int64_t)md.N * md.n_g * 2;
  hipLaunchKernelGGL(k_pack_env, dim3((unsigned)cdiv(nenv, 256)), dim3(256), 0, st, wenv,
                     dw ? theta + dw->offset : (const double*)nullptr,
                     dw ? theta + dw->offset + dw->size : (const double*)nullptr, nenv, dp + md.wenv_off);
  NQS_LAUNCH_CHECK();
  const Leaf* lw = find_leaf(md, "params/det_logw");
  if (lw) {
    NQS_TRY(copy2d<double>(theta + lw->offset, md.K, 1, 1, 0, dp + md.dlw_off, 2, st));
    NQS_TRY(copy2d<double>(theta + lw->offset + lw->size, md.K, 1, 1, 0, dp + md.dlw_off + 1, 2, st));
  }
  if (md.m.use_jastrow) {
    if (md.m.jastrow_kind == 0) {
      const Leaf* cs = find_leaf(md, "params/jas_c_same");
      const Leaf* co = find_leaf(md, "params/jas_c_opp");
      NQS_TRY(copy2d<double>(theta + cs->offset, 1, cs->size, 0, 1, dp + md.jas_off, 0, st));
      NQS_TRY(copy2d<double>(theta + co->offset, 1, co->size, 0, 1, dp + md.jas_off + md.m.n_knots, 0, st));
    } else {
      const Leaf* la = find_leaf(md, "params/jastrow_loga");
      NQS_TRY(copy2d<double>(theta + la->offset, 1, 1, 0, 1, dp + md.jas_off, 0, st));
    }
  }
  return NQS_OK;
}

DetParams det_params(const Model& md, const void* packed) {
  DetParams p;
  p.N = md.N;
  p.n_g = md.n_g;
  p.K = md.K;
  p.use_jastrow = md.m.use_jastrow;
  p.jastrow_kind = md.m.jastrow_kind;
  p.n_knots = md.m.n_knots;
  for (int i = 0; i < 4; ++i) {
    p.A[i] = md.A[i];
    p.Ainv[i] = md.Ainv[i];
  }
  p.cusp = md.m.cusp;
  const double area = fabs(md.A[0] * md.A[3] - md.A[1] * md.A[2]);
  p.rbf_rcut = 0.5 * sqrt(area);
  p.rbf_w = p.rbf_rcut / (double)std::max(md.m.n_knots - 1, 1);
  const double* dp = (const double*)((const char*)packed + md.dbl_off);
  p.gs = dp + md.gs_off;
  p.wenv = dp + md.wenv_off;
  p.det_logw = dp + md.dlw_off;
  p.jas = dp + md.jas_off;
  return p;
}

// ---- workspace of one jet forward
static int64_t al(int64_t b) { return round_up(b, 256); }

int64_t fwd_bytes(const Model& md, int64_t n_cfg, int C, int nt) {
  const int64_t rows = n_cfg * C * md.N;
  const int64_t wbig = std::max<int64_t>({3LL * md.d, (int64_t)md.ff * md.d, 2LL * md.KN});
  const int64_t watt = std::max<int64_t>(md.d, (int64_t)md.ff * md.d);
  const int64_t t = md.tsize;
  return al(rows * md.kpad * t) + 2 * al(rows * md.d * t) + al(rows * wbig * t) + al(rows * watt * t) +
         al(n_cfg * 16) + al(n_cfg * (int64_t)std::max(nt, 1) * 16) + al(n_cfg * 16);
}

template <typename T>
FwdBufs<T> carve_fwd(const Model& md, char*& p, int64_t n_cfg, int C, int nt) {
  const int64_t rows = n_cfg * C * md.N;
  const int64_t wbig = std::max<int64_t>({3LL * md.d, (int64_t)md.ff * md.d, 2LL * md.KN});
  const int64_t watt = std::max<int64_t>(md.d, (int64_t)md.ff * md.d);
  FwdBufs<T> b;
  auto grab = [&](int64_t bytes) {
    char* r = p;
    p += al(bytes);
    return r;
  };
  b.X0 = (T*)grab(rows * md.kpad * sizeof(T));
  b.x = (T*)grab(rows * md.d * sizeof(T));
  b.h = (T*)grab(rows * md.d * sizeof(T));
  b.big = (T*)grab(rows * wbig * sizeof(T));
  b.att = (T*)grab(rows * watt * sizeof(T));
  b.lp = (double*)grab(n_cfg * 16);
  b.tan = (double*)grab(n_cfg * (int64_t)std::max(nt, 1) * 16);
  b.lap = (double*)grab(n_cfg * 16);
  return b;
}

template <typename T>
static GemmArgs<T> gargs(const Model& md, const T* A, const T* B, T* C, int64_t M, int64_t N, int64_t K, int64_t lda,
                         int64_t ldb, int64_t ldc, const T* bias, int64_t period, int64_t blen, T beta) {
  GemmArgs<T> g;
  g.A = A;
  g.B = B;
  g.C = C;
  g.M = M;
  g.N = N;
  g.K = K;
  g.lda = lda;
  g.ldb = ldb;
  g.ldc = ldc;
  g.strideA = g.strideB = g.strideC = 0;
  g.alpha = (T)1;
  g.beta = beta;
  g.bias = bias;
  g.bias_period = period;
  g.bias_len = blen;
  g.tri = 0;
  g.mirror = 0;
  g.split = sizeof(T) == 4 ? md.split : 0;
  return g;
}

template <typename T>
int run_forward(const Model& md, const void* packed, const JetGeom& jg, const double* R, const int32_t* S, int64_t w0,
                const FwdBufs<T>& b, hipStream_t st) {
  const T* pk = (const T*)packed;
  const int d = md.d, ffd = md.ff * md.d, KN = md.KN;
  const int64_t rows = jg.n_cfg * jg.C * md.N;
  const int64_t per = (int64_t)jg.C * md.N, blen = md.N;
  NQS_TRY(launch_features<T>(jg, R, S, md.Ainv, md.F, md.kpad, w0, b.X0, st));
  if constexpr (sizeof(T) == 4) {
    // value-only jets of the headline architecture: the whole backbone in one launch
    const bool fv = fused_value_ok(md, jg), fj = !fv && fused_jet1_ok(md, jg);
    if (fv || fj) {
      if (fv)
        NQS_TRY(launch_fused_value(md, pk, b.X0, rows, b.big, st));
      else
        NQS_TRY(launch_fused_jet1(md, pk, b.X0, jg.n_cfg, b.big, st));
      const DetParams dp = det_params(md, packed);
      const int64_t watt = std::max<int64_t>(md.d, (int64_t)md.ff * md.d);
      return launch_det_jet<T>(jg, dp, R, S, w0, b.big, b.lp, b.tan, b.lap, b.att, (size_t)(rows * watt * sizeof(T)),
                               st);
    }
  }
  NQS_TRY(gemm_nt<T>(gargs<T>(md, b.X0, pk + md.embT, b.x, rows, d, md.kpad, md.kpad, md.kpad, d, pk + md.emb_b, per,
                              blen, (T)0),
                     1, st));
  for (int l = 0; l < md.depth; ++l) {
    const LayerPack& P = md.layers[l];
    NQS_TRY(launch_layernorm_jet<T>(jg, d, b.x, pk + P.ln0_s, pk + P.ln0_b, b.h, st));
    NQS_TRY(gemm_nt<T>(gargs<T>(md, b.h, pk + P.qkvT, b.big, rows, 3 * d, d, d, d, 3 * d, pk + P.qkv_b, per, blen, (T)0),
                       1, st));
    NQS_TRY(launch_attention_jet<T>(jg, d, md.nh, b.big, b.att, st));
    // residual GEMMs accumulate into x: the tile-shape tuner's trials write into h (free here)
    GemmArgs<T> go = gargs<T>(md, b.att, pk + P.oT, b.x, rows, d, d, d, d, d, pk + P.o_b, per, blen, (T)1);
    go.tune_scratch = b.h;
    NQS_TRY(gemm_nt<T>(go, 1, st));
    NQS_TRY(launch_layernorm_jet<T>(jg, d, b.x, pk + P.ln1_s, pk + P.ln1_b, b.h, st));
    NQS_TRY(gemm_nt<T>(gargs<T>(md, b.h, pk + P.w0T, b.big, rows, ffd, d, d, d, ffd, pk + P.b0, per, blen, (T)0), 1, st));
    // without a Laplacian channel the GELU jet is elementwise given the token's value row, so the
    // bf16 Dense_1 applies it while staging its A operand (no pass over the hidden activations)
    const bool fuse_gelu = sizeof(T) == 4 && md.split != 0 && !jg.has_lap;
    if (!fuse_gelu) NQS_TRY(launch_gelu_jet<T>(jg, ffd, b.big, b.att, st));
    GemmArgs<T> g1 = gargs<T>(md, fuse_gelu ? b.big : b.att, pk + P.w1T, b.x, rows, d, ffd, ffd, ffd, d, pk + P.b1,
                              per, blen, (T)1);
    g1.tune_scratch = b.h;
    if (fuse_gelu) {
      g1.gelu_cn = jg.C * md.N;
      g1.gelu_n = md.N;
    }
    NQS_TRY(gemm_nt<T>(g1, 1, st));
  }
  NQS_TRY(launch_layernorm_jet<T>(jg, d, b.x, pk + md.lnf_s, pk + md.lnf_b, b.h, st));
  NQS_TRY(gemm_nt<T>(gargs<T>(md, b.h, pk + md.bfT, b.big, rows, 2 * KN, d, d, d, 2 * KN, pk + md.bf_b, per, blen, (T)0),
                     1, st));
  const DetParams dp = det_params(md, packed);
  // the attention buffer is free here: it holds the per-determinant results
  const int64_t watt = std::max<int64_t>(md.d, (int64_t)md.ff * md.d);
  NQS_TRY(launch_det_jet<T>(jg, dp, R, S, w0, b.big, b.lp, b.tan, b.lap, b.att, (size_t)(rows * watt * sizeof(T)), st));
  return NQS_OK;
}

template FwdBufs<double> carve_fwd<double>(const Model&, char*&, int64_t, int, int);
template FwdBufs<float> carve_fwd<float>(const Model&, char*&, int64_t, int, int);
template int run_forward<double>(const Model&, const void*, const JetGeom&, const double*, const int32_t*, int64_t,
                                 const FwdBufs<double>&, hipStream_t);
template int run_forward<float>(const Model&, const void*, const JetGeom&, const double*, const int32_t*, int64_t,
                                const FwdBufs<float>&, hipStream_t);


// ------------------------------------------------------------------ backward driver
int64_t grad_bytes(const Model& md, int64_t n_cfg) {
  const int64_t rows = n_cfg * md.N, t = md.tsize;
  const int64_t d = md.d, ffd = (int64_t)md.ff * md.d, KN2 = 2LL * md.KN;
  const int64_t wq = std::max<int64_t>(3 * d, KN2);
  int64_t b = al(rows * md.kpad * t) + (md.depth + 1) * al(rows * d * t);
  b += 4 * al(rows * d * t) + al(rows * 3 * d * t) + 2 * al(rows * ffd * t) + al(rows * d * t) + al(rows * KN2 * t);
  b += al(2 * rows * d * t) + al(2 * rows * ffd * t) + al(2 * rows * d * t) + al(2 * rows * wq * t);
  return b;
}

// Extra workspace of the vector-Jacobian product: the slice partials of the weight sums and the
// compact per-walker rows of the LayerNorm and determinant-head leaves.
static int64_t vjp_det_cols(const Model& md) {
  const Leaf* jl = find_leaf(md, "params/jas_c_same");
  const Leaf* jc = find_leaf(md, "params/jastrow_loga");
  return (md.K > 1 ? 2LL * md.K : 0) + (find_leaf(md, "params/dW") ? 2LL * md.N * md.n_g * 2 : 0) +
         (jl ? 2 * jl->size : 0) + (jc ? 1 : 0);
}
int64_t vjp_bytes(const Model& md, int64_t n_cfg) {
  const int64_t d = md.d, ffd = (int64_t)md.ff * md.d;
  const int64_t maxio = std::max<int64_t>({(int64_t)md.kin * d, d * 2LL * md.KN, ffd * d, 3 * d * d});
  return al(vjp_slices(n_cfg) * maxio * md.tsize) + al(2 * n_cfg * std::max<int64_t>(1, vjp_det_cols(md)) * md.tsize) +
         al(2 * n_cfg * 2 * d * md.tsize);
}

template <typename T>
int run_backward(const Model& md, const void* packed, const double* R, const int32_t* S, int64_t w0, int64_t n_cfg,
                 int64_t W_total, T* O, int64_t ldo, double* out_lp, char* ws, hipStream_t st, const double* vjp_cot,
                 double* vjp_g) {
  const bool vjp = vjp_g != nullptr;
  const T* pk = (const T*)packed;
  const int N = md.N, d = md.d, ffd = md.ff * md.d, KN = md.KN, depth = md.depth;
  const int64_t rows = n_cfg * N;
  char* p = ws;
  auto grab = [&](int64_t elems) {
    T* r = (T*)p;
    p += al(elems * (int64_t)sizeof(T));
    return r;
  };
  T* X0 = grab(rows * md.kpad);
  std::vector<T*> xs(depth + 1);
  for (int l = 0; l <= depth; ++l) xs[l] = grab(rows * d);
  T* hA = grab(rows * d);
  T* attA = grab(rows * d);
  T* xmid = grab(rows * d);
  T* hB = grab(rows * d);
  T* qkv = grab(rows * 3 * d);
  T* U = grab(rows * ffd);
  T* Gv = grab(rows * ffd);
  T* xf = grab(rows * d);
  T* bfo = grab(rows * 2 * KN);
  T* xb = grab(2 * rows * d);
  T* gb = grab(2 * rows * ffd);
  T* hb = grab(2 * rows * d);
  T* qb = grab(2 * rows * std::max(3 * d, 2 * KN));
  // vector-Jacobian product: slice partials and compact small-leaf rows after the backward buffers
  const int64_t n_slices = vjp_slices(n_cfg);
  const int64_t ld_det = std::max<int64_t>(1, vjp_det_cols(md)), ld_ln = 2LL * d;
  T* part = nullptr;
  T* osm_det = nullptr;
  T* osm_ln = nullptr;
  if (vjp) {
    const int64_t maxio = std::max<int64_t>({(int64_t)md.kin * d, (int64_t)d * 2 * KN, (int64_t)ffd * d, 3LL * d * d});
    part = grab(n_slices * maxio);
    osm_det = grab(2 * n_cfg * ld_det);
    osm_ln = grab(2 * n_cfg * ld_ln);
  }
  // the O-row kernels index rows c * W_total + w0 + cfg; a compact buffer is addressed through a
  // base shifted back by w0 rows, with W_total = n_cfg
  auto shifted = [&](T* buf, int64_t ld) { return (T*)((uintptr_t)buf - (uintptr_t)(w0 * ld * (int64_t)sizeof(T))); };
  const JetGeom jg{n_cfg, N, 1, 0, 0, 0};
  const int64_t per = N, blen = N;
  auto leaf = [&](const std::string& n) -> int64_t {
    const Leaf* l = find_leaf(md, n);
    return l ? l->offset : -1;
  };
  auto wg = [&](const T* X, int64_t ldx, int in, const T* G, int out, std::initializer_list<int64_t> bases,
                int seg_w) {
    WgradArgs<T> a;
    a.X = X;
    a.ldx = ldx;
    a.in = in;
    a.G = G;
    a.ldg = out;
    a.out = out;
    a.N = N;
    a.n_cfg = n_cfg;
    a.w0 = w0;
    a.W_total = W_total;
    a.O = O;
    a.ldo = ldo;
    int q = 0;
    for (int64_t b : bases) a.seg_base[q++] = b;
    for (; q < 3; ++q) a.seg_base[q] = 0;
    a.seg_w = seg_w;
    return a;
  };
  auto lnb = [&](const T* x, const T* ybar, T* xbar, const T* scale, int acc, const std::string& name) -> int {
    LnBwdArgs<T> a;
    a.x = x;
    a.ybar = ybar;
    a.xbar = xbar;
    a.scale = scale;
    a.d = d;
    a.N = N;
    a.accumulate = acc;
    a.n_cfg = n_cfg;
    a.w0 = w0;
    a.W_total = vjp ? n_cfg : W_total;
    a.O = vjp ? shifted(osm_ln, ld_ln) : O;
    a.ldo = vjp ? ld_ln : ldo;
    a.off_scale = vjp ? 0 : leaf(name + "/scale");
    a.off_bias = vjp ? d : leaf(name + "/bias");
    NQS_TRY(launch_ln_bwd<T>(a, st));
    if (vjp) {
      OsmMap mp;
      mp.n = 2;
      mp.start[0] = 0;
      mp.start[1] = d;
      mp.start[2] = 2 * d;
      mp.dst[0] = leaf(name + "/scale");
      mp.dst[1] = leaf(name + "/bias");
      NQS_TRY(launch_vjp_osm_reduce<T>(osm_ln, ld_ln, n_cfg, mp, nullptr, w0, vjp_g, st));
    }
    return NQS_OK;
  };
  // weight and bias gradients: O rows, or (vector-Jacobian product) sums over the rows into g
  auto wgrad = [&](const WgradArgs<T>& a, hipStream_t) -> int {
    return vjp ? launch_vjp_wsum<T>(a, part, n_slices, vjp_g, st) : launch_wgrad<T>(a, st);
  };
  auto bgrad = [&](WgradArgs<T> a, hipStream_t) -> int {
    if (!vjp) return launch_bgrad<T>(a, st);
    a.X = nullptr;
    return launch_vjp_wsum<T>(a, part, n_slices, vjp_g, st);
  };
  auto gemm = [&](const T* A, const T* B, T* C, int64_t M, int64_t Nn, int64_t K, const T* bias, T beta) {
    return gemm_nt<T>(gargs<T>(md, A, B, C, M, Nn, K, K, K, Nn, bias, per, blen, beta), 1, st);
  };
  // block forward from x_in into x_out (x_out = x_in + attn + mlp), keeping hA, qkv, attA, xmid, hB, U, Gv
  auto block_fwd = [&](int l, const T* x_in, T* x_out) -> int {
    const LayerPack& P = md.layers[l];
    NQS_TRY(launch_layernorm_jet<T>(jg, d, x_in, pk + P.ln0_s, pk + P.ln0_b, hA, st));
    NQS_TRY(gemm(hA, pk + P.qkvT, qkv, rows, 3 * d, d, pk + P.qkv_b, (T)0));
    NQS_TRY(launch_attention_jet<T>(jg, d, md.nh, qkv, attA, st));
    NQS_CHECK_HIP(hipMemcpyAsync(xmid, x_in, rows * d * sizeof(T), hipMemcpyDeviceToDevice, st));
    NQS_TRY(gemm(attA, pk + P.oT, xmid, rows, d, d, pk + P.o_b, (T)1));
    NQS_TRY(launch_layernorm_jet<T>(jg, d, xmid, pk + P.ln1_s, pk + P.ln1_b, hB, st));
    NQS_TRY(gemm(hB, pk + P.w0T, U, rows, ffd, d, pk + P.b0, (T)0));
    NQS_TRY(launch_gelu_jet<T>(jg, ffd, U, Gv, st));
    if (x_out) {
      NQS_CHECK_HIP(hipMemcpyAsync(x_out, xmid, rows * d * sizeof(T), hipMemcpyDeviceToDevice, st));
      NQS_TRY(gemm(Gv, pk + P.w1T, x_out, rows, d, ffd, pk + P.b1, (T)1));
    }
    return NQS_OK;
  };
  // ---- forward with block-input checkpoints
  NQS_TRY(launch_features<T>(jg, R, S, md.Ainv, md.F, md.kpad, w0, X0, st));
  NQS_TRY(gemm_nt<T>(gargs<T>(md, X0, pk + md.embT, xs[0], rows, d, md.kpad, md.kpad, md.kpad, d, pk + md.emb_b, per, blen,
                              (T)0),
                     1, st));
  for (int l = 0; l < depth; ++l) NQS_TRY(block_fwd(l, xs[l], xs[l + 1]));
  NQS_TRY(launch_layernorm_jet<T>(jg, d, xs[depth], pk + md.lnf_s, pk + md.lnf_b, xf, st));
  NQS_TRY(gemm(xf, pk + md.bfT, bfo, rows, 2 * KN, d, pk + md.bf_b, (T)0));
  // ---- determinant head
  DetBwdOffsets off;
  off.dlw = leaf("params/det_logw");
  off.dW = leaf("params/dW");
  off.jas_same = leaf("params/jas_c_same");
  off.jas_opp = leaf("params/jas_c_opp");
  off.jas_loga = leaf("params/jastrow_loga");
  const DetParams dp = det_params(md, packed);
  if (!vjp) {
    NQS_TRY(launch_det_bwd<T>(n_cfg, dp, R, S, w0, bfo, qb, out_lp, O, ldo, W_total, off, st));
  } else {
    // the determinant-head leaves into compact rows (both channels), weighted by the cotangent
    DetBwdOffsets so;
    OsmMap mp;
    mp.n = 0;
    int64_t c = 0;
    auto seg = [&](int64_t orig, int64_t len, int64_t& slot) {
      if (orig < 0) {
        slot = -1;
        return;
      }
      slot = c;
      mp.start[mp.n] = c;
      mp.dst[mp.n] = orig;
      ++mp.n;
      c += len;
    };
    const Leaf* lj = find_leaf(md, "params/jas_c_same");
    seg(off.dlw, 2LL * md.K, so.dlw);
    seg(off.dW, 2LL * N * md.n_g * 2, so.dW);
    seg(off.jas_same, lj ? lj->size : 0, so.jas_same);
    seg(off.jas_opp, lj ? lj->size : 0, so.jas_opp);
    seg(off.jas_loga, 1, so.jas_loga);
    mp.start[mp.n] = c;
    NQS_TRY(launch_det_bwd<T>(n_cfg, dp, R, S, w0, bfo, qb, out_lp, shifted(osm_det, ld_det), ld_det, n_cfg, so, st));
    NQS_TRY(launch_vjp_osm_reduce<T>(osm_det, ld_det, n_cfg, mp, vjp_cot, w0, vjp_g, st));
    NQS_TRY(launch_vjp_fold<T>(n_cfg, N, 2 * KN, w0, vjp_cot, qb, st));
  }
  // ---- backflow Dense (2KN real outputs = [Re | Im]; theta: re block then im block)
  const int64_t bk = leaf("params/backflow/kernel"), bb = leaf("params/backflow/bias");
  NQS_TRY(wgrad(wg(xf, d, d, qb, 2 * KN, {bk, bk + (int64_t)d * KN}, KN), st));
  NQS_TRY(bgrad(wg(xf, d, d, qb, 2 * KN, {bb}, 2 * KN), st));
  NQS_TRY(gemm(qb, pk + md.bfW, hb, 2 * rows, d, 2 * KN, nullptr, (T)0));
  NQS_TRY(lnb(xs[depth], hb, xb, pk + md.lnf_s, 0, "params/LayerNorm_0"));
  // ---- blocks in reverse, recomputing each block's activations from its checkpoint
  for (int l = depth - 1; l >= 0; --l) {
    const LayerPack& P = md.layers[l];
    const std::string b = "params/TransformerBlock_" + std::to_string(l) + "/";
    NQS_TRY(block_fwd(l, xs[l], nullptr));
    // MLP out: x_out = xmid + Gv W1 + b1
    NQS_TRY(gemm(xb, pk + P.w1, gb, 2 * rows, ffd, d, nullptr, (T)0));
    NQS_TRY(wgrad(wg(Gv, ffd, ffd, xb, d, {leaf(b + "Dense_1/kernel")}, d), st));
    NQS_TRY(bgrad(wg(Gv, ffd, ffd, xb, d, {leaf(b + "Dense_1/bias")}, d), st));
    NQS_TRY(launch_gelu_bwd<T>(n_cfg, (int64_t)N * ffd, U, gb, gb, st));
    NQS_TRY(gemm(gb, pk + P.w0, hb, 2 * rows, d, ffd, nullptr, (T)0));
    NQS_TRY(wgrad(wg(hB, d, d, gb, ffd, {leaf(b + "Dense_0/kernel")}, ffd), st));
    NQS_TRY(bgrad(wg(hB, d, d, gb, ffd, {leaf(b + "Dense_0/bias")}, ffd), st));
    NQS_TRY(lnb(xmid, hb, xb, pk + P.ln1_s, 1, b + "LayerNorm_1"));
    // attention out: xmid = x_in + attA Wo + bo
    NQS_TRY(gemm(xb, pk + P.o, hb, 2 * rows, d, d, nullptr, (T)0));
    NQS_TRY(wgrad(wg(attA, d, d, xb, d, {leaf(b + "attn/o/kernel")}, d), st));
    NQS_TRY(bgrad(wg(attA, d, d, xb, d, {leaf(b + "attn/o/bias")}, d), st));
    NQS_TRY(launch_attn_bwd<T>(n_cfg, N, d, md.nh, qkv, hb, qb, st));
    NQS_TRY(gemm(qb, pk + P.qkv, hb, 2 * rows, d, 3 * d, nullptr, (T)0));
    NQS_TRY(wgrad(wg(hA, d, d, qb, 3 * d,
                               {leaf(b + "attn/q/kernel"), leaf(b + "attn/k/kernel"), leaf(b + "attn/v/kernel")}, d),
                            st));
    NQS_TRY(bgrad(
        wg(hA, d, d, qb, 3 * d, {leaf(b + "attn/q/bias"), leaf(b + "attn/k/bias"), leaf(b + "attn/v/bias")}, d), st));
    NQS_TRY(lnb(xs[l], hb, xb, pk + P.ln0_s, 1, b + "LayerNorm_0"));
  }
  // ---- embedding
  NQS_TRY(wgrad(wg(X0, md.kpad, md.kin, xb, d, {leaf("params/embed/kernel")}, d), st));
  NQS_TRY(bgrad(wg(X0, md.kpad, md.kin, xb, d, {leaf("params/embed/bias")}, d), st));
  return NQS_OK;
}

template int run_backward<double>(const Model&, const void*, const double*, const int32_t*, int64_t, int64_t, int64_t,
                                  double*, int64_t, double*, char*, hipStream_t, const double*, double*);
template int run_backward<float>(const Model&, const void*, const double*, const int32_t*, int64_t, int64_t, int64_t,
                                 float*, int64_t, double*, char*, hipStream_t, const double*, double*);

}  // namespace nqs
