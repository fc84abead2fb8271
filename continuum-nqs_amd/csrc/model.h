// Model plan: real_ravel parameter layout, packed kernel layout, workspace plans.
#pragma once
#include <string>
#include <vector>
#include "nqs_common.h"
#include "kernels.h"

namespace nqs {

struct Leaf {
  std::string name;  // "params/TransformerBlock_0/attn/q/kernel"
  int64_t shape[3];
  int ndim;
  bool cplx;
  int64_t offset;  // in theta
  int64_t size;    // elements per real/imag part
};

// Offsets (in elements of the backbone type T) of the packed backbone tensors.
struct LayerPack {
  int64_t ln0_s, ln0_b, qkvT, qkv_b, oT, o_b, ln1_s, ln1_b, w0T, b0, w1T, b1;
  int64_t qkv, o, w0, w1;  // non-transposed (in x out) copies for the backward pass
  // fused value-only forward: the transposed weights again as bf16 hi/lo MFMA operand fragments
  // (fragment order, see fused.hip), when Model::fused
  int64_t fz_qkv, fz_o, fz_w0, fz_w1;
};

struct Model {
  nqs_model_desc m;
  int N, d, depth, nh, ff, F, K, n_g, kin, kpad, KN;
  double A[4], Ainv[4];
  int64_t P;
  std::vector<Leaf> leaves;
  // packed layout: backbone part in T elements, then a float64 part (byte offset dbl_off)
  int64_t embT, emb_b, emb, lnf_s, lnf_b, bfT, bf_b, bfW;
  bool fused;              // the fused value-only forward takes this architecture
  int64_t fz_emb, fz_bf;   // its embedding (K padded to 64) and backflow operand fragments
  std::vector<LayerPack> layers;
  int64_t n_backbone;  // elements of T
  int64_t dbl_off;     // bytes
  int64_t gs_off, wenv_off, dlw_off, jas_off, n_dbl;  // offsets in doubles from dbl_off
  int64_t packed_bytes;
  int tsize;           // sizeof(T)
  int split;           // GemmArgs::split of the backbone dense layers (float32 models)
};

int build_model(const nqs_model_desc* m, Model& md);
const Leaf* find_leaf(const Model& md, const std::string& name);

// Per-chunk buffers of a jet forward.
template <typename T>
struct FwdBufs {
  T *X0, *x, *h, *big, *att;
  double *lp, *tan, *lap;
};

DetParams det_params(const Model& md, const void* packed);

template <typename T>
int run_forward(const Model& md, const void* packed, const JetGeom& jg, const double* R, const int32_t* S,
                int64_t w0, const FwdBufs<T>& b, hipStream_t st);

// The fused value-only forward (fused.hip): whether it takes this architecture (float32 bf16x3
// backbone with N = 16, d = 128, 8 heads, d_ff = 256) and this jet (one channel;
// NQS_LAYERWISE_FORWARD=1 in the environment turns it off), the packing of its weight fragments,
// and the launch from the features X0 (rows x kpad) to the backflow activations out (rows x 2 K N).
bool fused_model_ok(const Model& md);
bool fused_value_ok(const Model& md, const JetGeom& jg);
// the spin-flip pass's jets (value and two tangents per configuration) on the fused kernel
bool fused_jet1_ok(const Model& md, const JetGeom& jg);
int launch_fused_jet1(const Model& md, const float* packed, const float* X0, int64_t n_cfg, float* out,
                      hipStream_t st);
int pack_fused_weights(const Model& md, float* packed, hipStream_t st);
int launch_fused_value(const Model& md, const float* packed, const float* X0, int64_t rows, float* out,
                       hipStream_t st);

// bytes of FwdBufs for n_cfg configurations with C channels and nt tangents
int64_t fwd_bytes(const Model& md, int64_t n_cfg, int C, int nt);
template <typename T>
FwdBufs<T> carve_fwd(const Model& md, char*& p, int64_t n_cfg, int C, int nt);

// Per-walker parameter gradient: O rows (Re at w, Im at W_total + w) for walkers
// [w0, w0 + n_cfg), plus log psi of those walkers.  With vjp_g != NULL it computes instead the
// vector-Jacobian product vjp_g += sum_w (cot[2w] d Re L_w / d theta + cot[2w+1] d Im L_w / d theta)
// over these walkers (O and ldo unused; the workspace needs vjp_bytes more).
int64_t grad_bytes(const Model& md, int64_t n_cfg);
int64_t vjp_bytes(const Model& md, int64_t n_cfg);
template <typename T>
int run_backward(const Model& md, const void* packed, const double* R, const int32_t* S, int64_t w0,
                 int64_t n_cfg, int64_t W_total, T* O, int64_t ldo, double* out_lp, char* ws, hipStream_t st,
                 const double* vjp_cot = nullptr, double* vjp_g = nullptr);

}  // namespace nqs
