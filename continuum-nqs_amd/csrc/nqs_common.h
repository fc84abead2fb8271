// Shared device/host helpers for the gfx950 VMC kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include <string>

#include "../../include/nqs_hip.h"

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef uint32_t u4v __attribute__((ext_vector_type(4)));

namespace nqs {

// ---- error plumbing: thread-local last error, no C++ exceptions across the ABI
void set_error(const std::string& msg);
#define NQS_CHECK_HIP(expr)                                                            \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess) {                                                            \
      ::nqs::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));             \
      return NQS_ERR_HIP;                                                              \
    }                                                                                  \
  } while (0)
#define NQS_REQUIRE(cond, code, msg)                                                   \
  do {                                                                                 \
    if (!(cond)) {                                                                     \
      ::nqs::set_error(msg);                                                           \
      return code;                                                                     \
    }                                                                                  \
  } while (0)
#define NQS_LAUNCH_CHECK()                                                             \
  do {                                                                                 \
    hipError_t _e = hipGetLastError();                                                 \
    if (_e != hipSuccess) {                                                            \
      ::nqs::set_error(std::string("kernel launch: ") + hipGetErrorString(_e));        \
      return NQS_ERR_HIP;                                                              \
    }                                                                                  \
  } while (0)

#define NQS_TRY(x)                 \
  do {                             \
    int _rc = (x);                 \
    if (_rc != NQS_OK) return _rc; \
  } while (0)

static inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
__device__ inline int64_t cdiv_dev(int64_t a, int64_t b) { return (a + b - 1) / b; }
static inline int64_t round_up(int64_t a, int64_t b) { return cdiv(a, b) * b; }

// ---- complex double arithmetic on a plain struct (no std::complex on device)
struct cd {
  double re, im;
};
__host__ __device__ inline cd mk(double r, double i) { return cd{r, i}; }
__host__ __device__ inline cd operator+(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline cd operator-(cd a, cd b) { return cd{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline cd operator*(cd a, cd b) {
  return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline cd operator*(double s, cd a) { return cd{s * a.re, s * a.im}; }
__host__ __device__ inline cd cdiv_(cd a, cd b) {
  // Smith's algorithm for a / b
  if (fabs(b.re) >= fabs(b.im)) {
    double r = b.im / b.re, den = b.re + b.im * r;
    return cd{(a.re + a.im * r) / den, (a.im - a.re * r) / den};
  } else {
    double r = b.re / b.im, den = b.re * r + b.im;
    return cd{(a.re * r + a.im) / den, (a.im * r - a.re) / den};
  }
}
__host__ __device__ inline double cabs_(cd a) { return hypot(a.re, a.im); }
__device__ inline cd cexp_(cd a) {
  double e = exp(a.re), s, c;
  sincos(a.im, &s, &c);
  return cd{e * c, e * s};
}
__device__ inline cd clog_(cd a) { return cd{log(hypot(a.re, a.im)), atan2(a.im, a.re)}; }

// ---- wave-level reductions (64 lanes)
// One DPP move of a 32- or 64-bit value inside each 16-lane row (dpp_ctrl CTRL: quad_perm,
// row_mirror, row_half_mirror, ...), on the VALU without LDS traffic.
template <int CTRL, typename T>
__device__ __forceinline__ T row_dpp(T x) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xF, 0xF, false));
  } else {
    const long long b = __builtin_bit_cast(long long, x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(T, (long long)(unsigned)lo | ((long long)hi << 32));
  }
}

// Sum over the 64 lanes, the same bits in every lane: pairwise within quads (quad_perm [1,0,3,2]
// and [2,3,0,1]), quad with quad (row_half_mirror: lane i with 7 - i), half-row with half-row
// (row_mirror: i with 15 - i) as DPP, then row with row by two shuffles (lane i with i ^ 16, then
// i ^ 32).  Every step adds the values of two lanes that pair with each other, so both get the
// same sum.
template <typename T>
__device__ inline T wave_sum(T v) {
  v += row_dpp<0xB1>(v);
  v += row_dpp<0x4E>(v);
  v += row_dpp<0x141>(v);
  v += row_dpp<0x140>(v);
  v += __shfl_xor(v, 16, 64);
  v += __shfl_xor(v, 32, 64);
  return v;
}

}  // namespace nqs
