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
template <typename T>
__device__ inline T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

}  // namespace nqs
