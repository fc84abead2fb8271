// Launch-shape autotuner.  The first launch of a tuned kernel for a given shape class times each
// candidate launch shape once with hipEvents on the caller's stream, keeps the fastest in a
// process-wide table and uses it from then on.  Tuned kernels are pure functions of their inputs
// (no accumulation into their outputs), so every candidate writes the same result and the call
// stays correct.  Tuning synchronises the stream once per shape class; it is skipped (default
// shape) while the stream is being captured into a graph, and switched off by
// nqs_autotune_enable(0).
#pragma once
#include <functional>
#include <string>
#include "nqs_common.h"

namespace nqs {

struct TuneKey {
  int kernel;  // TUNE_* id
  int64_t a, b, c;
  bool operator<(const TuneKey& o) const {
    if (kernel != o.kernel) return kernel < o.kernel;
    if (a != o.a) return a < o.a;
    if (b != o.b) return b < o.b;
    return c < o.c;
  }
};

enum { TUNE_ATTENTION = 1, TUNE_DET16 = 2, TUNE_DET32 = 3 };

// Launch candidate `choice` (0..n-1) and return NQS_OK or an error code.
int tune_run(const TuneKey& key, int n_candidates, int default_choice, const char* const* names,
             const std::function<int(int)>& launch, hipStream_t st);

int64_t tune_bucket(int64_t n);  // log2 size class

}  // namespace nqs
