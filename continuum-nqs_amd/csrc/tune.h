// Launch-shape autotuner.  The first launch of a tuned kernel for a given shape class times each
// candidate launch shape once with hipEvents on the caller's stream, keeps the fastest in a
// process-wide table and uses it from then on: waves per workgroup of the attention jet and the
// determinant heads, tile shapes of the bf16 dense-layer GEMM.  Every candidate computes the same
// result bit for bit; kernels that accumulate into their output run their trials into a scratch
// buffer (tune_run_trials).  Tuning synchronises the stream once per shape class; it is skipped
// (default shape) while the stream is being captured into a graph, and switched off by
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

enum { TUNE_ATTENTION = 1, TUNE_DET16 = 2, TUNE_DET32 = 3, TUNE_GEMM_BF = 4 };

// Launch candidate `choice` (0..n-1) and return NQS_OK or an error code.
int tune_run(const TuneKey& key, int n_candidates, int default_choice, const char* const* names,
             const std::function<int(int)>& launch, hipStream_t st);

// As tune_run for kernels whose trial launches cannot write the real output (they accumulate into
// it): trials call launch(c, true), and the chosen candidate is then always launched with
// launch(choice, false).  With trials_ok = false a missing entry is not tuned: the default runs.
int tune_run_trials(const TuneKey& key, int n_candidates, int default_choice, const char* const* names,
                    const std::function<int(int, bool)>& launch, bool trials_ok, hipStream_t st);

int64_t tune_bucket(int64_t n);  // log2 size class

}  // namespace nqs
