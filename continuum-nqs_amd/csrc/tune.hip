// Launch-shape autotuner (see tune.h).
#include <map>
#include <mutex>
#include <sstream>
#include "tune.h"

namespace nqs {

namespace {
struct Entry {
  int choice;
  std::string report;
};
std::mutex g_mu;
std::map<TuneKey, Entry> g_table;
bool g_enabled = true;
}  // namespace

int64_t tune_bucket(int64_t n) {
  int64_t b = 0;
  while ((1LL << (b + 1)) <= n) ++b;
  return b;
}

int tune_run(const TuneKey& key, int n, int def, const char* const* names, const std::function<int(int)>& launch,
             hipStream_t st) {
  return tune_run_trials(key, n, def, names, [&](int c, bool trial) { return launch(c); }, true, st);
}

int tune_run_trials(const TuneKey& key, int n, int def, const char* const* names,
                    const std::function<int(int, bool)>& launch_t, bool trials_ok, hipStream_t st) {
  auto launch = [&](int c) { return launch_t(c, true); };
  int choice = -1;
  bool enabled;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    enabled = g_enabled;
    auto it = g_table.find(key);
    if (it != g_table.end()) choice = it->second.choice;
  }
  if (choice >= 0) return launch_t(choice, false);
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  NQS_CHECK_HIP(hipStreamIsCapturing(st, &cs));
  if (!enabled || !trials_ok || cs != hipStreamCaptureStatusNone || n <= 1) return launch_t(def, false);
  hipEvent_t e[2];
  NQS_CHECK_HIP(hipEventCreate(&e[0]));
  NQS_CHECK_HIP(hipEventCreate(&e[1]));
  float best = 1e30f;
  int best_c = def;
  std::ostringstream rep;
  rep << "kernel " << key.kernel << " shape (" << key.a << ", " << key.b << ", " << key.c << "):";
  for (int c = 0; c < n; ++c) {
    NQS_CHECK_HIP(hipEventRecord(e[0], st));
    const int rc = launch(c);
    NQS_CHECK_HIP(hipEventRecord(e[1], st));
    NQS_CHECK_HIP(hipEventSynchronize(e[1]));
    float ms = 0;
    NQS_CHECK_HIP(hipEventElapsedTime(&ms, e[0], e[1]));
    rep << " " << names[c] << "=" << (rc == NQS_OK ? std::to_string(ms) : std::string("failed")) << "ms";
    if (rc == NQS_OK && ms < best) {
      best = ms;
      best_c = c;
    }
  }
  NQS_CHECK_HIP(hipEventDestroy(e[0]));
  NQS_CHECK_HIP(hipEventDestroy(e[1]));
  rep << " -> " << names[best_c];
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_table[key] = Entry{best_c, rep.str()};
  }
  // the kept choice writes the outputs (for pure kernels the trials already did; once per class)
  return launch_t(best_c, false);
}

}  // namespace nqs

extern "C" {

int nqs_autotune_enable(int32_t enable) {
  std::lock_guard<std::mutex> lk(nqs::g_mu);
  nqs::g_enabled = enable != 0;
  return NQS_OK;
}

int nqs_autotune_reset(void) {
  std::lock_guard<std::mutex> lk(nqs::g_mu);
  nqs::g_table.clear();
  return NQS_OK;
}

int64_t nqs_autotune_report(char* buf, int64_t len) {
  std::string s;
  {
    std::lock_guard<std::mutex> lk(nqs::g_mu);
    for (const auto& kv : nqs::g_table) s += kv.second.report + "\n";
  }
  if (buf && len > 0) {
    const int64_t n = std::min<int64_t>(len - 1, (int64_t)s.size());
    s.copy(buf, n);
    buf[n] = 0;
  }
  return (int64_t)s.size();
}

}  // extern "C"
