// Autotuner of the v1/v2 fold schedule.
//
// The schedules differ only in how the fold is mapped onto the GPU: the
// fused kernel (one workgroup per (signal, loop) folding into LDS and
// transforming there) or the signal-major streamed fold with R resident
// workgroups followed by the bucket-FFT batches.  Every lane sums its bucket
// in the same tap order in all of them, so the choice never changes a bit of
// the result; B, the loop counts and the filters are the reference's and are
// not tuned.  Candidates are timed on synthetic noise signals with a private
// RNG stream, so tuning does not advance the plan's own draws.
#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "internal.hpp"

namespace sfa {

namespace {

struct TuneKey {
  int n, k, version, precision;
  bool operator<(const TuneKey &o) const
  {
    return std::tie(n, k, version, precision) < std::tie(o.n, o.k, o.version, o.precision);
  }
};

std::mutex g_tune_mu;
std::map<TuneKey, std::pair<int, int>> g_tuned;  // key -> (fold_mode, stream_blocks)

TuneKey key_of(const Plan *P)
{
  return TuneKey{P->n, P->k, (int)P->version, (int)P->precision};
}

}  // namespace

void apply_tuned_schedule(Plan *P)
{
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tuned.find(key_of(P));
  if (it != g_tuned.end()) {
    P->fold_mode = it->second.first;
    P->stream_blocks = it->second.second;
  }
}

int autotune_plan(Plan *P, int batch)
{
  if (P->version == SFFT_VERSION_3) return 0;  // one schedule only
  const V12Params &p = P->v12;
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  // synthetic batch: at most 8 GiB of input
  const int64_t per = (int64_t)p.n * 16;
  const int S = (int)std::max<int64_t>(1, std::min<int64_t>(std::max(batch, 1), (8LL << 30) / per));
  double2 *x = nullptr;
  if (hipMalloc(&x, (size_t)S * per) != hipSuccess) {
    set_error("autotune: hipMalloc of the synthetic batch failed");
    return -1;
  }
  int rc = k::fill_noise(x, (int64_t)S * p.n, 0x5eed, P->stream);
  std::vector<const double2 *> in(S);
  for (int s = 0; s < S; s++) in[s] = x + (int64_t)s * p.n;
  Rng rng;
  rng.seed(1, 1);
  std::vector<Rng *> rngs(S, &rng);
  std::vector<int64_t> counts(S);
  BatchIO io;
  io.S = S;
  io.rng = rngs.data();
  io.in = in.data();
  io.counts = counts.data();

  std::vector<std::pair<int, int>> cand;
  if (std::max(p.B_loc, p.B_est) <= kMaxLdsFft64) cand.push_back({0, 0});
  for (int r : {1, 2, 4}) cand.push_back({1, r * cus});
  if (P->precision == SFFT_PRECISION_FP32_REFINE && p.comb) cand.push_back({2, cus});
  const int mode0 = P->fold_mode, blocks0 = P->stream_blocks;
  double best = 1e30;
  std::pair<int, int> pick{mode0, blocks0};
  for (const auto &c : cand) {
    if (rc != 0) break;
    P->fold_mode = c.first;
    P->stream_blocks = c.second;
    double t = 1e30;
    for (int rep = 0; rep < 3 && rc == 0; rep++) {  // first run warms up
      const auto t0 = std::chrono::steady_clock::now();
      rc = exec_batch(P, io);
      const double ms =
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      if (rep > 0) t = std::min(t, ms);
    }
    if (rc == 0 && t < best) {
      best = t;
      pick = c;
    }
  }
  (void)hipFree(x);
  if (rc != 0) {
    P->fold_mode = mode0;
    P->stream_blocks = blocks0;
    return -1;
  }
  P->fold_mode = pick.first;
  P->stream_blocks = pick.second;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned[key_of(P)] = pick;
  return 0;
}

}  // namespace sfa
