// extern "C" entry points: the reference API (src/sfft.h:158-167) plus the
// sparse / precision / RNG / stream extensions declared in include/sfft.h.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "internal.hpp"

namespace sfa {
const char *last_error();
}

using namespace sfa;

namespace {

Plan *get(const sfft_plan *plan)
{
  if (!plan || !plan->data) {
    set_error("null plan");
    return nullptr;
  }
  return reinterpret_cast<Plan *>(plan->data);
}

// Calls on one plan from several threads take turns (the reference let
// OpenMP threads call sfft_exec on one plan concurrently).
std::unique_lock<std::mutex> hold(Plan *P) { return std::unique_lock<std::mutex>(P->mu); }


// Signals per GPU batch: bounded by workspace and, for host inputs, by the
// staging buffer (whole signals are copied to the device).
int chunk_signals(const Plan *P, bool host_inputs)
{
  const int64_t budget = 4LL << 30;
  int64_t per = std::max<int64_t>(1, P->bytes_per_signal());
  if (host_inputs) per += (int64_t)P->n * 16;
  return (int)std::max<int64_t>(1, std::min<int64_t>(4096, budget / per));
}

// Runs num signals.  Exactly one of (dense, sparse) is non-null.
int run(Plan *P, int num, const complex_t *const *in, complex_t *const *dense,
        sfft_output_c *sparse)
{
  P->last_status.clear();
  if (num <= 0) return 0;
  const int64_t n = P->n;
  const int64_t cap = P->output_capacity();
  bool host_in = false;
  PtrKinds kinds;
  std::vector<char> in_dev(num);
  for (int s = 0; s < num; s++) {
    if (!in[s]) {
      set_error("null input signal");
      return -1;
    }
    in_dev[s] = kinds.is_device(in[s]);
    host_in |= !in_dev[s];
  }
  const int chunk = chunk_signals(P, host_in);
  hipStream_t st = P->stream;
  // on an error, host dense outputs not yet delivered read as zero (the
  // reference zero-fills before anything else, src/sfft.cc:121-123)
  auto fail = [&](int from) {
    if (dense)
      for (int s = from; s < num; s++)
        if (dense[s] && !kinds.is_device(dense[s])) std::memset(dense[s], 0, (size_t)n * 16);
    return -1;
  };
  bool partial = false;  // a v3 map overflowed: results written, error set
  for (int c0 = 0; c0 < num; c0 += chunk) {
    const int S = std::min(chunk, num - c0);
    std::vector<const double2 *> ins(S);
    std::vector<int32_t *> fr(S, nullptr);
    std::vector<double2 *> va(S, nullptr), de(S, nullptr);
    std::vector<int64_t> counts(S, 0);
    std::vector<char> via_ws(S, 0);  // results staged in the plan's buffers
    // inputs
    int staged = 0;
    for (int s = 0; s < S; s++) staged += !in_dev[c0 + s];
    if (staged && !P->ws->stage.ensure((size_t)staged * n * 16)) return fail(c0);
    int slot = 0;
    for (int s = 0; s < S; s++) {
      const complex_t *src = in[c0 + s];
      if (in_dev[c0 + s]) {
        ins[s] = reinterpret_cast<const double2 *>(src);
      } else {
        double2 *dst = P->ws->stage.as<double2>() + (size_t)slot++ * n;
        if (hipMemcpyAsync(dst, src, n * 16, hipMemcpyHostToDevice, st) != hipSuccess) {
          set_error("input copy failed");
          return fail(c0);
        }
        ins[s] = dst;
      }
    }
    // outputs
    int nws = 0;
    for (int s = 0; s < S; s++) {
      if (dense) {
        complex_t *o = dense[c0 + s];
        if (!o) {
          set_error("null output signal");
          return fail(c0);
        }
        if (kinds.is_device(o)) {
          de[s] = reinterpret_cast<double2 *>(o);
          if (k::zero_dense(de[s], n, st) != 0) return fail(c0);
        } else {
          via_ws[s] = 1;
          nws++;
        }
      } else {
        sfft_output_c &o = sparse[c0 + s];
        if (o.freq && o.val && o.capacity >= cap && kinds.is_device(o.freq) &&
            kinds.is_device(o.val)) {
          fr[s] = o.freq;
          va[s] = reinterpret_cast<double2 *>(o.val);
        } else {
          via_ws[s] = 1;
          nws++;
        }
      }
    }
    if (nws) {
      if (!P->ws->out_freq.ensure((size_t)nws * cap * 4) ||
          !P->ws->out_val.ensure((size_t)nws * cap * 16))
        return fail(c0);
      int q = 0;
      for (int s = 0; s < S; s++) {
        if (!via_ws[s]) continue;
        fr[s] = P->ws->out_freq.as<int32_t>() + (size_t)q * cap;
        va[s] = P->ws->out_val.as<double2>() + (size_t)q * cap;
        q++;
      }
    }
    BatchIO io;
    io.S = S;
    io.in = ins.data();
    const bool any_sparse = nws > 0 || !dense;
    io.freq = any_sparse ? fr.data() : nullptr;
    io.val = any_sparse ? va.data() : nullptr;
    io.dense = dense ? de.data() : nullptr;
    io.cap = cap;
    io.counts = counts.data();
    const int rc = exec_batch(P, io);
    if (rc < 0) return fail(c0);
    partial |= rc > 0;
    // results that went through the plan's buffers
    for (int s = 0; s < S; s++) {
      const int64_t cnt = std::min(counts[s], cap);
      if (dense) {
        if (!via_ws[s]) continue;
        std::vector<int32_t> hf(cnt);
        std::vector<double2> hv(cnt);
        if (cnt && (hipMemcpy(hf.data(), fr[s], cnt * 4, hipMemcpyDeviceToHost) != hipSuccess ||
                    hipMemcpy(hv.data(), va[s], cnt * 16, hipMemcpyDeviceToHost) != hipSuccess)) {
          set_error("result copy failed");
          return fail(c0 + s);
        }
        double2 *o = reinterpret_cast<double2 *>(dense[c0 + s]);
        std::memset(o, 0, (size_t)n * 16);  // src/sfft.cc:121-123
        for (int64_t j = 0; j < cnt; j++) o[hf[j]] = hv[j];
      } else {
        sfft_output_c &o = sparse[c0 + s];
        o.count = counts[s];
        if (!via_ws[s]) continue;
        const int64_t m = std::min<int64_t>(cnt, std::max<int64_t>(0, o.capacity));
        if (m && o.freq &&
            hipMemcpy(o.freq, fr[s], m * 4, hipMemcpyDefault) != hipSuccess) {
          set_error("result copy failed");
          return -1;
        }
        if (m && o.val &&
            hipMemcpy(o.val, va[s], m * 16, hipMemcpyDefault) != hipSuccess) {
          set_error("result copy failed");
          return -1;
        }
      }
    }
  }
  return partial ? -1 : 0;
}

}  // namespace

extern "C" {

void *sfft_malloc(size_t s)
{
  void *p = nullptr;
  if (posix_memalign(&p, 16, s ? s : 16) != 0) return nullptr;
  return p;
}

void sfft_free(void *p) { free(p); }

sfft_plan *sfft_make_plan(int n, int k, sfft_version version,
                          int fftw_optimization)
{
  set_error("");
  Plan *P = make_plan(n, k, version);
  if (!P) return nullptr;
  if (fftw_optimization == 0) {  // FFTW_MEASURE: tune now
    if (autotune_plan(P, 64) != 0) {
      free_plan(P);
      return nullptr;
    }
  } else {
    apply_tuned_schedule(P);
  }
  sfft_plan *plan = (sfft_plan *)std::malloc(sizeof(sfft_plan));
  if (!plan) {
    free_plan(P);
    set_error("out of host memory");
    return nullptr;
  }
  plan->version = version;
  plan->n = (unsigned)n;
  plan->k = (unsigned)k;
  plan->data = P;
  return plan;
}

void sfft_free_plan(sfft_plan *plan)
{
  if (!plan) return;
  free_plan(reinterpret_cast<Plan *>(plan->data));
  std::free(plan);
}

void sfft_exec(sfft_plan *plan, complex_t *in, complex_t *out)
{
  set_error("");
  Plan *P = get(plan);
  if (!P) return;
  auto lk = hold(P);
  const complex_t *ins[1] = {in};
  complex_t *outs[1] = {out};
  (void)run(P, 1, ins, outs, nullptr);
}

void sfft_exec_many(sfft_plan *plan, int num, complex_t **in, complex_t **out)
{
  set_error("");
  Plan *P = get(plan);
  if (!P) return;
  auto lk = hold(P);
  (void)run(P, num, in, out, nullptr);
}

int64_t sfft_plan_output_capacity(const sfft_plan *plan)
{
  Plan *P = get(plan);
  return P ? P->output_capacity() : -1;
}

int sfft_exec_sparse(sfft_plan *plan, const complex_t *in, sfft_output_c *out)
{
  set_error("");
  Plan *P = get(plan);
  if (!P || !out) return -1;
  auto lk = hold(P);
  const complex_t *ins[1] = {in};
  return run(P, 1, ins, nullptr, out);
}

int sfft_exec_many_sparse(sfft_plan *plan, int num, const complex_t *const *in,
                          sfft_output_c *out)
{
  set_error("");
  Plan *P = get(plan);
  if (!P || !out) return -1;
  auto lk = hold(P);
  return run(P, num, in, nullptr, out);
}

int sfft_exec_sparse_noalloc(sfft_plan *plan, const complex_t *in, sfft_output_c *out)
{
  set_error("");
  Plan *P = get(plan);
  if (!P || !out) {
    if (P) set_error("null output");
    return -1;
  }
  if (P->version == SFFT_VERSION_3) {
    set_error("sfft_exec_sparse_noalloc runs v1/v2 plans only");
    return -1;
  }
  const int64_t cap = P->output_capacity();
  {
    PtrKinds kinds;  // fixed-size cache: no allocation
    if (!kinds.covers(in, (size_t)P->n * 16) || !kinds.covers(out->freq, (size_t)cap * 4) ||
        !kinds.covers(out->val, (size_t)cap * 16) || out->capacity < cap) {
      set_error("sfft_exec_sparse_noalloc takes a device input of n values and device "
                "outputs of sfft_plan_output_capacity entries");
      return -1;
    }
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    set_error("no current device");
    return -1;
  }
  void *scr = k::noalloc_scratch(dev);
  if (!scr) {
    set_error("no scratch region on this device: build a plan on it first");
    return -1;
  }
  auto lk = hold(P);
  // the workspace of this call: views into the device scratch region and a
  // thread-local host staging buffer (params, pointer tables, count)
  alignas(256) static thread_local char host_stage[4096];
  Arena arena;
  arena.base = static_cast<char *>(scr);
  arena.size = k::kNoAllocScratch;
  Workspace ws;
  ws.use_arena(&arena);
  ws.pinned = host_stage;
  ws.pinned_bytes = sizeof(host_stage);
  ws.pinned_view = true;
  const double2 *ins[1] = {reinterpret_cast<const double2 *>(in)};
  int32_t *frs[1] = {out->freq};
  double2 *vas[1] = {reinterpret_cast<double2 *>(out->val)};
  int64_t count = 0;
  BatchIO io;
  io.S = 1;
  io.in = ins;
  io.freq = frs;
  io.val = vas;
  io.cap = out->capacity;
  io.counts = &count;
  io.fixed_ws = true;
  Workspace *own = P->ws;
  P->ws = &ws;
  const int rc = exec_batch(P, io);
  P->ws = own;
  P->last_S = 0;  // the buckets of this call live in the scratch region
  if (rc == 0) out->count = count;
  return rc;
}

int sfft_exec_many_plans(int num, sfft_plan *const *plans, const complex_t *const *in,
                         sfft_output_c *out)
{
  set_error("");
  if (num <= 0) return 0;
  if (!plans || !in || !out) {
    set_error("null argument");
    return -1;
  }
  // group the signals by plan, in order of first appearance
  std::vector<Plan *> order;
  std::vector<std::vector<int>> members;
  PtrKinds kinds;
  for (int i = 0; i < num; i++) {
    Plan *P = get(plans[i]);
    if (!P) return -1;
    if (!in[i] || !kinds.is_device(in[i]) || !out[i].freq || !out[i].val ||
        !kinds.is_device(out[i].freq) || !kinds.is_device(out[i].val)) {
      set_error("sfft_exec_many_plans takes device inputs and device sparse outputs");
      return -1;
    }
    if (out[i].capacity < P->output_capacity()) {
      set_error("output capacity below the plan's");
      return -1;
    }
    size_t g = 0;
    while (g < order.size() && order[g] != P) g++;
    if (g == order.size()) {
      order.push_back(P);
      members.emplace_back();
    }
    members[g].push_back(i);
  }
  const size_t G = order.size();
  std::vector<Plan *> by_addr(order);
  std::sort(by_addr.begin(), by_addr.end());
  std::vector<std::unique_lock<std::mutex>> locks;
  for (Plan *P : by_addr) locks.push_back(hold(P));  // address order: no deadlock
  for (Plan *P : order) P->last_status.clear();
  std::vector<std::vector<const double2 *>> ins(G);
  std::vector<std::vector<int32_t *>> frs(G);
  std::vector<std::vector<double2 *>> vas(G);
  std::vector<std::vector<int64_t>> counts(G);
  std::vector<BatchIO> ios(G);
  std::vector<Plan *> v3plans;
  std::vector<BatchIO> v3ios;
  // v1/v2 groups are enqueued one after another without waiting (each on
  // its plan's stream, from its plan's staging), then waited for together
  std::vector<V12Pending> pend(G);
  std::vector<size_t> v12g;
  auto drain = [&]() {  // on an error: nothing of this call left in flight
    for (size_t g : v12g) (void)hipStreamSynchronize(order[g]->stream);
    return -1;
  };
  for (size_t g = 0; g < G; g++) {
    for (int i : members[g]) {
      ins[g].push_back(reinterpret_cast<const double2 *>(in[i]));
      frs[g].push_back(out[i].freq);
      vas[g].push_back(reinterpret_cast<double2 *>(out[i].val));
    }
    counts[g].assign(members[g].size(), 0);
    BatchIO &io = ios[g];
    io.S = (int)members[g].size();
    io.in = ins[g].data();
    io.freq = frs[g].data();
    io.val = vas[g].data();
    io.cap = order[g]->output_capacity();
    io.counts = counts[g].data();
    if (order[g]->version == SFFT_VERSION_3) {
      v3plans.push_back(order[g]);
      v3ios.push_back(io);
    } else {
      if (exec_batch_async(order[g], io, pend[g]) != 0) return drain();
      v12g.push_back(g);
    }
  }
  int v3rc = 0;
  if (!v3plans.empty()) {
    // one launch on the first v3 plan's stream; it first waits for the work
    // the caller enqueued on the other v3 plans' streams (their inputs).  A
    // v3 plan's private stream holds no work between calls (every v3
    // exec synchronises it before returning), so it is not waited for.
    hipStream_t st = v3plans[0]->stream;
    std::vector<hipStream_t> waited;
    for (Plan *P : v3plans) {
      if (P->stream == st || P->own_stream ||
          std::find(waited.begin(), waited.end(), P->stream) != waited.end())
        continue;
      waited.push_back(P->stream);
      hipEvent_t ev = nullptr;
      if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        set_error("sfft_exec_many_plans: hipEventCreate failed");
        return drain();
      }
      const bool ok = hipEventRecord(ev, P->stream) == hipSuccess &&
                      hipStreamWaitEvent(st, ev, 0) == hipSuccess;
      (void)hipEventDestroy(ev);  // released once the wait has passed
      if (!ok) {
        set_error("sfft_exec_many_plans: stream ordering failed");
        return drain();
      }
    }
    v3rc = exec_multi_v3(v3plans.data(), v3ios.data(), (int)v3plans.size(), st,
                         v3plans[0]->ws);
    if (v3rc < 0) return drain();
  }
  for (size_t g : v12g)
    if (exec_batch_wait(order[g], ios[g], pend[g]) != 0) return drain();
  // with a v3 map overflow (v3rc 1) every count and result is written, the
  // overflowed signals' truncated (status bit 1), and the call fails
  for (size_t g = 0; g < G; g++)
    for (size_t j = 0; j < members[g].size(); j++) out[members[g][j]].count = counts[g][j];
  return v3rc ? -1 : 0;
}

int sfft_plan_set_schedule(sfft_plan *plan, int fold_mode, int stream_blocks)
{
  Plan *P = get(plan);
  if (!P) return -1;
  auto lk = hold(P);
  if (fold_mode < 0 || fold_mode > 2 || stream_blocks < 0 || stream_blocks > (1 << 20)) {
    set_error("fold_mode must be 0, 1 or 2 and stream_blocks in [0, 2^20]");
    return -1;
  }
  P->fold_mode = fold_mode;
  P->stream_blocks = stream_blocks;
  return 0;
}

int sfft_plan_get_schedule(const sfft_plan *plan, int *fold_mode, int *stream_blocks)
{
  Plan *P = get(plan);
  if (!P) return -1;
  auto lk = hold(P);
  if (fold_mode) *fold_mode = P->fold_mode;
  if (stream_blocks) *stream_blocks = P->stream_blocks;
  return 0;
}

int sfft_plan_enable_graphs(sfft_plan *plan, int enabled)
{
  Plan *P = get(plan);
  if (!P) return -1;
  auto lk = hold(P);
  P->use_graphs = enabled != 0;
  return 0;
}

int sfft_plan_autotune(sfft_plan *plan, int batch)
{
  set_error("");
  Plan *P = get(plan);
  if (!P) return -1;
  auto lk = hold(P);
  return autotune_plan(P, batch);
}

int sfft_plan_set_precision(sfft_plan *plan, sfft_precision precision)
{
  Plan *P = get(plan);
  if (!P) return -1;
  auto lk = hold(P);
  if (precision != SFFT_PRECISION_FP64 && precision != SFFT_PRECISION_FP32_REFINE) {
    set_error("unknown precision");
    return -1;
  }
  P->precision = precision;
  return 0;
}

int sfft_plan_seed(sfft_plan *plan, unsigned int random_seed, long drand48_seed)
{
  Plan *P = get(plan);
  if (!P) return -1;
  auto lk = hold(P);
  P->rng.seed(random_seed, drand48_seed);
  return 0;
}

int sfft_plan_set_stream(sfft_plan *plan, void *hip_stream)
{
  Plan *P = get(plan);
  if (!P) return -1;
  auto lk = hold(P);
  (void)hipStreamSynchronize(P->stream);
  if (P->own_stream && P->stream) (void)hipStreamDestroy(P->stream);
  P->own_stream = false;
  P->stream = reinterpret_cast<hipStream_t>(hip_stream);
  if (!P->stream) {
    if (hipStreamCreate(&P->stream) != hipSuccess) {
      set_error("hipStreamCreate failed");
      return -1;
    }
    P->own_stream = true;
  }
  return 0;
}

int sfft_plan_set_reference_conjugate(sfft_plan *plan, int enabled)
{
  Plan *P = get(plan);
  if (!P) return -1;
  auto lk = hold(P);
  P->ref_conj = enabled != 0;
  return 0;
}

int sfft_plan_memory(const sfft_plan *plan, int64_t *plan_bytes,
                     int64_t *bytes_per_signal)
{
  Plan *P = get(plan);
  if (!P) return -1;
  auto lk = hold(P);
  if (plan_bytes) *plan_bytes = P->plan_bytes();
  if (bytes_per_signal) *bytes_per_signal = P->bytes_per_signal();
  return 0;
}

const char *sfft_last_error(void) { return sfa::last_error(); }

int sfft_plan_params(const sfft_plan *plan, int *p10)
{
  Plan *P = get(plan);
  if (!P || !p10) return -1;
  std::memset(p10, 0, 10 * sizeof(int));
  if (P->version == SFFT_VERSION_3) {
    const V3Params &q = P->v3;
    int v[10] = {q.W_Man, q.B_g1, q.B_g2, q.w_g1, q.w_g2, 0, 0, 0, 0, 0};
    std::memcpy(p10, v, sizeof(v));
  } else {
    const V12Params &q = P->v12;
    int v[10] = {q.B_loc, q.B_est, q.B_thresh, q.W_comb, q.comb_loops,
                 q.loc_loops, q.est_loops, q.thr_loops, q.w_loc, q.w_est};
    std::memcpy(p10, v, sizeof(v));
  }
  return 0;
}

int sfft_derive_params(int n, int k, sfft_version version, int *p10)
{
  if (!p10) return -1;
  std::memset(p10, 0, 10 * sizeof(int));
  std::string why;
  if (version == SFFT_VERSION_3) {
    const V3Params q = derive_v3(n, k);
    int v[10] = {q.W_Man, q.B_g1, q.B_g2, q.w_g1, q.w_g2, 0, 0, 0, 0, 0};
    std::memcpy(p10, v, sizeof(v));
    why = (n > 0 && (n & (n - 1)) == 0) ? check_v3(q) : "n must be a power of two";
  } else if (version == SFFT_VERSION_1 || version == SFFT_VERSION_2) {
    const V12Params q = derive_v12(n, k, version == SFFT_VERSION_2);
    int v[10] = {q.B_loc, q.B_est, q.B_thresh, q.W_comb, q.comb_loops,
                 q.loc_loops, q.est_loops, q.thr_loops, q.w_loc, q.w_est};
    std::memcpy(p10, v, sizeof(v));
    why = (n > 0 && (n & (n - 1)) == 0) ? check_v12(q) : "n must be a power of two";
  } else {
    set_error("unknown sfft_version");
    return -1;
  }
  if (!why.empty()) {
    set_error(why);
    return 1;
  }
  return 0;
}

int sfft_plan_enable_timing(sfft_plan *plan, int enabled)
{
  Plan *P = get(plan);
  if (!P) return -1;
  auto lk = hold(P);
  P->timing = enabled != 0;
  for (int i = 0; i < Plan::kStages; i++) {
    P->stage_ms[i] = 0;
    P->stage_launches[i] = 0;
  }
  return 0;
}

int sfft_plan_last_status(const sfft_plan *plan, int *flags, int num)
{
  Plan *P = get(plan);
  if (!P) return -1;
  auto lk = hold(P);
  const int m = (int)P->last_status.size();
  for (int i = 0; i < std::min(m, std::max(num, 0)) && flags; i++) flags[i] = P->last_status[i];
  return m;
}

int sfft_plan_stage_times(const sfft_plan *plan, double *ms6, int64_t *launches6)
{
  Plan *P = get(plan);
  if (!P) return -1;
  auto lk = hold(P);
  for (int i = 0; i < Plan::kStages; i++) {
    if (ms6) ms6[i] = P->stage_ms[i];
    if (launches6) launches6[i] = P->stage_launches[i];
  }
  return 0;
}

int sfft_plan_debug_comb_votes(sfft_plan *plan, int enabled)
{
  Plan *P = get(plan);
  if (!P) return -1;
  if (P->version != SFFT_VERSION_2) {
    set_error("comb votes exist in v2 plans only");
    return -1;
  }
  auto lk = hold(P);
  P->debug_votes = enabled != 0;
  return 0;
}

int sfft_plan_debug_v3(sfft_plan *plan, int serial_subtractions, int map_limit)
{
  Plan *P = get(plan);
  if (!P) return -1;
  if (P->version != SFFT_VERSION_3) {
    set_error("the v3 diagnostics apply to v3 plans only");
    return -1;
  }
  if (map_limit < 0) {
    set_error("map_limit must be >= 0");
    return -1;
  }
  auto lk = hold(P);
  P->debug_v3_serial = serial_subtractions != 0;
  P->debug_v3_map_limit = map_limit;
  return 0;
}

int sfft_plan_debug_v3_inline_pre(sfft_plan *plan, int enabled)
{
  Plan *P = get(plan);
  if (!P) return -1;
  if (P->version != SFFT_VERSION_3) {
    set_error("the v3 diagnostics apply to v3 plans only");
    return -1;
  }
  auto lk = hold(P);
  P->debug_v3_inline_pre = enabled != 0;
  return 0;
}

int64_t sfft_debug_comb_votes(const sfft_plan *plan, int signal, int32_t *hits, int64_t cap)
{
  Plan *P = get(plan);
  if (!P || signal < 0 || (!hits && cap > 0)) return -1;
  auto lk = hold(P);
  if (signal >= P->last_S || P->votes_stride == 0) {
    set_error("no comb votes recorded for this signal (enable them on a v2 FP64 plan first)");
    return -1;
  }
  const int64_t stride = P->votes_stride;
  std::vector<int64_t> kv((size_t)stride);
  if (hipMemcpy(kv.data(), P->ws->votes.as<int64_t>() + (size_t)signal * stride, stride * 8,
                hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("vote copy failed");
    return -1;
  }
  std::vector<std::pair<int64_t, int32_t>> app;
  for (int64_t i = 0; i < stride / 2; i++)
    if (kv[2 * i] >= 0) app.push_back({kv[2 * i], (int32_t)kv[2 * i + 1]});
  std::sort(app.begin(), app.end());
  for (int64_t i = 0; i < std::min<int64_t>(cap, (int64_t)app.size()); i++) hits[i] = app[i].second;
  return (int64_t)app.size();
}

int64_t sfft_debug_buckets(const sfft_plan *plan, int signal, complex_t *dst,
                           int64_t count)
{
  Plan *P = get(plan);
  if (!P || !dst || P->version == SFFT_VERSION_3 || signal < 0) return -1;
  auto lk = hold(P);
  const int64_t xs = P->v12.x_samp_size();
  if (signal >= P->last_S) {
    set_error("no such signal in the last batch");
    return -1;
  }
  const bool fp32 = P->last_fp32;
  // loops [0, n64) are FP64 buffers, the rest FP32 (see exec_batch)
  const int64_t n64 = !fp32 ? xs : (!P->v12.comb ? (int64_t)P->v12.loc_loops * P->v12.B_loc : 0);
  const int64_t m = std::min(count, xs);
  std::vector<double2> t64(std::min(m, n64));
  std::vector<float2> t32(std::max<int64_t>(0, m - n64));
  if ((!t64.empty() &&
       hipMemcpy(t64.data(), P->ws->xsamp.as<double2>() + (size_t)signal * xs,
                 t64.size() * 16, hipMemcpyDeviceToHost) != hipSuccess) ||
      (!t32.empty() &&
       hipMemcpy(t32.data(), P->ws->xsamp32.as<float2>() + (size_t)signal * xs + n64,
                 t32.size() * 8, hipMemcpyDeviceToHost) != hipSuccess)) {
    set_error("bucket copy failed");
    return -1;
  }
  for (size_t i = 0; i < t64.size(); i++) dst[i] = complex_t(t64[i].x, t64[i].y);
  for (size_t i = 0; i < t32.size(); i++) dst[n64 + i] = complex_t(t32[i].x, t32[i].y);
  if (P->last_rows_layout && m == xs) {
    // back to natural bucket order: bucket b = r + d*q of loop l was stored
    // at r*rows + (q * u^-1 mod rows), u = ai mod rows (the batch's draws
    // are still in the staging buffer)
    const V12Params &p = P->v12;
    const int L = p.loops(), stride = 2 * L + std::max(p.comb_loops, 1);
    const int *hp = reinterpret_cast<const int *>(P->ws->pinned);
    const int rows = p.n / p.W_comb;
    std::vector<complex_t> tmp;
    int64_t off = 0;
    for (int l = 0; l < L; l++) {
      const int B = l < p.loc_loops ? p.B_loc : p.B_est, d = B / rows;
      uint32_t ai = (uint32_t)hp[(size_t)signal * stride + L + l], inv = ai;
      for (int it = 0; it < 4; it++) inv *= 2u - ai * inv;
      tmp.assign(dst + off, dst + off + B);
      for (int b = 0; b < B; b++)
        dst[off + b] = tmp[(size_t)(b % d) * rows + (((uint32_t)(b / d) * inv) & (uint32_t)(rows - 1))];
      off += B;
    }
  }
  return m;
}

int sfft_plan_filter(const sfft_plan *plan, int which, complex_t *time, int time_cap,
                     complex_t *resp, int resp_cap, int *half)
{
  Plan *P = get(plan);
  if (!P || which < 0 || which > 1) return -1;
  const std::vector<cd> &t = P->f_time[which], &r = P->f_resp[which];
  if (half) *half = P->resp_half[which];
  if (time) {
    if (time_cap < (int)t.size()) {
      set_error("time buffer too small");
      return -1;
    }
    for (size_t i = 0; i < t.size(); i++) time[i] = t[i];
  }
  if (resp) {
    if (resp_cap < (int)r.size()) {
      set_error("response buffer too small");
      return -1;
    }
    for (size_t i = 0; i < r.size(); i++) resp[i] = r[i];
  }
  return (int)t.size();
}

int sfft_design_filter(double lobefrac, double tolerance, int n, int b,
                       complex_t *time, int time_cap, complex_t *freq)
{
  const int w = dolph_chebyshev_width(lobefrac, tolerance);
  if (!time || time_cap < w || n <= 0) {
    set_error("time buffer too small");
    return -1;
  }
  HostFilter f = design_filter(lobefrac, tolerance, n, b);
  for (int i = 0; i < w; i++) time[i] = f.time[i];
  if (freq)
    for (int i = 0; i < n; i++) freq[i] = f.freq[i];
  return w;
}

}  // extern "C"
