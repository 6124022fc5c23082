// Serving loop: continuous batching of single-signal requests over the plan
// machinery of host_plan.cpp (see include/sfft_server.h for the contract).
//
// Worker threads each own a HIP stream (non-blocking, default priority) and
// one Workspace shared by that worker's plan clones; the device tables of a
// plan are built once per (n, k, version, precision) and shared by every
// worker's clone.  All shared state is guarded by one mutex; plan building
// and batch execution run with it released.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/sfft_server.h"
#include "internal.hpp"

namespace sfa {
const char *last_error();
}

using namespace sfa;

namespace {

using Clock = std::chrono::steady_clock;

struct Key {
  int n, k, version, precision;
  bool operator<(const Key &o) const
  {
    return std::tie(n, k, version, precision) < std::tie(o.n, o.k, o.version, o.precision);
  }
  bool operator==(const Key &o) const
  {
    return n == o.n && k == o.k && version == o.version && precision == o.precision;
  }
};

struct Req {
  int64_t id = 0;
  Key key{};
  sfft_request r{};
  Rng rng;
  Clock::time_point t_submit, t_done;
  int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
  int64_t count = 0;
  std::string error;
};

struct Resident {
  Plan *master = nullptr;
  bool building = false;
  int users = 0;  // batches running on it
  int64_t table_bytes = 0;
  uint64_t last_use = 0;
  std::string error;  // non-empty: the build failed
};

struct Worker {
  std::thread th;
  hipStream_t stream = nullptr;
  Workspace ws;                  // shared by the v1/v2 clones
  std::map<Key, Plan *> clones;  // v3 clones hold their own workspace
  int64_t reserved = 0;          // workspace bytes the running batch may grow to
  bool busy = false;
  int64_t held() const
  {
    int64_t b = (int64_t)ws.device_bytes();
    for (const auto &kv : clones) b += (int64_t)kv.second->own_ws.device_bytes();
    return b;
  }
};

}  // namespace

struct sfft_server {
  int device = 0;
  int64_t budget = 0;
  int max_batch = 1;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Req *> queue;
  std::unordered_map<int64_t, Req *> reqs;
  int64_t next_id = 0;
  std::map<Key, Resident> plans;
  std::vector<Worker *> workers;
  bool stop = false;
  uint64_t tick = 0;
  int running = 0;
  sfft_server_stats st{};

  int64_t tables_bytes() const
  {
    int64_t b = 0;
    for (const auto &kv : plans) b += kv.second.table_bytes;
    return b;
  }
  // HBM held or about to be held: tables plus, per worker, the larger of its
  // workspace and what its running batch reserved.
  int64_t in_use() const
  {
    int64_t b = tables_bytes();
    for (const Worker *w : workers) b += std::max<int64_t>(w->held(), w->reserved);
    return b;
  }
  void note_use()
  {
    st.hbm_in_use = in_use();
    st.hbm_peak = std::max(st.hbm_peak, st.hbm_in_use);
  }
};

namespace {

Key key_of(const sfft_request &r) { return Key{r.n, r.k, (int)r.version, (int)r.precision}; }

// Feasibility and capacity without building a plan.
int64_t capacity_of(int n, int k, sfft_version v, std::string &why)
{
  if (v != SFFT_VERSION_1 && v != SFFT_VERSION_2 && v != SFFT_VERSION_3) {
    why = "unknown sfft_version";
    return -1;
  }
  if (n <= 0 || (n & (n - 1)) != 0) {
    why = "n must be a power of two";
    return -1;
  }
  if (v == SFFT_VERSION_3) {
    const V3Params q = derive_v3(n, k);
    why = check_v3(q);
    return why.empty() ? v3_map_capacity(q) : -1;
  }
  const V12Params p = derive_v12(n, k, v == SFFT_VERSION_2);
  why = check_v12(p);
  return why.empty() ? v12_output_capacity(p) : -1;
}

// Called on a thread whose current device is the server's.
Plan *build_master(const Key &K)
{
  Plan *P = make_plan(K.n, K.k, (sfft_version)K.version);
  if (P) P->precision = (sfft_precision)K.precision;
  return P;
}

// Frees the plan tables nobody is using, least recently used first, until
// `need` more bytes fit.  Called with the lock held.
void evict_for(sfft_server *S, int64_t need)
{
  while (S->in_use() + need > S->budget) {
    auto victim = S->plans.end();
    for (auto it = S->plans.begin(); it != S->plans.end(); ++it) {
      const Resident &r = it->second;
      if (!r.master || r.building || r.users > 0) continue;
      if (victim == S->plans.end() || r.last_use < victim->second.last_use) victim = it;
    }
    if (victim == S->plans.end()) return;
    for (Worker *w : S->workers) {
      auto c = w->clones.find(victim->first);
      if (c != w->clones.end()) {
        free_plan(c->second);  // a clone frees no tables and no stream
        w->clones.erase(c);
      }
    }
    free_plan(victim->second.master);
    S->plans.erase(victim);
    S->st.plans_evicted++;
  }
}

void finish(sfft_server *S, Req *q, bool ok, int64_t count, const std::string &err)
{
  q->t_done = Clock::now();
  q->state = ok ? 2 : 3;
  q->count = count;
  q->error = err;
  if (ok) S->st.completed++;
  else S->st.failed++;
}

void worker_main(sfft_server *S, Worker *W)
{
  (void)hipSetDevice(S->device);
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    // the oldest queued request whose plan is not being built elsewhere
    Req *head = nullptr;
    for (Req *q : S->queue) {
      auto it = S->plans.find(q->key);
      if (it != S->plans.end() && it->second.building) continue;
      head = q;
      break;
    }
    if (!head) {
      if (S->stop && S->queue.empty()) break;
      S->cv.wait(lk);
      continue;
    }
    const Key K = head->key;
    Resident &R = S->plans[K];
    if (!R.master) {
      // build the tables with the lock released; other workers skip K
      R.building = true;
      lk.unlock();
      const auto t0 = Clock::now();
      Plan *P = build_master(K);
      const double ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
      const std::string err = P ? "" : std::string(last_error());
      lk.lock();
      Resident &R2 = S->plans[K];
      R2.building = false;
      S->st.plan_build_ms += ms;
      if (!P) {
        // fail every queued request of this key
        for (auto it = S->queue.begin(); it != S->queue.end();) {
          if ((*it)->key == K) {
            finish(S, *it, false, 0, "plan: " + err);
            it = S->queue.erase(it);
          } else {
            ++it;
          }
        }
        S->plans.erase(K);
      } else {
        R2.master = P;
        R2.table_bytes = P->plan_bytes();
        R2.last_use = ++S->tick;
        S->st.plans_created++;
        R2.users++;  // never evict the plan just built for its own requests
        evict_for(S, 0);
        R2.users--;
        S->note_use();
      }
      S->cv.notify_all();
      continue;
    }
    // The batch, in arrival order: the queued requests of K, or, when K is a
    // v3 key, the queued v3 requests of every key with resident tables (they
    // run as one multi-plan launch).
    const bool multi = K.version == (int)SFFT_VERSION_3;
    std::vector<Req *> batch;
    for (Req *q : S->queue) {
      if ((int)batch.size() >= S->max_batch) break;
      if (q->key == K) {
        batch.push_back(q);
      } else if (multi && q->key.version == (int)SFFT_VERSION_3) {
        auto it = S->plans.find(q->key);
        if (it != S->plans.end() && it->second.master && !it->second.building)
          batch.push_back(q);
      }
    }
    // pin every key of the batch (eviction skips pinned plans)
    std::vector<Key> keys;
    for (Req *q : batch)
      if (std::find(keys.begin(), keys.end(), q->key) == keys.end()) {
        keys.push_back(q->key);
        S->plans[q->key].users++;
      }
    // admission by HBM footprint: the batch's workspace is counted as new
    // memory on top of what is held; cut the batch, then evict unused plans,
    // then wait for running batches
    auto need = [&](size_t cnt) {
      int64_t b = 0;
      for (size_t i = 0; i < cnt; i++) b += S->plans[batch[i]->key].master->bytes_per_signal();
      return b;
    };
    auto fits = [&](size_t cnt) { return S->in_use() + need(cnt) <= S->budget; };
    // this worker's idle workspaces count as held: give them back first
    if (!fits(batch.size())) {
      W->ws.release();
      for (auto &kv : W->clones) kv.second->own_ws.release();
    }
    while (batch.size() > 1 && !fits(batch.size())) batch.pop_back();
    if (!fits(batch.size())) evict_for(S, need(batch.size()));
    if (!fits(batch.size())) {
      for (const Key &k : keys) S->plans[k].users--;
      if (S->running > 0) {  // running batches will give memory back
        S->cv.wait(lk);
        continue;
      }
      S->queue.erase(std::find(S->queue.begin(), S->queue.end(), batch[0]));
      finish(S, batch[0], false, 0, "request does not fit in the HBM budget");
      S->note_use();
      S->cv.notify_all();
      continue;
    }
    // unpin the keys the cut removed
    for (auto it = keys.begin(); it != keys.end();) {
      bool used = false;
      for (Req *q : batch) used |= q->key == *it;
      if (!used) {
        S->plans[*it].users--;
        it = keys.erase(it);
      } else {
        ++it;
      }
    }
    for (Req *q : batch) {
      S->queue.erase(std::find(S->queue.begin(), S->queue.end(), q));
      q->state = 1;
    }
    W->reserved = W->held() + need(batch.size());
    W->busy = true;
    S->running++;
    S->st.batches++;
    S->st.batched_requests += (int64_t)batch.size();
    S->st.max_concurrent_batches = std::max<int64_t>(S->st.max_concurrent_batches, S->running);
    // one clone per key: v1/v2 clones share the worker's workspace, v3
    // clones hold their own (a multi-plan launch keeps every key's state)
    std::vector<Plan *> clones;
    for (const Key &k : keys) {
      Resident &Rk = S->plans[k];
      Rk.last_use = ++S->tick;
      auto ci = W->clones.find(k);
      if (ci == W->clones.end())
        ci = W->clones.emplace(k, clone_plan(Rk.master, multi ? nullptr : &W->ws, W->stream))
                 .first;
      clones.push_back(ci->second);
    }
    S->note_use();
    lk.unlock();

    // per key: its requests in arrival order
    const size_t G = keys.size();
    std::vector<std::vector<Req *>> grp(G);
    for (Req *q : batch)
      grp[std::find(keys.begin(), keys.end(), q->key) - keys.begin()].push_back(q);
    std::vector<std::vector<const double2 *>> in(G);
    std::vector<std::vector<int32_t *>> fr(G);
    std::vector<std::vector<double2 *>> va(G);
    std::vector<std::vector<Rng *>> rng(G);
    std::vector<std::vector<int64_t>> counts(G);
    std::vector<BatchIO> ios(G);
    for (size_t g = 0; g < G; g++) {
      bool any_out = false;
      for (Req *q : grp[g]) {
        in[g].push_back(reinterpret_cast<const double2 *>(q->r.in));
        fr[g].push_back(q->r.freq);
        va[g].push_back(reinterpret_cast<double2 *>(q->r.val));
        rng[g].push_back(&q->rng);
        any_out |= q->r.freq != nullptr || q->r.val != nullptr;
      }
      counts[g].assign(grp[g].size(), 0);
      BatchIO &io = ios[g];
      io.S = (int)grp[g].size();
      io.rng = rng[g].data();
      io.in = in[g].data();
      io.freq = any_out ? fr[g].data() : nullptr;
      io.val = any_out ? va[g].data() : nullptr;
      io.cap = clones[g]->output_capacity();
      io.counts = counts[g].data();
    }
    for (Plan *c : clones) c->last_status.clear();
    const int rc = multi ? exec_multi_v3(clones.data(), ios.data(), (int)G, W->stream, &W->ws)
                         : exec_batch(clones[0], ios[0]);
    const std::string err = rc == 0 ? "" : std::string(last_error());

    lk.lock();
    for (size_t g = 0; g < G; g++) {
      // rc 1: some v3 map overflowed; the other requests' results are whole
      const std::vector<int> &stat = clones[g]->last_status;
      for (size_t i = 0; i < grp[g].size(); i++) {
        const bool ok = rc == 0 || (rc == 1 && i < stat.size() && !(stat[i] & 2));
        finish(S, grp[g][i], ok, counts[g][i], ok ? std::string() : err);
      }
      S->plans[keys[g]].users--;
    }
    W->reserved = 0;
    W->busy = false;
    S->running--;
    S->note_use();
    S->cv.notify_all();
  }
}

}  // namespace

extern "C" {

int64_t sfft_output_capacity(int n, int k, sfft_version version)
{
  std::string why;
  const int64_t c = capacity_of(n, k, version, why);
  set_error(why);
  return c;
}

sfft_server *sfft_server_create(int64_t hbm_budget, int workers, int max_batch)
{
  set_error("");
  if (workers < 1 || max_batch < 1) {
    set_error("workers and max_batch must be >= 1");
    return nullptr;
  }
  sfft_server *S = new sfft_server;
  (void)hipGetDevice(&S->device);
  if (hbm_budget <= 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      set_error("hipMemGetInfo failed");
      delete S;
      return nullptr;
    }
    hbm_budget = (int64_t)(0.8 * (double)fr);
  }
  S->budget = hbm_budget;
  S->max_batch = max_batch;
  S->st.hbm_budget = hbm_budget;
  for (int i = 0; i < workers; i++) {
    Worker *W = new Worker;
    if (hipStreamCreateWithFlags(&W->stream, hipStreamNonBlocking) != hipSuccess) {
      set_error("hipStreamCreate failed");
      delete W;
      sfft_server_destroy(S);
      return nullptr;
    }
    S->workers.push_back(W);
  }
  for (Worker *W : S->workers) W->th = std::thread(worker_main, S, W);
  return S;
}

int sfft_server_prepare(sfft_server *S, int n, int k, sfft_version version,
                        sfft_precision precision)
{
  set_error("");
  if (!S) {
    set_error("null server");
    return -1;
  }
  std::string why;
  if (capacity_of(n, k, version, why) < 0) {
    set_error(why);
    return -1;
  }
  const Key K{n, k, (int)version, (int)precision};
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    Resident &R = S->plans[K];
    if (R.master) return 0;
    if (!R.building) break;
    S->cv.wait(lk);
  }
  S->plans[K].building = true;
  lk.unlock();
  const auto t0 = Clock::now();
  Plan *P = build_master(K);
  const double ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
  const std::string err = P ? "" : std::string(last_error());
  lk.lock();
  Resident &R = S->plans[K];
  R.building = false;
  S->st.plan_build_ms += ms;
  if (P) {
    R.master = P;
    R.table_bytes = P->plan_bytes();
    R.last_use = ++S->tick;
    S->st.plans_created++;
    R.users++;  // never evict the plan just built
    evict_for(S, 0);
    R.users--;
    S->note_use();
  } else {
    S->plans.erase(K);
  }
  S->cv.notify_all();
  lk.unlock();
  if (!P) {
    set_error(err);
    return -1;
  }
  return 0;
}

int64_t sfft_server_submit(sfft_server *S, const sfft_request *r)
{
  set_error("");
  if (!S || !r) {
    set_error("null server or request");
    return -1;
  }
  std::string why;
  const int64_t cap = capacity_of(r->n, r->k, r->version, why);
  if (cap < 0 || !r->in ||
      (r->precision != SFFT_PRECISION_FP64 && r->precision != SFFT_PRECISION_FP32_REFINE) ||
      ((r->freq || r->val) && r->capacity < cap)) {
    if (cap >= 0)
      why = !r->in ? "null input" : (r->capacity < cap ? "capacity below the plan's" : "unknown precision");
    std::lock_guard<std::mutex> lk(S->mu);
    S->st.rejected++;
    set_error(why);
    return -1;
  }
  // the kernels dereference these: host memory would fault the GPU
  {
    PtrKinds kinds;
    const char *bad = nullptr;
    if (!kinds.covers(r->in, (size_t)r->n * 16)) bad = "input is not device memory holding n values";
    else if (r->freq && !kinds.covers(r->freq, (size_t)r->capacity * 4))
      bad = "freq is not device memory holding capacity entries";
    else if (r->val && !kinds.covers(r->val, (size_t)r->capacity * 16))
      bad = "val is not device memory holding capacity entries";
    if (bad) {
      std::lock_guard<std::mutex> lk(S->mu);
      S->st.rejected++;
      set_error(bad);
      return -1;
    }
  }
  Req *q = new Req;
  q->key = key_of(*r);
  q->r = *r;
  q->rng.seed(r->random_seed, r->drand48_seed);
  q->t_submit = Clock::now();
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->stop) {
    delete q;
    set_error("server is shutting down");
    return -1;
  }
  q->id = S->next_id++;
  S->reqs[q->id] = q;
  S->queue.push_back(q);
  S->st.submitted++;
  S->cv.notify_all();
  return q->id;
}

int sfft_server_poll(sfft_server *S, int64_t ticket)
{
  if (!S) return -1;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(ticket);
  if (it == S->reqs.end()) return -1;
  const int s = it->second->state;
  return s == 2 ? 1 : (s == 3 ? -1 : 0);
}

int sfft_server_wait(sfft_server *S, int64_t ticket, int64_t *count, double *latency_ms)
{
  set_error("");
  if (!S) {
    set_error("null server");
    return -1;
  }
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->reqs.find(ticket);
  if (it == S->reqs.end()) {
    set_error("unknown ticket");
    return -1;
  }
  Req *q = it->second;
  S->cv.wait(lk, [&] { return q->state >= 2; });
  S->reqs.erase(it);
  lk.unlock();
  if (count) *count = q->count;
  if (latency_ms)
    *latency_ms = std::chrono::duration<double, std::milli>(q->t_done - q->t_submit).count();
  const bool ok = q->state == 2;
  if (!ok) set_error(q->error);
  delete q;
  return ok ? 0 : -1;
}

void sfft_server_get_stats(sfft_server *S, sfft_server_stats *out)
{
  if (!S || !out) return;
  std::lock_guard<std::mutex> lk(S->mu);
  S->note_use();
  S->st.plans_resident = (int64_t)S->plans.size();
  *out = S->st;
}

void sfft_server_destroy(sfft_server *S)
{
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
    S->cv.notify_all();
  }
  for (Worker *W : S->workers)
    if (W->th.joinable()) W->th.join();
  for (Worker *W : S->workers) {
    if (W->stream) (void)hipStreamSynchronize(W->stream);
    for (auto &kv : W->clones) free_plan(kv.second);
    W->ws.release();
    if (W->stream) (void)hipStreamDestroy(W->stream);
    delete W;
  }
  for (auto &kv : S->plans)
    if (kv.second.master) free_plan(kv.second.master);
  for (auto &kv : S->reqs) delete kv.second;
  delete S;
}

}  // extern "C"
