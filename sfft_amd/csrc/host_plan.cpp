// Plan construction (device-resident tables), RNG draws and batch execution
// of sFFT v1/v2 on the GPU.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>

#include "internal.hpp"

namespace sfa {

// ---------------------------------------------------------------- errors
static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
const char *last_error() { return g_last_error.c_str(); }

#define HIP_OK(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) {                                                 \
      set_error(std::string(#call) + ": " + hipGetErrorString(e_));         \
      return -1;                                                            \
    }                                                                       \
  } while (0)

// ---------------------------------------------------------------- LDS limit
void raise_lds_limit(const void *kernel)
{
  // a fixed table: nothing is allocated here, whoever calls
  static std::mutex mu;
  static std::pair<int, const void *> done[512];
  static int ndone = 0;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return;
  std::lock_guard<std::mutex> lk(mu);
  for (int i = 0; i < ndone; i++)
    if (done[i].first == dev && done[i].second == kernel) return;
  // gfx950: 160 KiB of LDS per workgroup, static and dynamic together
  size_t stat = 0;
  hipFuncAttributes fa;
  if (hipFuncGetAttributes(&fa, kernel) == hipSuccess) stat = fa.sharedSizeBytes;
  else (void)hipGetLastError();
  const int lim = (int)(160 * 1024 - std::min<size_t>(stat, 160 * 1024));
  if (hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, lim) != hipSuccess)
    (void)hipGetLastError();  // the launch reports it
  else if (ndone < 512)
    done[ndone++] = {dev, kernel};
}

// ---------------------------------------------------------------- RNG
int gcd_int(int a, int b)
{
  // gcd(0, n) = n, as src/utils.cc:41-46 (which forces the first draw).
  while (a % b != 0) {
    int t = a % b;
    a = b;
    b = t;
  }
  return b;
}

int mod_inverse_int(int a, int n)
{
  // Extended Euclid on (n, a); result in [0, n).
  int64_t r0 = n, r1 = a, s0 = 0, s1 = 1;
  while (r1 != 0) {
    int64_t q = r0 / r1, t = r0 - q * r1;
    r0 = r1;
    r1 = t;
    t = s0 - q * s1;
    s0 = s1;
    s1 = t;
  }
  int64_t v = s0 % n;
  if (v < 0) v += n;
  return (int)v;
}

// The inverse of odd a modulo n = 2^p (the only n a plan accepts): Newton's
// iteration x <- x (2 - a x) doubles the correct low bits, starting from
// x = a (a*a = 1 mod 8).  The same unique value as mod_inverse_int.
static inline int mod_inverse_pow2(int a, int n)
{
  uint32_t x = (uint32_t)a;
  for (int i = 0; i < 4; i++) x *= 2u - (uint32_t)a * x;  // 3 -> 48 bits
  return (int)(x & (uint32_t)(n - 1));
}

void Rng::seed(unsigned rs, long ds)
{
  std::memset(&rd, 0, sizeof(rd));
  std::memset(state, 0, sizeof(state));
  initstate_r(rs, state, sizeof(state), &rd);  // TYPE_3, as srandom()
  srand48_r(ds, &dd);
  seeded = true;
}

long Rng::next_random()
{
  if (!seeded) return ::random();
  int32_t r = 0;
  random_r(&rd, &r);
  return r;
}

double Rng::next_drand48()
{
  if (!seeded) return ::drand48();
  double r = 0;
  drand48_r(&dd, &r);
  return r;
}

// ---------------------------------------------------------------- buffers
bool DevBuf::ensure(size_t nbytes)
{
  if (nbytes <= bytes && p) return true;
  if (arena) {
    // carved once per exec from the fixed region, 256-byte aligned
    const size_t at = (arena->used + 255) & ~size_t(255);
    if (p || at + nbytes > arena->size) {
      set_error("allocation-free exec: the plan needs more than the " +
                std::to_string(arena->size >> 20) + " MiB scratch region");
      return false;
    }
    p = arena->base + at;
    bytes = nbytes;
    arena->used = at + nbytes;
    return true;
  }
  release();
  if (nbytes == 0) nbytes = 256;
  if (hipMalloc(&p, nbytes) != hipSuccess) {
    p = nullptr;
    bytes = 0;
    set_error("hipMalloc of " + std::to_string(nbytes) + " bytes failed");
    return false;
  }
  bytes = nbytes;
  return true;
}

void DevBuf::release()
{
  if (p && !arena) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

template <typename T>
static bool upload(DevBuf &b, const std::vector<T> &v)
{
  if (!b.ensure(v.size() * sizeof(T))) return false;
  if (hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) !=
      hipSuccess) {
    set_error("table upload failed");
    return false;
  }
  return true;
}

static std::vector<double2> to_d2(const std::vector<cd> &v, size_t cnt)
{
  std::vector<double2> o(cnt);
  for (size_t i = 0; i < cnt; i++) o[i] = make_double2(v[i].real(), v[i].imag());
  return o;
}

static std::vector<float2> to_f2(const std::vector<cd> &v, size_t cnt)
{
  std::vector<float2> o(cnt);
  for (size_t i = 0; i < cnt; i++)
    o[i] = make_float2((float)v[i].real(), (float)v[i].imag());
  return o;
}

// Taps by bucket: t[b * R + k] = tap[b + k*B] for b + k*B < w (zero past
// the last tap), R = ceil(w / B).  The fold reads a bucket's taps from one
// contiguous run.
template <typename T2>
static std::vector<T2> taps_by_bucket(const std::vector<cd> &v, int B, int &R)
{
  const int w = (int)v.size();
  R = (w + B - 1) / B;
  std::vector<T2> o((size_t)B * R);
  for (auto &t : o) t.x = t.y = 0;
  for (int i = 0; i < w; i++) {
    T2 &t = o[(size_t)(i % B) * R + i / B];
    t.x = (decltype(t.x))v[i].real();
    t.y = (decltype(t.y))v[i].imag();
  }
  return o;
}

// Estimation window of one filter: win[d + m/2] = freq[(n - d) mod n] for
// d = -m/2..m/2 (m = n/B); estimate_values reads only these
// (src/computefourier-1.0-2.0.cc:371-385).
static std::vector<double2> est_window(const std::vector<cd> &freq, int n, int B)
{
  const int m = n / B;
  std::vector<double2> w(m + 1);
  for (int d = -(m / 2); d <= m / 2; d++) {
    const cd v = freq[(size_t)(((int64_t)n - d) % n)];
    w[d + m / 2] = make_double2(v.real(), v.imag());
  }
  return w;
}

// ---------------------------------------------------------------- plan
static int pow2_ceil(int64_t v)
{
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

int64_t v12_output_capacity(const V12Params &p)
{
  if (p.comb) return (int64_t)p.comb_loops * p.B_thresh * (p.n / p.W_comb);
  // v1: every hit is selected in >= thr of the location loops.
  const int64_t sel = (int64_t)p.loc_loops * p.B_thresh * (p.n / p.B_loc);
  return std::min<int64_t>(p.n, sel / std::max(1, p.thr_loops) + 1);
}

// v3 map of located frequencies: room for several passes' spurious finds.
int64_t v3_map_capacity(const V3Params &q)
{
  return pow2_ceil(8LL * q.k + q.W_Man + 2LL * q.B_g1 + 2LL * q.B_g2 + 64);
}

int64_t Plan::output_capacity() const
{
  return version == SFFT_VERSION_3 ? v3_Mcap : v12_output_capacity(v12);
}

int64_t Plan::plan_bytes() const
{
  return (int64_t)(taps_loc.bytes + taps_est.bytes + tapsT_loc32.bytes +
                   tapsT_est32.bytes + win_loc.bytes + win_est.bytes +
                   tw64.bytes + tw32.bytes + v3_time1.bytes + v3_time2.bytes +
                   v3_freq1.bytes + v3_freq2.bytes);
}

int64_t Plan::bytes_per_signal() const
{
  if (version == SFFT_VERSION_3) {
    // Mansour buckets 2W; above the LDS size the pre-pass copy and its
    // transform scratch (2 or, refined, kV3RefineOffsets + 1 offsets each);
    // refined below it, the fresh buckets at kV3RefineOffsets offsets
    const bool ref = precision == SFFT_PRECISION_FP32_REFINE;
    const bool big = v3.W_Man > kMaxLdsFft64;
    const int64_t off = ref ? kV3RefineOffsets : 0;
    const int64_t ms = 32LL + (big ? 32LL * (ref ? off + 1 : 2) : 16LL * off);
    return ms * v3.W_Man + (int64_t)v3_Mcap * 28 + (int64_t)v3_Rcap * 236 +
           (int64_t)v3_Fcap * 40 + 16LL * (v3.w_g2 + 2) + 32LL * (v3.B_g1 + v3.B_g2) + 768;
  }
  const V12Params &p = v12;
  const bool fp32 = precision == SFFT_PRECISION_FP32_REFINE;
  const bool need64 = !fp32 || !p.comb;
  const int64_t xs = p.x_samp_size();
  const int Bm = std::max(p.B_loc, p.B_est);
  int64_t b = (need64 ? xs * 16 : 0) + (fp32 ? xs * 8 : 0);
  if (need64 && Bm > kMaxLdsFft64) b += xs * 16;
  if (fp32 && Bm > 2 * kMaxLdsFft64) b += xs * 8;
  if (p.comb && p.W_comb > k::kMaxLdsComb) b += (int64_t)p.comb_loops * p.W_comb * 40;
  if (!p.comb && p.B_loc > k::kMaxLdsSelect) b += (int64_t)p.loc_loops * p.B_loc * 8;
  if (!p.comb) b += 4LL * k::vote_chunks(n, p.B_loc, p.B_thresh, p.loc_loops, p.thr_loops);
  return b + (int64_t)p.comb_loops * p.B_thresh * 4 + (int64_t)p.loc_loops * p.B_thresh * 4 +
         output_capacity() * (4 + 16 + 4) + (2 * p.loops() + p.comb_loops + 8) * 4;
}

// Filter response window at signed offsets -half..half (freq[d mod n]);
// v3 reads offsets within 3n/2B of a bucket centre (cf3 371-375, 611-613).
static std::vector<double2> centred_window(const std::vector<cd> &freq, int n, int half)
{
  std::vector<double2> w(2 * (size_t)half + 1);
  for (int d = -half; d <= half; d++) {
    const cd v = freq[(size_t)(((int64_t)d % n + n) % n)];
    w[d + half] = make_double2(v.real(), v.imag());
  }
  return w;
}

// Host copy of a filter for sfft_plan_filter: taps and the response at
// offsets -half..half.
static void keep_filter(Plan *P, int which, const HostFilter &f, int n, int half)
{
  P->f_time[which] = f.time;
  P->f_resp[which].resize(2 * (size_t)half + 1);
  for (int d = -half; d <= half; d++)
    P->f_resp[which][d + half] = f.freq[(size_t)(((int64_t)d % n + n) % n)];
  P->resp_half[which] = half;
}

static std::vector<cd> twiddles(int nmax)
{
  std::vector<cd> tw(nmax / 2);
  for (int j = 0; j < nmax / 2; j++) {
    // sincos() explicitly: whether a compiler fuses sin+cos into it varies,
    // and the two differ in the last bit for rare arguments.
    const double ang = 2.0 * M_PI * (double)j / (double)nmax;
    double s, c;
    ::sincos(ang, &s, &c);
    tw[j] = cd(c, -s);
  }
  return tw;
}

// v3 plan tables (src/sfft.cc:513-560): two filters with b = n/B.
static int make_v3_tables(Plan *P)
{
  const V3Params &q = P->v3;
  const int n = q.n;
  HostFilter f1, f2;
  if (!design_filter_gpu(0.5 / q.BB1, 1e-8, n, q.b_g1, f1) ||
      !design_filter_gpu(0.5 / q.BB2, 1e-8, n, q.b_g2, f2))
    return -1;
  P->v3_half1 = 3 * (n / q.B_g1) / 2;
  P->v3_half2 = 3 * (n / q.B_g2) / 2;
  keep_filter(P, 0, f1, n, P->v3_half1);
  keep_filter(P, 1, f2, n, P->v3_half2);
  P->tw_nmax = std::max(q.W_Man, std::max(q.B_g1, q.B_g2));
  P->v3_Mcap = (int)v3_map_capacity(q);
  P->v3_Hcap = 2 * P->v3_Mcap;
  P->v3_Fcap = std::max(q.W_Man, std::max(q.B_g1, q.B_g2));
  P->v3_Rcap = std::max(P->v3_Mcap, P->v3_Fcap);
  // the Mansour FFTs run in LDS up to kMaxLdsFft64 points, else before the
  // persistent kernel as batched radix-2 passes
  P->v3_lds = std::max<size_t>(q.W_Man <= kMaxLdsFft64 ? (size_t)q.W_Man * 16 : 0,
                               (size_t)(2 * q.B_g1 + 2 * q.B_g2) * 16);
  P->v3_lds = std::max<size_t>(P->v3_lds, std::min<size_t>((size_t)P->v3_Mcap * 8, 131072));
  // room past the Gauss bucket arrays for the Gauss-2 phase chain, where it
  // fits
  const size_t gauss = (size_t)(2 * q.B_g1 + 2 * q.B_g2) * 16;
  P->v3_lds = std::max<size_t>(P->v3_lds,
                               std::min<size_t>(gauss + (size_t)(q.w_g2 + 1) * 16, kV3MaxDynLds));
  if (P->v3_lds > kV3MaxDynLds) {
    set_error("v3: the four Gauss bucket arrays (2 B_g1 + 2 B_g2 complex doubles) exceed the "
              "LDS of a workgroup; k is too large for this plan");
    return -1;
  }
  const bool ok = upload(P->v3_time1, to_d2(f1.time, f1.time.size())) &&
                  upload(P->v3_time2, to_d2(f2.time, f2.time.size())) &&
                  upload(P->v3_freq1, centred_window(f1.freq, n, P->v3_half1)) &&
                  upload(P->v3_freq2, centred_window(f2.freq, n, P->v3_half2)) &&
                  upload(P->tw64, to_d2(twiddles(P->tw_nmax), P->tw_nmax / 2));
  return ok ? 0 : -1;
}

Plan *make_plan(int n, int k, sfft_version v)
{
  if (v != SFFT_VERSION_1 && v != SFFT_VERSION_2 && v != SFFT_VERSION_3) {
    set_error("unknown sfft_version");
    return nullptr;
  }
  if (n <= 0 || (n & (n - 1)) != 0) {
    set_error("n must be a power of two");
    return nullptr;
  }
  Plan *P = new Plan();
  P->version = v;
  P->n = n;
  P->k = k;
  if (v == SFFT_VERSION_3) {
    P->v3 = derive_v3(n, k);
    std::string why = check_v3(P->v3);
    if (!why.empty()) {
      set_error(why);
      delete P;
      return nullptr;
    }
    if (make_v3_tables(P) != 0 || hipStreamCreate(&P->stream) != hipSuccess) {
      if (P->stream == nullptr && sfa::last_error()[0] == 0) set_error("hipStreamCreate failed");
      free_plan(P);
      return nullptr;
    }
    P->own_stream = true;
    return P;
  }
  // tuning overrides for experiments: fold schedule and its resident blocks
  if (const char *e = std::getenv("SFFT_FOLD_MODE")) P->fold_mode = std::min(2, std::max(0, std::atoi(e)));
  if (const char *e = std::getenv("SFFT_FOLD_BLOCKS")) P->stream_blocks = std::atoi(e);
  P->v12 = derive_v12(n, k, v == SFFT_VERSION_2);
  std::string why = check_v12(P->v12);
  if (!why.empty()) {
    set_error(why);
    delete P;
    return nullptr;
  }
  const V12Params &p = P->v12;
  HostFilter fl, fe;
  if (!design_filter_gpu(0.5 / p.BB_loc, p.tol_loc, p.n, p.b_loc, fl) ||
      !design_filter_gpu(0.5 / p.BB_est, p.tol_est, p.n, p.b_est, fe)) {
    free_plan(P);
    return nullptr;
  }
  keep_filter(P, 0, fl, p.n, (p.n / p.B_loc) / 2);
  keep_filter(P, 1, fe, p.n, (p.n / p.B_est) / 2);
  int nmax = std::max(p.B_loc, p.B_est);
  if (p.comb) nmax = std::max(nmax, p.W_comb);
  const std::vector<cd> tw = twiddles(nmax);
  P->tw_nmax = nmax;
  bool ok = upload(P->taps_loc, to_d2(fl.time, fl.time.size())) &&
            upload(P->taps_est, to_d2(fe.time, fe.time.size())) &&
            upload(P->tapsT_loc32, taps_by_bucket<float2>(fl.time, p.B_loc, P->tR_loc)) &&
            upload(P->tapsT_est32, taps_by_bucket<float2>(fe.time, p.B_est, P->tR_est)) &&
            upload(P->win_loc, est_window(fl.freq, p.n, p.B_loc)) &&
            upload(P->win_est, est_window(fe.freq, p.n, p.B_est)) &&
            upload(P->tw64, to_d2(tw, tw.size())) &&
            upload(P->tw32, to_f2(tw, tw.size()));
  if (!ok || hipStreamCreate(&P->stream) != hipSuccess) {
    if (ok) set_error("hipStreamCreate failed");
    free_plan(P);
    return nullptr;
  }
  P->own_stream = true;
  // the allocation-free exec's scratch address on this device, looked up
  // here (plan construction is its set-up) rather than in that exec
  if (k::noalloc_setup() != 0) {
    free_plan(P);
    return nullptr;
  }
  return P;
}

size_t Workspace::device_bytes() const
{
  const DevBuf *bufs[] = {&params, &xsamp, &xsamp32, &approved, &numcomb, &J, &hits,
                          &ptrs, &out_freq, &out_val, &count, &stage, &v3, &v3_status,
                          &comb, &mag, &tmp64, &tmp32, &multi, &votes};
  size_t b = 0;
  for (const DevBuf *d : bufs) b += d->bytes;
  return b;
}

void Workspace::use_arena(Arena *a)
{
  DevBuf *bufs[] = {&params, &xsamp, &xsamp32, &approved, &numcomb, &J, &hits,
                    &ptrs, &out_freq, &out_val, &count, &stage, &v3, &v3_status,
                    &comb, &mag, &tmp64, &tmp32, &multi, &votes};
  for (DevBuf *d : bufs) d->arena = a;
}

void Workspace::release()
{
  DevBuf *bufs[] = {&params, &xsamp, &xsamp32, &approved, &numcomb, &J, &hits,
                    &ptrs, &out_freq, &out_val, &count, &stage, &v3, &v3_status,
                    &comb, &mag, &tmp64, &tmp32, &multi, &votes};
  for (DevBuf *d : bufs) d->release();
  if (pinned && !pinned_view) (void)hipHostFree(pinned);
  pinned = nullptr;
  pinned_bytes = 0;
}

Plan *clone_plan(const Plan *src, Workspace *ws, hipStream_t stream)
{
  Plan *P = new Plan;
  P->version = src->version;
  P->n = src->n;
  P->k = src->k;
  P->v12 = src->v12;
  P->v3 = src->v3;
  P->precision = src->precision;
  P->fold_mode = src->fold_mode;
  P->stream_blocks = src->stream_blocks;
  P->ref_conj = src->ref_conj;
  P->stream = stream;
  P->own_stream = false;
  P->owns_tables = false;
  P->taps_loc = src->taps_loc;
  P->taps_est = src->taps_est;
  P->tapsT_loc32 = src->tapsT_loc32;
  P->tapsT_est32 = src->tapsT_est32;
  P->tR_loc = src->tR_loc;
  P->tR_est = src->tR_est;
  P->seg_rows = src->seg_rows;
  P->win_loc = src->win_loc;
  P->win_est = src->win_est;
  P->tw64 = src->tw64;
  P->tw32 = src->tw32;
  P->tw_nmax = src->tw_nmax;
  P->v3_time1 = src->v3_time1;
  P->v3_time2 = src->v3_time2;
  P->v3_freq1 = src->v3_freq1;
  P->v3_freq2 = src->v3_freq2;
  P->v3_half1 = src->v3_half1;
  P->v3_half2 = src->v3_half2;
  P->v3_Mcap = src->v3_Mcap;
  P->v3_Hcap = src->v3_Hcap;
  P->v3_Fcap = src->v3_Fcap;
  P->v3_Rcap = src->v3_Rcap;
  P->v3_lds = src->v3_lds;
  P->ws = ws ? ws : &P->own_ws;
  return P;
}

void free_plan(Plan *P)
{
  if (!P) return;
  // a clone's stream belongs to its server worker, which waits for it
  if (P->stream && (P->own_stream || P->owns_tables)) (void)hipStreamSynchronize(P->stream);
  if (P->owns_tables) {
    DevBuf *bufs[] = {&P->taps_loc, &P->taps_est, &P->tapsT_loc32, &P->tapsT_est32,
                      &P->win_loc, &P->win_est, &P->tw64, &P->tw32,
                      &P->v3_time1, &P->v3_time2, &P->v3_freq1, &P->v3_freq2};
    for (DevBuf *b : bufs) b->release();
  }
  for (const auto &g : P->graphs) (void)hipGraphExecDestroy(g.exec);
  P->own_ws.release();
  if (P->ev_made) {
    for (hipEvent_t e : P->ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : P->ev_comb) (void)hipEventDestroy(e);
  }
  if (P->ev_fork) (void)hipEventDestroy(P->ev_fork);
  if (P->ev_join) (void)hipEventDestroy(P->ev_join);
  for (hipEvent_t e : P->ev_part)
    if (e) (void)hipEventDestroy(e);
  if (P->ev_est) (void)hipEventDestroy(P->ev_est);
  if (P->stream2) {
    (void)hipStreamSynchronize(P->stream2);
    (void)hipStreamDestroy(P->stream2);
  }
  if (P->stream3) {
    (void)hipStreamSynchronize(P->stream3);
    (void)hipStreamDestroy(P->stream3);
  }
  if (P->own_stream && P->stream) (void)hipStreamDestroy(P->stream);
  delete P;
}

// ---------------------------------------------------------------- exec
static int ensure_pinned(Plan *P, size_t bytes)
{
  if (bytes <= P->ws->pinned_bytes) return 0;
  if (P->ws->pinned_view) {
    set_error("host staging of the allocation-free exec too small");
    return -1;
  }
  if (P->ws->pinned) (void)hipHostFree(P->ws->pinned);
  P->ws->pinned = nullptr;
  P->ws->pinned_bytes = 0;
  HIP_OK(hipHostMalloc(&P->ws->pinned, bytes, hipHostMallocDefault));
  P->ws->pinned_bytes = bytes;
  return 0;
}

// A v3 batch between its uploads and its results.
struct V3Pending {
  k::V3Launch a{};
  int64_t *h_counts = nullptr;
  int *h_status = nullptr;
};

// v3: draws (cf3 800-810) in signal order, the workspace layout and the
// uploads of one batch (one workgroup per signal) on stream st.
// Host and device staging of one v3 call: the groups' descriptors and
// block list, each group's draws and pointer tables (uploaded with one
// copy), then each group's counts and status flags (read back with one).
struct V3Stage {
  char *h = nullptr, *d = nullptr;  // pinned host / device, same offsets
  size_t up = 0, dn = 0;            // next free offsets of the two regions
};
static size_t al256(size_t b) { return (b + 255) & ~size_t(255); }
static size_t v3_up_bytes(int S) { return al256((size_t)S * 6 * 4) + 4 * al256((size_t)S * 8); }
static size_t v3_dn_bytes(int S) { return al256((size_t)S * 8) + al256((size_t)S * 4); }

static int v3_prepare(Plan *P, BatchIO &io, V3Pending &pd, V3Stage &sg)
{
  const int S = io.S;
  const V3Params &q = P->v3;
  const int n = q.n;
  const size_t param_bytes = al256((size_t)S * 6 * sizeof(int));
  const size_t tab_bytes = al256((size_t)S * sizeof(void *));
  const size_t cnt_bytes = al256((size_t)S * sizeof(int64_t));
  char *pin = sg.h + sg.up, *dup = sg.d + sg.up;
  sg.up += param_bytes + 4 * tab_bytes;
  char *pdn = sg.h + sg.dn, *ddn = sg.d + sg.dn;
  sg.dn += cnt_bytes + al256((size_t)S * 4);
  int *hp = reinterpret_cast<int *>(pin);
  const void **t_in = reinterpret_cast<const void **>(pin + param_bytes);
  const void **t_freq = reinterpret_cast<const void **>(pin + param_bytes + tab_bytes);
  const void **t_val = reinterpret_cast<const void **>(pin + param_bytes + 2 * tab_bytes);
  const void **t_dense = reinterpret_cast<const void **>(pin + param_bytes + 3 * tab_bytes);
  int64_t *h_counts = reinterpret_cast<int64_t *>(pdn);
  int *h_status = reinterpret_cast<int *>(pdn + cnt_bytes);
  for (int s = 0; s < S; s++) {
    int *d = hp + (size_t)s * 6;
    Rng &R = io.rng && io.rng[s] ? *io.rng[s] : P->rng;
    // gcd(a, n) == 1 (src/utils.cc:41-46) is "a odd" for n = 2^p; a = 0
    // forces the first draw as gcd(0, n) = n does
    int a = 0;
    while ((a & 1) == 0) a = (int)(R.next_random() % n);
    const int ai = mod_inverse_pow2(a, n);
    const int b = (int)(R.next_random() % n);
    d[0] = a;
    d[1] = ai;
    d[2] = b;
    d[3] = (int)(((int64_t)ai * b) % n);  // ((ai*b) % n + n) % n with int wrap
    d[4] = (int)(unsigned)std::floor(R.next_drand48() * n);
    d[5] = (int)(unsigned)std::floor(R.next_drand48() * n);
    t_in[s] = io.in[s];
    t_freq[s] = io.freq ? io.freq[s] : nullptr;
    t_val[s] = io.val ? io.val[s] : nullptr;
    t_dense[s] = io.dense ? io.dense[s] : nullptr;
  }
  // workspace: each array is S x (its per-signal count), stride = that count
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t Sz = (size_t)S;
  const size_t b_ms = al(Sz * 2 * q.W_Man * 16), b_keys = al(Sz * P->v3_Mcap * 4),
               b_vals = al(Sz * P->v3_Mcap * 16), b_hidx = al(Sz * P->v3_Hcap * 4),
               b_fi = al(Sz * P->v3_Fcap * 4), b_fv = al(Sz * P->v3_Fcap * 16),
               b_rem = al(Sz * P->v3_Rcap * 14 * 16), b_remi = al(Sz * P->v3_Rcap * 3 * 4);
  const int Pcap = q.w_g2 + 2;
  const size_t b_ph = al(Sz * Pcap * 16);
  const bool big_ms = q.W_Man > kMaxLdsFft64;
  // the refine form: fresh Mansour buckets at kV3RefineOffsets offsets (in
  // the pre-pass arrays when W is above the LDS size)
  const bool refine = P->precision == SFFT_PRECISION_FP32_REFINE;
  const int ms_loops = refine && big_ms ? kV3RefineOffsets + 1 : 2;
  const size_t b_pre = big_ms ? al(Sz * ms_loops * q.W_Man * 16) : 0;
  const size_t b_ref = refine && !big_ms ? al(Sz * kV3RefineOffsets * q.W_Man * 16) : 0;
  const size_t b_gpre = al(Sz * (2 * (size_t)q.B_g1 + 2 * (size_t)q.B_g2) * 16);
  const size_t total = b_ms + b_keys + b_vals + b_hidx + 2 * b_fi + 2 * b_fv + b_rem + b_remi +
                       b_ph + 2 * b_pre + b_ref + b_gpre;
  if (!P->ws->v3.ensure(total)) return -1;
  pd.h_counts = h_counts;
  pd.h_status = h_status;
  char *dtab = dup + param_bytes;
  char *w = reinterpret_cast<char *>(P->ws->v3.p);
  k::V3Launch &a = pd.a;
  a = k::V3Launch{};
  a.n = n;
  a.S = S;
  a.k = q.k;
  a.W = q.W_Man;
  a.B1 = q.B_g1;
  a.B2 = q.B_g2;
  a.w1 = q.w_g1;
  a.w2 = q.w_g2;
  a.x = reinterpret_cast<const double2 *const *>(dtab);
  a.params = reinterpret_cast<const int *>(dup);
  a.taps1 = P->v3_time1.as<const double2>();
  a.taps2 = P->v3_time2.as<const double2>();
  a.win1 = P->v3_freq1.as<const double2>();
  a.win2 = P->v3_freq2.as<const double2>();
  a.win1_half = P->v3_half1;
  a.win2_half = P->v3_half2;
  a.tw64 = P->tw64.as<const double2>();
  a.tw_nmax = P->tw_nmax;
  a.ms = reinterpret_cast<double2 *>(w);
  w += b_ms;
  a.ms_pre = big_ms ? reinterpret_cast<double2 *>(w) : nullptr;
  w += b_pre;
  a.ms_tmp = big_ms ? reinterpret_cast<double2 *>(w) : nullptr;
  w += b_pre;
  a.map_keys = reinterpret_cast<int *>(w);
  w += b_keys;
  a.map_vals = reinterpret_cast<double2 *>(w);
  w += b_vals;
  a.map_hidx = reinterpret_cast<int *>(w);
  w += b_hidx;
  a.cf = reinterpret_cast<int *>(w);
  w += b_fi;
  a.ef = reinterpret_cast<int *>(w);
  w += b_fi;
  a.cv = reinterpret_cast<double2 *>(w);
  w += b_fv;
  a.ev = reinterpret_cast<double2 *>(w);
  w += b_fv;
  a.rem = reinterpret_cast<double2 *>(w);
  w += b_rem;
  a.remi = reinterpret_cast<int *>(w);
  w += b_remi;
  a.phase = reinterpret_cast<double2 *>(w);
  w += b_ph;
  a.ms_loops = ms_loops;
  a.refine = refine ? 1 : 0;
  a.serial_all = P->debug_v3_serial ? 1 : 0;
  a.inline_pre = P->debug_v3_inline_pre ? 1 : 0;
  a.map_limit = P->debug_v3_map_limit > 0 ? std::min(P->debug_v3_map_limit, P->v3_Mcap)
                                          : P->v3_Mcap;
  a.ref_ms = refine && !big_ms ? reinterpret_cast<double2 *>(w) : nullptr;
  w += b_ref;
  a.gpre = reinterpret_cast<double2 *>(w);
  w += b_gpre;
  {
    // refine offset stride: odd, near 0.618 n/W, below n/W
    const int np = n / q.W_Man;
    int t = ((int)std::lround(np * 0.6180339887498949)) | 1;
    if (t >= np) t = np - 1 > 0 ? np - 1 : 1;
    unsigned inv = (unsigned)t;  // Newton: t inv = 1 mod 2^32 after 5 steps
    for (int it = 0; it < 5; it++) inv *= 2u - (unsigned)t * inv;
    a.ref_t = t;
    a.ref_tinv = (int)(inv & (unsigned)(np - 1));
  }
  a.Pcap = Pcap;
  a.Mcap = P->v3_Mcap;
  a.Hcap = P->v3_Hcap;
  a.Fcap = P->v3_Fcap;
  a.Rcap = P->v3_Rcap;
  a.max_passes = 3 * 4096;  // the oracle's cap of the round-robin loop
  a.count = reinterpret_cast<int64_t *>(ddn);
  a.status = reinterpret_cast<int *>(ddn + cnt_bytes);
  const bool want_sparse = io.freq != nullptr || io.val != nullptr;
  a.o_freq = want_sparse ? reinterpret_cast<int32_t *const *>(dtab + tab_bytes) : nullptr;
  a.o_val = want_sparse ? reinterpret_cast<double2 *const *>(dtab + 2 * tab_bytes) : nullptr;
  a.o_dense = io.dense ? reinterpret_cast<double2 *const *>(dtab + 3 * tab_bytes) : nullptr;
  a.o_cap = io.cap;
  a.lds_bytes = P->v3_lds;
  P->last_S = S;
  return 0;
}

// After the stream has drained.  Returns 1 when some signal's map
// overflowed (error set; every signal's count and result are written, the
// overflowed ones truncated, status bit 1).
static int v3_finish(Plan *P, BatchIO &io, const V3Pending &pd)
{
  const int S = io.S;
  P->last_status.insert(P->last_status.end(), pd.h_status, pd.h_status + S);
  if (io.counts)
    for (int s = 0; s < S; s++) io.counts[s] = pd.h_counts[s];
  for (int s = 0; s < S; s++) {
    if (pd.h_status[s] & 2) {
      set_error("v3 result map overflow (more located frequencies than the plan holds)");
      return 1;
    }
    if (pd.h_status[s] & 16) {
      set_error("v3: a transform gave up waiting for a bucket task (library defect)");
      return 1;
    }
  }
  return 0;
}

static int exec_batch_v3(Plan *P, BatchIO &io)
{
  Plan *plans[1] = {P};
  return exec_multi_v3(plans, &io, 1, P->stream, P->ws, P);
}

static int ensure_pinned_ws(Workspace *ws, size_t bytes)
{
  if (bytes <= ws->pinned_bytes) return 0;
  if (ws->pinned_view) {
    set_error("host staging of the allocation-free exec too small");
    return -1;
  }
  if (ws->pinned) (void)hipHostFree(ws->pinned);
  ws->pinned = nullptr;
  ws->pinned_bytes = 0;
  HIP_OK(hipHostMalloc(&ws->pinned, bytes, hipHostMallocDefault));
  ws->pinned_bytes = bytes;
  return 0;
}

int exec_multi_v3(Plan *const *plans, BatchIO *ios, int G, hipStream_t st, Workspace *ws,
                  Plan *timer)
{
  std::vector<V3Pending> pd(G);
  std::vector<k::V3Launch> groups(G);
  std::vector<int2> blocks;
  size_t up = 0, dn = 0;
  for (int g = 0; g < G; g++) {
    if (plans[g]->version != SFFT_VERSION_3) {
      set_error("exec_multi_v3: not a v3 plan");
      return -1;
    }
    for (int s = 0; s < ios[g].S; s++) blocks.push_back(make_int2(g, s));
    up += v3_up_bytes(ios[g].S);
    dn += v3_dn_bytes(ios[g].S);
  }
  if (blocks.empty()) return 0;
  const size_t desc_bytes = al256((size_t)G * sizeof(k::V3Launch));
  const size_t blk_bytes = al256(blocks.size() * sizeof(int2));
  const size_t clm_bytes = al256(blocks.size() * 4 * sizeof(int));  // zeroed claim words
  const size_t up_total = desc_bytes + blk_bytes + clm_bytes + up;
  if (!ws->multi.ensure(up_total + dn) || ensure_pinned_ws(ws, up_total + dn) != 0) return -1;
  V3Stage sg;
  sg.h = reinterpret_cast<char *>(ws->pinned);
  sg.d = reinterpret_cast<char *>(ws->multi.p);
  sg.up = desc_bytes + blk_bytes + clm_bytes;
  sg.dn = up_total;
  for (int g = 0; g < G; g++) {
    if (v3_prepare(plans[g], ios[g], pd[g], sg) != 0) return -1;
    groups[g] = pd[g].a;
  }
  std::memcpy(sg.h, groups.data(), (size_t)G * sizeof(k::V3Launch));
  std::memcpy(sg.h + desc_bytes, blocks.data(), blocks.size() * sizeof(int2));
  std::memset(sg.h + desc_bytes + blk_bytes, 0, clm_bytes);
  HIP_OK(hipMemcpyAsync(sg.d, sg.h, up_total, hipMemcpyHostToDevice, st));
  const bool timing = timer && timer->timing;
  if (timing && !timer->ev_made) {
    for (hipEvent_t &e : timer->ev) HIP_OK(hipEventCreate(&e));
    timer->ev_made = true;
  }
  if (timing) (void)hipEventRecord(timer->ev[0], st);
  if (k::v3_multi(groups.data(), G, reinterpret_cast<k::V3Launch *>(sg.d),
                  reinterpret_cast<int2 *>(sg.d + desc_bytes), (int)blocks.size(),
                  reinterpret_cast<int *>(sg.d + desc_bytes + blk_bytes), st) != 0)
    return -1;
  if (timing) (void)hipEventRecord(timer->ev[1], st);
  HIP_OK(hipMemcpyAsync(sg.h + up_total, sg.d + up_total, dn, hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  if (timing) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, timer->ev[0], timer->ev[1]) == hipSuccess) {
      timer->stage_ms[5] += ms;
      timer->stage_launches[5] += 1;
    }
  }
  int rc = 0;  // 1: some map overflowed (every result written)
  for (int g = 0; g < G; g++)
    if (v3_finish(plans[g], ios[g], pd[g]) != 0) rc = 1;
  return rc;
}

// Addresses a captured batch depends on (FNV-1a over the pointers).
static uint64_t workspace_signature(const Workspace &w)
{
  const void *ptrs[] = {w.params.p, w.xsamp.p, w.xsamp32.p, w.approved.p, w.numcomb.p, w.J.p,
                        w.hits.p, w.ptrs.p, w.count.p, w.comb.p, w.mag.p, w.tmp64.p,
                        w.tmp32.p, w.pinned};
  uint64_t h = 1469598103934665603ull;
  for (const void *q : ptrs) {
    h ^= (uint64_t)(uintptr_t)q;
    h *= 1099511628211ull;
  }
  return h;
}

// After the stream work of a v1/v2 batch: wait, stage times, counts.
static int v12_finish(Plan *P, BatchIO &io, const V12Pending &q)
{
  HIP_OK(hipStreamSynchronize(q.st));
  if (q.timing) {
    for (int i = 0; i < 5; i++) {
      // only stages that ran: an event never recorded fails the query, and
      // that error would be reported by the next launch check
      if (!q.launches[i]) continue;
      float ms = 0;
      const hipError_t e = i == 0 ? hipEventElapsedTime(&ms, P->ev_comb[0], P->ev_comb[1])
                                  : hipEventElapsedTime(&ms, P->ev[i], P->ev[i + 1]);
      if (e == hipSuccess) {
        P->stage_ms[i] += ms;
        P->stage_launches[i] += q.launches[i];
      }
    }
  }
  if (io.counts)
    for (int s = 0; s < q.S; s++) io.counts[s] = q.h_counts[s];
  return 0;
}

// Signals [base, base + cnt) of a batch's launch: every per-signal array
// advanced by base signals (the kernels index them by signal alone).
static k::V12Launch sub_launch(const k::V12Launch &a, int base, int cnt, int stride, int vchunks)
{
  k::V12Launch b = a;
  const int64_t s0 = base;
  b.S = cnt;
  b.x = a.x + s0;
  b.params = a.params + s0 * stride;
  if (a.xs64) b.xs64 = a.xs64 + s0 * a.xs;
  if (a.xs32) b.xs32 = a.xs32 + s0 * a.xs;
  if (a.tmp64) b.tmp64 = a.tmp64 + s0 * a.xs;
  if (a.tmp32) b.tmp32 = a.tmp32 + s0 * a.xs;
  if (a.approved) b.approved = a.approved + s0 * a.comb_loops * a.num;
  if (a.num_comb) b.num_comb = a.num_comb + s0;
  if (a.J) b.J = a.J + s0 * a.Lloc * a.num;
  if (a.hits) b.hits = a.hits + s0 * a.hits_cap;
  if (a.vote_cnt) b.vote_cnt = a.vote_cnt + s0 * vchunks;
  b.count = a.count + s0;
  if (a.o_freq) b.o_freq = a.o_freq + s0;
  if (a.o_val) b.o_val = a.o_val + s0;
  if (a.o_dense) b.o_dense = a.o_dense + s0;
  return b;
}

// A v1/v2 batch: draws, uploads and stream work; with pd, returns once the
// work is enqueued (exec_batch_wait finishes it), else waits and finishes.
static int exec_batch_v12(Plan *P, BatchIO &io, V12Pending *pd)
{
  const int S = io.S;
  const V12Params &p = P->v12;
  const int L = p.loops(), cl = p.comb ? p.comb_loops : 0;
  const int n = p.n, num = p.B_thresh;
  const int stride = 2 * L + std::max(cl, 1);
  const int64_t cap = P->output_capacity();
  const int64_t xs = p.x_samp_size();
  const bool fp32 = P->precision == SFFT_PRECISION_FP32_REFINE;

  // Pinned staging: params | 4 pointer tables | counts.
  const size_t param_bytes = ((size_t)S * stride * sizeof(int) + 255) & ~size_t(255);
  const size_t tab_bytes = ((size_t)S * sizeof(void *) + 255) & ~size_t(255);
  const size_t cnt_bytes = (size_t)S * sizeof(int64_t);
  if (ensure_pinned(P, param_bytes + 4 * tab_bytes + cnt_bytes) != 0) return -1;
  char *pin = reinterpret_cast<char *>(P->ws->pinned);
  int *hp = reinterpret_cast<int *>(pin);
  const void **t_in = reinterpret_cast<const void **>(pin + param_bytes);
  const void **t_freq = reinterpret_cast<const void **>(pin + param_bytes + tab_bytes);
  const void **t_val = reinterpret_cast<const void **>(pin + param_bytes + 2 * tab_bytes);
  const void **t_dense = reinterpret_cast<const void **>(pin + param_bytes + 3 * tab_bytes);
  int64_t *h_counts = reinterpret_cast<int64_t *>(pin + param_bytes + 4 * tab_bytes);

  const int vchunks = p.comb ? 0 : k::vote_chunks(n, p.B_loc, num, p.loc_loops, p.thr_loops);
  // v1 keeps its location loops in FP64 under fp32 (selection-critical)
  const bool loc64 = fp32 && !p.comb;
  const bool need64 = !fp32 || loc64;
  const bool want_sparse = io.freq != nullptr || io.val != nullptr;
  const bool want_dense = io.dense != nullptr;
  for (int s = 0; s < S; s++) {
    t_in[s] = io.in[s];
    t_freq[s] = io.freq ? io.freq[s] : nullptr;
    t_val[s] = io.val ? io.val[s] : nullptr;
    t_dense[s] = io.dense ? io.dense[s] : nullptr;
  }

  // params and the pointer tables share one device buffer (one upload)
  if (!P->ws->params.ensure(param_bytes + 4 * tab_bytes) ||
      (need64 && !P->ws->xsamp.ensure((size_t)S * xs * 16)) ||
      (fp32 && !P->ws->xsamp32.ensure((size_t)S * xs * 8)) ||
      !P->ws->approved.ensure((size_t)S * std::max(cl, 1) * num * 4) ||
      !P->ws->numcomb.ensure((size_t)S * 4) ||
      !P->ws->J.ensure((size_t)S * p.loc_loops * num * 4) ||
      !P->ws->hits.ensure((size_t)S * (p.comb ? 1 : cap) * 4) ||
      !P->ws->count.ensure((size_t)S * 8 + (p.comb ? 0 : (size_t)S * vchunks * 4)))
    return -1;
  // v2 comb-vote diagnostic (FP64 buckets in natural order)
  const bool dbg_votes = P->debug_votes && p.comb && !fp32;
  if (dbg_votes && !P->ws->votes.ensure((size_t)S * 2 * cap * 8)) return -1;
  // scratch for the sizes above the LDS limits
  const bool big_comb = p.comb && p.W_comb > k::kMaxLdsComb;
  const bool big_sel = (!p.comb || dbg_votes) && p.B_loc > k::kMaxLdsSelect;
  const bool big64 = need64 && std::max(p.B_loc, p.B_est) > kMaxLdsFft64;
  const bool big32 = fp32 && std::max(p.B_loc, p.B_est) > 2 * kMaxLdsFft64;
  const size_t comb_elems = big_comb ? (size_t)S * cl * p.W_comb : 0;
  const size_t mag_elems =
      std::max(comb_elems, big_sel ? (size_t)S * p.loc_loops * p.B_loc : (size_t)0);
  if ((big_comb && !P->ws->comb.ensure(comb_elems * 2 * sizeof(double2))) ||
      (mag_elems && !P->ws->mag.ensure(mag_elems * sizeof(double))) ||
      (big64 && !P->ws->tmp64.ensure((size_t)S * xs * 16)) ||
      (big32 && !P->ws->tmp32.ensure((size_t)S * xs * 8)))
    return -1;
  // RNG draws in signal order (src/computefourier-1.0-2.0.cc:465-474, 62),
  // made once every buffer is in place so that a call failing above
  // consumes none.
  const int sigma = p.comb ? n / p.W_comb : 1;
  for (int s = 0; s < S; s++) {
    int *row = hp + (size_t)s * stride;
    Rng &R = io.rng && io.rng[s] ? *io.rng[s] : P->rng;
    for (int l = 0; l < L; l++) {
      int a = 0;
      while ((a & 1) == 0) a = (int)(R.next_random() % n);  // gcd(a, n) == 1, n = 2^p
      row[l] = a;
      row[L + l] = mod_inverse_pow2(a, n);
    }
    for (int c = 0; c < cl; c++)
      row[2 * L + c] = (int)(unsigned)std::floor(R.next_drand48() * sigma);
  }
  hipStream_t st = P->stream;
  char *dtab = reinterpret_cast<char *>(P->ws->params.p) + param_bytes;

  k::V12Launch a{};
  a.n = n;
  a.S = S;
  a.L = L;
  a.Lloc = p.loc_loops;
  a.B_loc = p.B_loc;
  a.B_est = p.B_est;
  a.w_loc = p.w_loc;
  a.w_est = p.w_est;
  a.num = num;
  a.W = p.W_comb;
  a.comb_loops = cl;
  a.thr = p.thr_loops;
  a.x = reinterpret_cast<const double2 *const *>(dtab);
  a.params = P->ws->params.as<const int>();
  a.taps_loc = P->taps_loc.as<const double2>();
  a.taps_est = P->taps_est.as<const double2>();
  a.tapsT_loc32 = P->tapsT_loc32.as<const float2>();
  a.tapsT_est32 = P->tapsT_est32.as<const float2>();
  a.tR_loc = P->tR_loc;
  a.tR_est = P->tR_est;
  a.seg_rows = P->seg_rows;
  a.win_loc = P->win_loc.as<const double2>();
  a.win_est = P->win_est.as<const double2>();
  a.tw64 = P->tw64.as<const double2>();
  a.tw32 = P->tw32.as<const float2>();
  a.tw_nmax = P->tw_nmax;
  a.fp32 = fp32;
  a.loc64 = loc64;
  a.ref_conj = P->ref_conj;
  a.conj_sign = P->ref_conj ? 1.0 : -1.0;
  a.xs64 = need64 ? P->ws->xsamp.as<double2>() : nullptr;
  a.xs32 = fp32 ? P->ws->xsamp32.as<float2>() : nullptr;
  a.xs = xs;
  a.approved = P->ws->approved.as<int>();
  a.num_comb = P->ws->numcomb.as<int>();
  a.J = P->ws->J.as<int>();
  a.hits = P->ws->hits.as<int>();
  a.hits_cap = cap;
  a.count = P->ws->count.as<int64_t>();
  a.vote_cnt = p.comb ? nullptr : reinterpret_cast<int *>(a.count + S);
  a.o_freq = want_sparse ? reinterpret_cast<int32_t *const *>(dtab + tab_bytes) : nullptr;
  a.o_val = want_sparse ? reinterpret_cast<double2 *const *>(dtab + 2 * tab_bytes) : nullptr;
  a.o_dense = want_dense ? reinterpret_cast<double2 *const *>(dtab + 3 * tab_bytes) : nullptr;
  a.o_cap = io.cap;
  a.est_cap = cap;
  a.comb_buf = big_comb ? P->ws->comb.as<double2>() : nullptr;
  a.comb_tmp = big_comb ? P->ws->comb.as<double2>() + comb_elems : nullptr;
  a.mag_buf = mag_elems ? P->ws->mag.as<double>() : nullptr;
  a.tmp64 = big64 ? P->ws->tmp64.as<double2>() : nullptr;
  a.tmp32 = big32 ? P->ws->tmp32.as<float2>() : nullptr;
  // the fused fold + FFT keeps a whole bucket array in LDS
  a.fold_mode = std::max(p.B_loc, p.B_est) > kMaxLdsFft64 ? std::max(1, P->fold_mode)
                                                          : P->fold_mode;
  // alias-row bucket layout (schedule 1): v2, W a multiple of n/B for both
  // loop kinds, at least one wavefront of alias rows, FFTs in LDS
  a.rows_layout = p.comb && a.fold_mode == 1 && !dbg_votes && p.W_comb % (n / p.B_loc) == 0 &&
                  p.W_comb % (n / p.B_est) == 0 && n / p.W_comb >= 64 &&
                  std::max(p.B_loc, p.B_est) <= (fp32 ? 2 : 1) * kMaxLdsFft64;
  P->last_rows_layout = a.rows_layout;
  a.stream_blocks = P->stream_blocks;
  if (a.stream_blocks <= 0) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
      (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    a.stream_blocks = 2 * cus;  // 8 resident wavefronts per CU
  }

  // an arena workspace creates no runtime objects: no timing events, no
  // second stream, no graphs
  const bool timing = P->timing && !io.fixed_ws;
  const bool two_streams = p.comb && !io.fixed_ws;
  if (timing && !P->ev_made) {
    for (hipEvent_t &e : P->ev) HIP_OK(hipEventCreate(&e));
    for (hipEvent_t &e : P->ev_comb) HIP_OK(hipEventCreate(&e));
    P->ev_made = true;
  }
  if (two_streams && !P->stream2) {
    HIP_OK(hipStreamCreateWithFlags(&P->stream2, hipStreamNonBlocking));
    HIP_OK(hipEventCreateWithFlags(&P->ev_fork, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&P->ev_join, hipEventDisableTiming));
  }
  int launches[5] = {0, 0, 0, 0, 0};
  // Large v2 FP32 batches (alias-row layout) run in parts: the fold and
  // bucket FFTs of each part on the plan's stream, each part's estimation on
  // stream3 once that part and the comb are done, so the estimation (FP64
  // arithmetic) of part q overlaps the fold (memory latency) of part q+1.
  // Every signal's kernels and inputs are the same, so results are too.
  // Not with stage timing, whose stages are whole-batch launches.
  // SFFT_PARTS overrides the part count (tuning experiments; 1: unparted).
  int parts = 1;
  if (two_streams && fp32 && a.rows_layout && !timing && !dbg_votes && S >= 64) {
    parts = 2;
    if (const char *e = std::getenv("SFFT_PARTS")) parts = std::atoi(e);
    parts = std::max(1, std::min({parts, Plan::kMaxParts, S / 32}));
  }
  if (parts > 1 && !P->stream3) {
    HIP_OK(hipStreamCreateWithFlags(&P->stream3, hipStreamNonBlocking));
    for (hipEvent_t &e : P->ev_part) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&P->ev_est, hipEventDisableTiming));
  }
  // The stream work of one batch: uploads, kernels, counts back.  The v2
  // comb (its own gathers and FFTs) runs on stream2 beside the fold.
  auto enqueue = [&](bool timed) -> int {
    auto mark = [&](int i) {
      if (timed) (void)hipEventRecord(P->ev[i], st);
    };
    HIP_OK(hipMemcpyAsync(P->ws->params.p, hp, param_bytes + 4 * tab_bytes,
                          hipMemcpyHostToDevice, st));
    if (parts > 1) {
      hipStream_t st2 = P->stream2, st3 = P->stream3;
      HIP_OK(hipEventRecord(P->ev_fork, st));
      HIP_OK(hipStreamWaitEvent(st2, P->ev_fork, 0));
      if (k::comb(a, st2) != 0) return -1;
      HIP_OK(hipEventRecord(P->ev_join, st2));
      HIP_OK(hipStreamWaitEvent(st3, P->ev_join, 0));
      for (int q = 0; q < parts; q++) {
        const int b0 = (int)((int64_t)S * q / parts), b1 = (int)((int64_t)S * (q + 1) / parts);
        const k::V12Launch sa = sub_launch(a, b0, b1 - b0, stride, vchunks);
        if (k::fold_fft(sa, st, 0) < 0 || k::fold_fft(sa, st, 1) < 0) return -1;
        HIP_OK(hipEventRecord(P->ev_part[q], st));
        HIP_OK(hipStreamWaitEvent(st3, P->ev_part[q], 0));
        if (k::estimate(sa, st3) != 0) return -1;
      }
      HIP_OK(hipEventRecord(P->ev_est, st3));
      HIP_OK(hipStreamWaitEvent(st, P->ev_est, 0));
      HIP_OK(hipMemcpyAsync(h_counts, a.count, cnt_bytes, hipMemcpyDeviceToHost, st));
      return 0;
    }
    if (two_streams) HIP_OK(hipEventRecord(P->ev_fork, st));
    mark(1);
    // the fold is enqueued before the comb: for a single signal the host's
    // launch calls are on the latency path, and the fold branch is longer
    if ((launches[1] = k::fold_fft(a, st, 0)) < 0) return -1;
    if (p.comb && !two_streams) {
      if (k::comb(a, st) != 0) return -1;
      launches[0] = cl > 1 ? 2 : 1;
    } else if (p.comb) {
      hipStream_t st2 = P->stream2;
      HIP_OK(hipStreamWaitEvent(st2, P->ev_fork, 0));
      if (timed) (void)hipEventRecord(P->ev_comb[0], st2);
      if (k::comb(a, st2) != 0) return -1;
      if (timed) (void)hipEventRecord(P->ev_comb[1], st2);
      HIP_OK(hipEventRecord(P->ev_join, st2));
      launches[0] = cl > 1 ? 2 : 1;
    }
    mark(2);
    if ((launches[2] = k::fold_fft(a, st, 1)) < 0) return -1;
    mark(3);
    if (!p.comb && k::select_vote_v1(a, st) != 0) return -1;
    launches[3] = p.comb ? 0 : 4;
    if (two_streams) HIP_OK(hipStreamWaitEvent(st, P->ev_join, 0));
    mark(4);
    if (k::estimate(a, st) != 0) return -1;
    launches[4] = 1;
    if (dbg_votes && k::comb_vote_keys(a, P->ws->votes.as<int64_t>(), 2 * cap, st) != 0)
      return -1;
    mark(5);
    HIP_OK(hipMemcpyAsync(h_counts, a.count, cnt_bytes, hipMemcpyDeviceToHost, st));
    return 0;
  };
  P->last_fp32 = fp32;
  P->last_S = S;
  P->votes_stride = dbg_votes ? 2 * cap : 0;
  if (P->use_graphs && !timing && !io.fixed_ws) {
    // replay the batch's stream work from a graph captured on the first
    // batch of this shape; the workspace and staging addresses are part of
    // the key (kernel arguments and copy nodes hold them)
    const uint64_t key = (uint64_t)S | (uint64_t)fp32 << 32 | (uint64_t)want_sparse << 33 |
                         (uint64_t)want_dense << 34 | (uint64_t)a.fold_mode << 35 |
                         (uint64_t)a.stream_blocks << 36;
    const uint64_t sig = workspace_signature(*P->ws) ^
                         ((uint64_t)io.cap * 0x9e3779b97f4a7c15ull + (uint64_t)a.ref_conj);
    hipGraphExec_t ex = nullptr;
    for (const auto &g : P->graphs)
      if (g.key == key && g.sig == sig) ex = g.exec;
    if (!ex) {
      HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
      const int rc = enqueue(false);
      hipGraph_t graph = nullptr;
      const hipError_t e = hipStreamEndCapture(st, &graph);
      if (rc != 0 || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (rc == 0) set_error(std::string("graph capture: ") + hipGetErrorString(e));
        return -1;
      }
      const hipError_t ei = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) {
        set_error(std::string("graph instantiate: ") + hipGetErrorString(ei));
        return -1;
      }
      if (P->graphs.size() >= 8) {
        (void)hipGraphExecDestroy(P->graphs.front().exec);
        P->graphs.erase(P->graphs.begin());
      }
      P->graphs.push_back(Plan::GraphEntry{key, sig, ex});
    }
    HIP_OK(hipGraphLaunch(ex, st));
  } else if (enqueue(timing) != 0) {
    return -1;
  }
  V12Pending local;
  V12Pending &q = pd ? *pd : local;
  q.st = st;
  q.timing = timing;
  q.S = S;
  q.h_counts = h_counts;
  std::copy(launches, launches + 5, q.launches);
  return pd ? 0 : v12_finish(P, io, q);
}

int exec_batch(Plan *P, BatchIO &io)
{
  if (P->version == SFFT_VERSION_3) return exec_batch_v3(P, io);
  return exec_batch_v12(P, io, nullptr);
}

int exec_batch_async(Plan *P, BatchIO &io, V12Pending &pd)
{
  if (P->version == SFFT_VERSION_3) {
    set_error("exec_batch_async: v1/v2 plans only");
    return -1;
  }
  return exec_batch_v12(P, io, &pd);
}

int exec_batch_wait(Plan *P, BatchIO &io, const V12Pending &pd)
{
  return v12_finish(P, io, pd);
}

}  // namespace sfa
