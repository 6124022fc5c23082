// Internal types of libsfft_amd: plan parameters, host filter design, the
// plan object with its device-resident tables, and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <mutex>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/sfft.h"

namespace sfa {

using cd = std::complex<double>;

// Largest bucket / comb / Mansour transform a single workgroup keeps in LDS.
constexpr int kMaxLdsFft64 = 8192;
// Largest number of loops (location + estimation) the median network takes.
constexpr int kMaxLoops = 32;
// Mansour offsets of the v3 refine form (equations per bucket).
constexpr int kV3RefineOffsets = 8;
// Dynamic LDS of the v3 kernels: 160 KiB per workgroup less 4 KiB kept for
// their static arrays (scan scratch, estimate ranks, loop flags)
constexpr size_t kV3MaxDynLds = 160 * 1024 - 4096;

void set_error(const std::string &msg);

// Lets `kernel` launch with up to the device's opt-in LDS per workgroup
// (160 KiB on gfx950).  The limit is raised once per kernel and device and
// never lowered, so launches from concurrent threads cannot undercut each
// other.  Kernels call it before every launch above the default 64 KiB.
void raise_lds_limit(const void *kernel);

// Device-pointer test with a cache of the allocation ranges met during ONE
// API call (a batch of rows of one tensor costs one query).  The cache must
// not outlive the call: once the caller frees an allocation, its address
// range can be handed out again as host memory, and a stale entry would
// make the kernels dereference a host address.
class PtrKinds {
 public:
  bool is_device(const void *p) { return covers(p, 1); }

  // p is device (or managed) memory and its allocation holds [p, p + bytes).
  bool covers(const void *p, size_t bytes)
  {
    if (!p) return false;
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    for (int i = 0; i < used_; i++)
      if (lo_[i] <= a && a < hi_[i]) return bytes <= hi_[i] - a;
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    const bool dev = attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
    if (dev) {
      hipDeviceptr_t base = nullptr;
      size_t size = 0;
      if (hipMemGetAddressRange(&base, &size, const_cast<void *>(p)) == hipSuccess && size) {
        lo_[next_] = reinterpret_cast<uintptr_t>(base);
        hi_[next_] = lo_[next_] + size;
        const bool fits = bytes <= hi_[next_] - a;
        next_ = (next_ + 1) % kSlots;
        used_ = std::min(used_ + 1, kSlots);
        return fits;
      }
      (void)hipGetLastError();
    }
    return dev;
  }

 private:
  static constexpr int kSlots = 128;  // a batch's inputs are often separate allocations
  uintptr_t lo_[kSlots] = {}, hi_[kSlots] = {};
  int used_ = 0, next_ = 0;
};

// ---------------------------------------------------------------- params
// sFFT v1/v2 sizes (src/sfft.cc:298-372 with src/parameters.cc).
struct V12Params {
  int n = 0, k = 0;
  bool comb = false;
  double Bcst_loc = 1, Bcst_est = 1, Comb_cst = 2, tol_loc = 1e-8, tol_est = 1e-8;
  int loc_loops = 4, est_loops = 16, thr_loops = 3, comb_loops = 1;
  double BB_loc = 0, BB_est = 0;
  int b_loc = 0, b_est = 0, B_loc = 0, B_est = 0, B_thresh = 0, W_comb = 0;
  int w_loc = 0, w_est = 0;
  int loops() const { return loc_loops + est_loops; }
  int64_t x_samp_size() const {
    return (int64_t)loc_loops * B_loc + (int64_t)est_loops * B_est;
  }
};

// sFFT v3 sizes (src/sfft.cc:506-560).
struct V3Params {
  int n = 0, k = 0;
  int W_Man = 0, B_g1 = 0, B_g2 = 0, w_g1 = 0, w_g2 = 0, b_g1 = 0, b_g2 = 0;
  double BB1 = 0, BB2 = 0;
};

int floor_to_pow2(double x);
int dolph_chebyshev_width(double lobefrac, double tolerance);
V12Params derive_v12(int n, int k, bool comb);
V3Params derive_v3(int n, int k);
// Empty string when the plan can run, else the reason.
std::string check_v12(const V12Params &p);
std::string check_v3(const V3Params &p);

// ---------------------------------------------------------------- host FFT
// In-place unnormalised DFT, e^{-2 pi i jk/n} (forward) or e^{+..}.
void host_fft(cd *x, int n, bool backward);

// ---------------------------------------------------------------- filters
struct HostFilter {
  std::vector<cd> time;  // w taps
  std::vector<cd> freq;  // n-point response
};
// Dolph-Chebyshev window flattened over b bins (src/filters.cc:70-160).
HostFilter design_filter(double lobefrac, double tolerance, int n, int b);
// The same design with its two n-point DFTs on the GPU (plan time); the
// running sum, normalisation and iterated phase ramp stay on the host.
// Returns false (set_error) on a HIP failure.
bool design_filter_gpu(double lobefrac, double tolerance, int n, int b, HostFilter &f);

// ---------------------------------------------------------------- RNG
// The reference's random() / drand48() draws, from the process-global glibc
// streams or from private streams seeded like srandom()/srand48().
struct Rng {
  bool seeded = false;
  struct random_data rd;
  char state[128];
  struct drand48_data dd;
  long next_random();
  double next_drand48();
  void seed(unsigned rs, long ds);
};
int gcd_int(int a, int b);
int mod_inverse_int(int a, int n);

// ---------------------------------------------------------------- device
// A fixed device region that DevBufs are carved from instead of allocated
// (the allocation-free exec).
struct Arena {
  char *base = nullptr;
  size_t size = 0, used = 0;
};

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  Arena *arena = nullptr;      // carve from here (never hipMalloc / hipFree)
  bool ensure(size_t nbytes);  // grow-only
  void release();
  template <typename T> T *as() const { return reinterpret_cast<T *>(p); }
};

// Per-exec device workspace (grow-only) and pinned staging.  A plan owns
// one; plan clones run by one server worker share that worker's.
struct Workspace {
  DevBuf params, xsamp, xsamp32, approved, numcomb, J, hits, ptrs, out_freq, out_val, count,
      stage, v3, v3_status;
  // scratch of the sizes above the LDS limits (comb FFT, v1 selection,
  // bucket FFT passes)
  DevBuf comb, mag, tmp64, tmp32;
  DevBuf multi;  // descriptors of a multi-plan launch
  DevBuf votes;  // v2 comb-vote diagnostic: S x 2 x located int64
  void *pinned = nullptr;
  size_t pinned_bytes = 0;
  bool pinned_view = false;  // pinned is the caller's host buffer (not freed)
  // every buffer carved from one fixed region (nothing allocated or freed)
  void use_arena(Arena *a);
  size_t device_bytes() const;
  void release();
};

struct Plan {
  // API calls on one plan are serialised on this (workspace, RNG streams and
  // pinned staging are per plan)
  std::mutex mu;
  sfft_version version = SFFT_VERSION_1;
  int n = 0, k = 0;
  V12Params v12;
  V3Params v3;
  sfft_precision precision = SFFT_PRECISION_FP64;
  int fold_mode = 1;       // see k::V12Launch::fold_mode
  int stream_blocks = 0;   // 0: derived from the CU count
  bool last_fp32 = false;  // precision of the buckets in ws_xsamp
  bool last_rows_layout = false;  // the last batch stored buckets by alias row
  bool ref_conj = true;
  bool debug_votes = false;  // v2 FP64: count the comb votes (diagnostic)
  bool debug_v3_serial = false;  // v3: every subtraction pass by one thread per array
  int debug_v3_map_limit = 0;     // v3: map capacity below Mcap (0: Mcap)
  bool debug_v3_inline_pre = false;  // v3: transforms run their own pre tasks
  int64_t votes_stride = 0;  // int64 per signal in ws->votes of the last batch
  Rng rng;
  hipStream_t stream = nullptr;
  bool own_stream = false;

  // v1/v2 device tables
  DevBuf taps_loc, taps_est;      // double2[w]
  DevBuf tapsT_loc32, tapsT_est32;  // float2[B * ceil(w/B)]: FP32 taps by bucket
  int tR_loc = 0, tR_est = 0;       // ceil(w/B) of each filter
  int seg_rows = 256;               // rows per chunk of the FP32 fold (sum order)
  DevBuf win_loc, win_est;        // double2[m+1]: freq[(n-d) mod n], d = -m/2..m/2
  DevBuf tw64, tw32;              // e^{-2 pi i j / Nmax}, j < Nmax/2
  int tw_nmax = 0;

  // host copies of the two filters: taps, and the response at signed offsets
  // -resp_half..resp_half (resp[d + resp_half] = freq[d mod n])
  std::vector<cd> f_time[2], f_resp[2];
  int resp_half[2] = {0, 0};

  // v3 device tables
  DevBuf v3_time1, v3_time2;  // double2[w_g*]
  DevBuf v3_freq1, v3_freq2;  // double2[3n/B + 1]: responses at offsets -3n/2B..3n/2B
  int v3_half1 = 0, v3_half2 = 0;
  int v3_Mcap = 0, v3_Hcap = 0, v3_Fcap = 0, v3_Rcap = 0;
  size_t v3_lds = 0;
  std::vector<int> last_status;  // v3: per signal of the last exec call

  Workspace own_ws;
  Workspace *ws = &own_ws;  // a clone points at its server worker's
  bool owns_tables = true;  // false for a clone: tables belong to the source
  int last_S = 0;           // signals in the last batch (for diagnostics)

  // v1/v2 batches replayed from captured HIP graphs (per batch shape and
  // workspace addresses; at most 8 kept)
  bool use_graphs = false;
  struct GraphEntry {
    uint64_t key, sig;
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs;

  // stage timing (hipEvents on the plan's stream)
  static constexpr int kStages = 6;
  bool timing = false;
  double stage_ms[kStages] = {0, 0, 0, 0, 0, 0};
  int64_t stage_launches[kStages] = {0, 0, 0, 0, 0, 0};
  hipEvent_t ev[kStages + 1] = {};
  hipEvent_t ev_comb[2] = {};  // comb start / end on stream2 (timing)
  bool ev_made = false;
  // The v2 comb runs on a second stream beside the fold and bucket FFTs;
  // fork after the uploads, join before estimation.
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // Large v2 FP32 batches in parts: the fold and bucket FFTs of part q+1
  // on the plan's stream while part q is estimated on stream3.
  static constexpr int kMaxParts = 8;
  hipStream_t stream3 = nullptr;
  hipEvent_t ev_part[kMaxParts] = {}, ev_est = nullptr;

  int64_t output_capacity() const;
  int64_t plan_bytes() const;
  int64_t bytes_per_signal() const;
};

int64_t v12_output_capacity(const V12Params &p);
int64_t v3_map_capacity(const V3Params &q);

Plan *make_plan(int n, int k, sfft_version v);
// Times the candidate fold schedules of a v1/v2 plan on `batch` synthetic
// signals and keeps the fastest (also for later plans of the same key).
// No effect on results or on the plan's RNG streams.  0 or -1.
int autotune_plan(Plan *P, int batch);
// Applies a schedule tuned earlier in this process for P's key, if any.
void apply_tuned_schedule(Plan *P);
// A plan sharing src's device tables, running on `stream` with workspace ws
// (both owned by the caller; null ws: a workspace of its own); src must
// outlive it.
Plan *clone_plan(const Plan *src, Workspace *ws, hipStream_t stream);
void free_plan(Plan *P);

// One batch of signals.  Every array here is a host array of S device
// pointers; a null freq/val/dense array (or entry) means "not wanted".
struct BatchIO {
  int S = 0;
  Rng *const *rng = nullptr;        // per-signal RNG streams (null: the plan's)
  const double2 *const *in = nullptr;
  int32_t *const *freq = nullptr;   // sparse frequencies, cap entries each
  double2 *const *val = nullptr;    // sparse values, cap entries each
  double2 *const *dense = nullptr;  // n values each, already zero-filled
  int64_t cap = 0;
  int64_t *counts = nullptr;        // host: S located counts (out)
  // the plan's workspace is an arena: no graph capture, no timing events,
  // no second stream (each would create runtime objects)
  bool fixed_ws = false;
};

// Draws the RNG values of S signals in signal order, runs the kernels on
// the plan's stream and waits for them.  Returns 0, -1 (set_error), or for
// v3 1 when a signal's map overflowed (every result written, error set).
int exec_batch(Plan *P, BatchIO &io);
// A v1/v2 batch in two halves: exec_batch_async draws and enqueues the
// stream work (uploads from the plan's pinned staging, kernels, the count
// copy back) and returns; exec_batch_wait waits for the plan's stream and
// delivers the counts.  Batches of different plans can be in flight
// together; the plan must not run anything else in between.
struct V12Pending {
  hipStream_t st = nullptr;
  bool timing = false;
  int S = 0;
  int launches[5] = {0, 0, 0, 0, 0};
  const int64_t *h_counts = nullptr;  // in the plan's pinned staging
};
int exec_batch_async(Plan *P, BatchIO &io, V12Pending &pd);
int exec_batch_wait(Plan *P, BatchIO &io, const V12Pending &pd);
// v3 batches of G plans (different n, k) as one launch on stream st; the
// descriptors live in ws (not any plan's).  Plans' own workspaces hold their
// signals' state.  Returns 0, -1, or 1 (a map overflowed, results written).
int exec_multi_v3(Plan *const *plans, BatchIO *ios, int G, hipStream_t st, Workspace *ws,
                  Plan *timer = nullptr);

// ---------------------------------------------------------------- kernels
namespace k {
struct V12Launch {
  int n, S, L, Lloc;
  int B_loc, B_est, w_loc, w_est, num, W, comb_loops, thr;
  const double2 *const *x;        // device array of S signal pointers
  const int *params;              // S x (2L + comb_loops): a[L], ai[L], off[cl]
  const double2 *taps_loc, *taps_est;
  // FP32 taps by bucket: tapsT[b * tR + k] = tap[b + k*B] (k < tR = ceil(w/B))
  const float2 *tapsT_loc32, *tapsT_est32;
  int tR_loc, tR_est;
  int seg_rows;    // rows per chunk of the FP32 fold (fixes its summation order)
  const double2 *win_loc, *win_est;
  const double2 *tw64;
  const float2 *tw32;
  int tw_nmax;
  bool fp32;       // estimation loops (and v2 location loops) in FP32
  bool loc64;      // v1 under fp32: location loops stay FP64
  bool ref_conj;
  double conj_sign;  // 1.0 with ref_conj, else -1.0 (the FP32 form's estimate sign)
  double2 *xs64;   // S x xs buckets of the FP64 loops
  float2 *xs32;    // S x xs buckets of the FP32 loops
  int64_t xs;
  int *approved;   // S x comb_loops*num
  int *num_comb;   // S
  int *J;          // S x Lloc x num
  int *hits;       // S x hits_cap
  int *vote_cnt;   // v1: S x vote_chunks hit counts, then offsets
  int64_t hits_cap;
  int64_t *count;  // S located counts
  int32_t *const *o_freq;   // S pointers or null
  double2 *const *o_val;    // S pointers or null
  double2 *const *o_dense;  // S pointers or null
  int64_t o_cap;            // sparse capacity per signal
  int64_t est_cap;          // upper bound of the located count per signal
  int fold_mode;            // 0: fused fold+FFT per (signal, loop) block;
                            // 1: signal-major streamed fold, then bucket FFTs
                            //    (alias-row layout where it applies);
                            // 2: as 1, buckets always in natural order
  int stream_blocks;        // resident blocks of the streamed fold
  // scratch for transforms and selections that do not fit in LDS (null
  // when unused)
  double2 *comb_buf, *comb_tmp;  // S x comb_loops x W sub-sampled comb signals
  double *mag_buf;               // |.|^2 of comb or v1 selection arrays
  double2 *tmp64;                // S x xs: radix-2 pass copy of xs64
  float2 *tmp32;                 // S x xs: radix-2 pass copy of xs32
  // v2 with W a multiple of n/B: buckets after the FFT stored by alias row
  // (see k_bucket_fft_rows); estimation reads them contiguously
  bool rows_layout;
};
// Blocks per signal of the v1 vote: its candidates (the ranges of the
// selected buckets of loops 0..Lloc-thr) in chunks of 8 * 256.
inline int vote_chunks(int n, int B_loc, int num, int Lloc, int thr)
{
  const int64_t total = (int64_t)std::max(Lloc - thr + 1, 0) * num * (n / B_loc);
  return (int)std::max<int64_t>(1, (total + 2047) / 2048);
}
// Largest comb transform / selection array kept in LDS.
constexpr int kMaxLdsComb = 8192;
constexpr int kMaxLdsSelect = 16384;
struct V3Launch {
  int n, S, k;
  int W, B1, B2, w1, w2;
  const double2 *const *x;   // S signal pointers
  const int *params;         // S x 6: a, ai, b, shift, init_offset, init_G_offset
  const double2 *taps1, *taps2;
  const double2 *win1, *win2;  // filter response at offsets -half..half
  int win1_half, win2_half;
  const double2 *tw64;
  int tw_nmax;
  // per-signal workspace, strided by signal
  double2 *ms;               // 2W (interleaved loops)
  double2 *ms_pre, *ms_tmp;  // W above the LDS size: S x ms_loops x W
                             // transforms computed before k_v3_run (else null)
  int ms_loops;              // Mansour offsets in ms_pre: the reference's
                             // two (init_offset + 0, 1), then with refine
                             // init_offset + l ref_t, 1 <= l < kV3RefineOffsets
  // refine form: the located values re-estimated from fresh Mansour buckets
  // at offsets init_offset + l ref_t, l < kV3RefineOffsets (S x offsets x W
  // in ref_ms; rows 0, 2, 3, ... of ms_pre when W is above the LDS size).
  // ref_t is odd, near 0.618 n/W: members of a bucket m W apart sit at
  // phase steps m ref_t / (n/W) mod 1, well apart for the small m at which
  // the reference places its false positives; ref_tinv = ref_t^-1 mod n/W.
  double2 *gpre;             // Gauss buckets from the pre tasks: S x (2 B1 + 2 B2)
  int refine;
  int ref_t, ref_tinv;
  int serial_all;            // every subtraction pass serial (diagnostic)
  int map_limit;             // map entries allowed (Mcap, or lower for diagnostics)
  int inline_pre;            // pre-task workgroups leave their tasks (diagnostic)
  double2 *ref_ms;
  int *map_keys;             // Mcap (insertion order)
  double2 *map_vals;         // Mcap
  int *map_hidx;             // Hcap hash slots -> entry
  int Mcap, Hcap;
  int *cf, *ef;              // Fcap candidates / found
  double2 *cv, *ev;
  int Fcap;
  double2 *rem;              // Rcap x 14 subtraction values
  int *remi;                 // Rcap x 3 bucket indices
  int Rcap;
  double2 *phase;            // Pcap phases of the permuted samples
  int Pcap;
  int max_passes;
  int64_t *count;            // S located counts
  int *status;               // S: 1 early return, 2 map full, 4 pass cap,
                             // 8 refine kept some reference values,
                             // 16 gave up waiting for a pre task (a defect)
  int32_t *const *o_freq;
  double2 *const *o_val;
  double2 *const *o_dense;
  int64_t o_cap;
  size_t lds_bytes;
};
// G groups in one launch; d_groups / d_blocks: the descriptors and the
// nblocks (group, signal) pairs, d_claims 4 nblocks zeroed ints, already on
// the device (enqueued on st).
int v3_multi(const V3Launch *h_groups, int G, const V3Launch *d_groups, const int2 *d_blocks,
             int nblocks, int *d_claims, hipStream_t st);

// Batched power-of-two FFTs; transform t = (outer, inner) lives at
// data + outer*so + inner*si.  In LDS for small forward sizes (tw: the
// plan's table for tw_nmax >= N), else Stockham passes ping-ponging with tmp
// (same layout); the result always ends in data.
int batch_fft64(double2 *data, double2 *tmp, int N, int64_t nb_outer, int nb_inner,
                int64_t so, int64_t si, bool backward, const double2 *tw, int tw_nmax,
                hipStream_t st);
int batch_fft32(float2 *data, float2 *tmp, int N, int64_t nb_outer, int nb_inner, int64_t so,
                int64_t si, bool backward, const float2 *tw, int tw_nmax, hipStream_t st);

// Forward radix-2 DIT FFTs of any power-of-two size with the plan's twiddle
// table, bit-identical to the in-LDS form; tmp (same layout as data) is
// needed when N exceeds kMaxLdsFft64 (FP64) / 2*kMaxLdsFft64 (FP32).
int radix2_fft64(double2 *data, double2 *tmp, int N, int64_t nb_outer, int nb_inner, int64_t so,
                 int64_t si, const double2 *tw, int tw_nmax, hipStream_t st);
int radix2_fft32(float2 *data, float2 *tmp, int N, int64_t nb_outer, int nb_inner, int64_t so,
                 int64_t si, const float2 *tw, int tw_nmax, hipStream_t st);

// The allocation-free exec's scratch region (kNoAllocScratch bytes of
// device memory per GPU, a file-scope __device__ array): noalloc_setup()
// looks up its address on the current device (once per device; a plan's
// construction calls it), noalloc_scratch() returns it or null.
constexpr size_t kNoAllocScratch = (size_t)32 << 20;
int noalloc_setup();
void *noalloc_scratch(int dev);

int comb(const V12Launch &a, hipStream_t st);
// part 0: the fold (fused with the bucket FFTs when fold_mode == 0);
// part 1: the bucket FFTs (nothing when fused).  Returns launches or -1.
int fold_fft(const V12Launch &a, hipStream_t st, int part);
int select_vote_v1(const V12Launch &a, hipStream_t st);
// Diagnostic of v2 plans in FP64 with buckets in natural order: the
// location-loop selections and, per located frequency (prefill order), the
// pair (append key, frequency) of the reference's comb votes, key -1 for a
// frequency never appended; keys holds S x stride int64.
int comb_vote_keys(const V12Launch &a, int64_t *keys, int64_t stride, hipStream_t st);
int estimate(const V12Launch &a, hipStream_t st);
int zero_dense(double2 *out, int64_t n, hipStream_t st);
// total values of uniform noise in [-1, 1)^2 (synthetic tuning input)
int fill_noise(double2 *x, int64_t total, uint64_t seed, hipStream_t st);
}  // namespace k

}  // namespace sfa
