/*
 * sfft_amd -- MI355X-native Sparse FFT behind the reference sFFT C API.
 *
 * The first block of entry points is the reference's public API
 * (src/sfft.h:158-167 of davidediger/sfft) with the same names, argument
 * meaning and enum values; a program written against the reference links
 * against libsfft_amd.so unchanged.  The second block adds what the
 * reference documents but does not export (a sparse result) and what a GPU
 * caller needs (precision, seeded RNG streams, streams, errors).
 *
 * Pointers passed to sfft_exec / sfft_exec_many may be host or device
 * (HIP) pointers; the library detects which on every call (one
 * hipPointerGetAttributes per allocation met in the call; nothing is
 * remembered between calls, because a freed device range can come back as
 * host memory).  Device-resident inputs avoid the PCIe copy of the whole
 * signal.
 *
 * Threads: calls on one plan from several threads take turns (the plan's
 * workspace, RNG streams and staging buffers are shared); different plans
 * run concurrently, each on its own stream.
 */
#ifndef SFFT_AMD_SFFT_H
#define SFFT_AMD_SFFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#include <complex>
typedef std::complex<double> complex_t;
extern "C" {
#else
#include <complex.h>
typedef double complex complex_t;
#endif

/* src/sfft.h:37-42 */
typedef enum sfft_version {
  SFFT_VERSION_1 = 0,
  SFFT_VERSION_2 = 1,
  SFFT_VERSION_3 = 2
} sfft_version;

/* src/sfft.h:44-50: same leading members; data is the library's plan. */
typedef struct sfft_plan {
  sfft_version version;
  unsigned int n;
  unsigned int k;
  void *data;
} sfft_plan;

/* ---- reference API ------------------------------------------------------ */

/* src/sfft.h:159 / src/sfft.cc:60-63: 16-byte aligned host allocation. */
void *sfft_malloc(size_t s);
/* src/sfft.h:160 / src/sfft.cc:65-69 */
void sfft_free(void *p);

/* src/sfft.h:162-163 / src/sfft.cc:71-101.  Builds parameters, filters and
 * device-resident tap tables.  Returns NULL (and sets sfft_last_error) for a
 * bad version, n not a power of two, or an (n,k) the algorithm cannot run;
 * the reference aborts or runs out of bounds on the last two.
 * fftw_optimization: FFTW_ESTIMATE (64) uses the schedule tuned earlier in
 * this process for the same key, if any; FFTW_MEASURE (0) runs the autotuner
 * (sfft_plan_autotune) on v1/v2 plans. */
sfft_plan *sfft_make_plan(int n, int k, sfft_version version,
                          int fftw_optimization);

/* src/sfft.h:164 / src/sfft.cc:103-117.  Frees everything the plan owns. */
void sfft_free_plan(sfft_plan *plan);

/* src/sfft.h:165 / src/sfft.cc:119-137.  out (n values) is zero-filled and
 * the located coefficients are written, as the reference does. */
void sfft_exec(sfft_plan *plan, complex_t *in, complex_t *out);

/* src/sfft.h:166 / src/sfft.cc:139-147.  All signals go to the GPU as one
 * batch; RNG draws are consumed in signal order 0..num-1. */
void sfft_exec_many(sfft_plan *plan, int num, complex_t **in,
                    complex_t **out);

/* ---- extensions --------------------------------------------------------- */

/* Sparse result: the documented sfft_output (src/sfft.h:34-35,
 * doc/usage.rst:64-90) as a C struct.  freq/val have room for capacity
 * entries; count is set by the exec call.  Buffers may be host or device
 * memory.  When count > capacity only capacity entries are written. */
typedef struct sfft_output_c {
  int64_t count;
  int64_t capacity;
  int32_t *freq;
  complex_t *val;
} sfft_output_c;

typedef enum sfft_precision {
  /* FP64 throughout; reproduces the reference's arithmetic order. */
  SFFT_PRECISION_FP64 = 0,
  /* v1/v2: FP32 bucketization and bucket FFTs (compensated FP32 sums), FP64
   * comb selection and FP64 estimation of the located values.  v3: the
   * reference's FP64 location (FP32 phases cannot locate at n = 2^24), then
   * the located values re-estimated in FP64 by least squares over fresh
   * Mansour buckets at 8 strided offsets with integer-exact phases, with up
   * to two unlocated tones per bucket found by a Prony step (located set
   * unchanged; false positives come out near zero). */
  SFFT_PRECISION_FP32_REFINE = 1
} sfft_precision;

/* Upper bound of sfft_output_c entries one exec of this plan can produce. */
int64_t sfft_plan_output_capacity(const sfft_plan *plan);

/* One signal, sparse result. Returns 0 on success, -1 on error. */
int sfft_exec_sparse(sfft_plan *plan, const complex_t *in, sfft_output_c *out);

/* num signals in one batch, sparse results. Returns 0 or -1. */
int sfft_exec_many_sparse(sfft_plan *plan, int num, const complex_t *const *in,
                          sfft_output_c *out);

/* sfft_exec_sparse without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * per-call workspace (buckets, comb residues, votes, counts) is carved from a
 * fixed 32 MiB device region of the library, one per GPU, whose address is
 * looked up when a plan is built on that GPU; no hipMalloc, hipFree or
 * allocating routine is called.  v1/v2 plans only; in, out->freq and
 * out->val must be device memory (n values; out->capacity >=
 * sfft_plan_output_capacity entries).  A plan whose workspace exceeds the
 * region is rejected with an error before any RNG draw or launch.  Every
 * call on one GPU uses the same region: calls that overlap on one GPU (from
 * several threads or plans) race.  Returns 0 or -1. */
int sfft_exec_sparse_noalloc(sfft_plan *plan, const complex_t *in, sfft_output_c *out);

/* Reference-compatible synthetic input (src/simulation.cc:95-112,
 * src/timing_many.cc:117-139): k frequencies drawn with replacement as
 * floor(drand48() * n) with X[f] = 1, x = the unnormalised inverse DFT
 * (e^{+2 pi i f t / n}); with std_noise != 0, complex Gaussian noise added
 * as AWGN does (src/utils.cc:250-280).  The drand48 stream is private,
 * seeded like srand48(drand48_seed).  Host pointers: x (n values), x_f (n
 * values or NULL), snr (NULL or the measured signal/noise power ratio, 1e9
 * when noiseless).  Returns 0 or -1. */
int sfft_generate_input(int n, int k, long drand48_seed, double std_noise, complex_t *x,
                        complex_t *x_f, double *snr);

/* Signals of different plans in one call (no counterpart in the
 * reference, whose sfft_exec_many shares one plan, src/sfft.cc:139-147):
 * signal i runs plans[i]; each plan draws for its signals in their order
 * here.  The v3 signals of all plans run as one launch (one workgroup per
 * signal, each reading its own plan's tables), so a batch can mix n and k;
 * that launch goes on the stream of the first v3 plan, after the work
 * enqueued so far on the other v3 plans' streams.  v1/v2 signals run as
 * one batch per plan, on that plan's stream.  in[i], out[i].freq and
 * out[i].val must be device pointers, with out[i].capacity at least
 * sfft_plan_output_capacity(plans[i]).  Returns 0 or -1. */
int sfft_exec_many_plans(int num, sfft_plan *const *plans, const complex_t *const *in,
                         sfft_output_c *out);

/* Precision of later execs of this plan (default SFFT_PRECISION_FP64). */
int sfft_plan_set_precision(sfft_plan *plan, sfft_precision precision);

/* Fold schedule of a v1/v2 plan (no counterpart in the reference; it maps
 * src/computefourier-1.0-2.0.cc:222-275 onto the GPU).  fold_mode 0: one
 * workgroup per (signal, loop) folds into LDS and transforms there; 1: a
 * signal-major streamed fold with stream_blocks resident workgroups (0: twice
 * the CU count), then the bucket FFTs, which for v2 store the buckets by
 * alias row so that estimation reads them contiguously; 2: as 1 with the
 * buckets in natural order.  Schedules never change results.
 * Returns 0 or -1. */
int sfft_plan_set_schedule(sfft_plan *plan, int fold_mode, int stream_blocks);
int sfft_plan_get_schedule(const sfft_plan *plan, int *fold_mode, int *stream_blocks);

/* HIP graphs (no counterpart in the reference): with enabled != 0, each
 * v1/v2 exec replays its uploads, kernels and count read-back from a graph
 * captured on the first batch of the same shape (signal count, precision,
 * outputs, schedule) and workspace addresses; the host still draws the RNG
 * values and writes the pinned staging buffers the graph copies from.  Cuts
 * the launch overhead of small batches; results are identical.  Stage timing
 * bypasses the graphs.  Returns 0 or -1. */
int sfft_plan_enable_graphs(sfft_plan *plan, int enabled);

/* Autotuner: times the candidate schedules on `batch` synthetic signals and
 * keeps the fastest, for this plan and for later plans of the same (n, k,
 * version, precision) in this process.  Draws nothing from the plan's RNG
 * streams.  sfft_make_plan with FFTW_MEASURE (0) runs it with batch 64, as
 * FFTW_MEASURE times FFTW's candidate plans (src/sfft.cc:274-295).  Returns
 * 0 or -1. */
int sfft_plan_autotune(sfft_plan *plan, int batch);

/* RNG streams.  By default a plan draws from the process-global glibc
 * random() and drand48() streams exactly as the reference does (a caller's
 * srand()/srand48() decide the draws).  Seeding gives the plan private
 * streams equal to srandom(random_seed) / srand48(drand48_seed) but
 * unaffected by other users of the global ones. */
int sfft_plan_seed(sfft_plan *plan, unsigned int random_seed,
                   long drand48_seed);

/* HIP stream (hipStream_t) the plan's work is enqueued on; NULL = the
 * library's own stream.  Exec calls return after their work completed. */
int sfft_plan_set_stream(sfft_plan *plan, void *hip_stream);

/* Values as the reference returns them (default, 1) or with the v1/v2
 * conjugation of src/computefourier-1.0-2.0.cc:383-398 undone (0).  v3 is
 * unaffected. */
int sfft_plan_set_reference_conjugate(sfft_plan *plan, int enabled);

/* Device bytes held by the plan and needed per signal of a batch. */
int sfft_plan_memory(const sfft_plan *plan, int64_t *plan_bytes,
                     int64_t *bytes_per_signal);

/* Message of the last error on this thread ("" if none). */
const char *sfft_last_error(void);

/* Parameters the plan derived (for tests and tools).  Returns 0 or -1.
 * p (10 ints): B_loc, B_est, B_thresh, W_comb, comb_loops, loc_loops,
 * est_loops, thr_loops, w_loc, w_est.  For v3: W_Man, B_g1, B_g2, w_g1,
 * w_g2, 0... */
int sfft_plan_params(const sfft_plan *plan, int *p10);

/* The same parameters for (n, k, version) without building a plan (host
 * only, no GPU needed).  Returns 0 when sfft_make_plan would accept the
 * triple, 1 when it would reject it (reason in sfft_last_error), -1 on a
 * bad version. */
int sfft_derive_params(int n, int k, sfft_version version, int *p10);

/* v3 outcome flags of the signals of the plan's last exec call, in signal
 * order (for sfft_exec_many_plans: this plan's signals in call order): bit 0
 * the reference's early return (src/computefourier-3.0.cc:859-860, empty
 * result), bit 1 result map overflow (the call returns -1 after writing
 * every signal's count and result, this signal's truncated to the map's
 * capacity), bit 2 the estimate/subtract loop stopped at its 3*4096-pass
 * cap instead of at repeated bucket counts, bit 3 the refine form kept the
 * reference's values in a bucket it could not solve, bit 4 the transform
 * gave up waiting (about a second) for a bucket task another workgroup had
 * started -- a library defect, never seen in normal runs; the call then
 * returns -1 as for bit 1, and that signal's result is not valid.  Copies up to num flags; returns the number of
 * signals recorded (0 for v1/v2 plans), or -1 for a bad plan. */
int sfft_plan_last_status(const sfft_plan *plan, int *flags, int num);

/* Stage timing with HIP events recorded on the plan's stream around each
 * kernel stage of every later exec (0 disables, any other value enables and
 * resets the totals). */
int sfft_plan_enable_timing(sfft_plan *plan, int enabled);

/* Totals since timing was enabled: for stage s (0 comb, 1 fold (gather and
 * bucket sums; includes the bucket FFTs when they are fused into it),
 * 2 bucket FFTs, 3 selection+votes, 4 estimation, 5 v3) ms[s] = summed
 * milliseconds and launches[s] = kernel launches timed.  Arrays of 6. */
int sfft_plan_stage_times(const sfft_plan *plan, double *ms6,
                          int64_t *launches6);

/* Diagnostics: the bucket values (after the bucket FFTs) of signal `signal`
 * of the plan's last v1/v2 batch, all loops concatenated as the reference's
 * x_samp (src/computefourier-1.0-2.0.cc:228-230): count entries into host
 * memory, in natural bucket order whatever the schedule's storage order.
 * Returns the number written or -1. */
int64_t sfft_debug_buckets(const sfft_plan *plan, int signal, complex_t *dst,
                           int64_t count);

/* Diagnostics of the v2 location votes (src/computefourier-1.0-2.0.cc:
 * 125-184), which never change the output (every frequency they append is
 * already in the comb prefill, cf12 505-512) and are skipped by the exec.
 * With enabled != 0, later FP64 execs of this v2 plan also select the top
 * buckets of the location loops and count the votes on the GPU.
 * sfft_debug_comb_votes then writes the frequencies the reference appends to
 * its hit list, in the reference's order, for signal `signal` of the last
 * batch (at most cap of them) and returns how many there are, or -1. */
int sfft_plan_debug_comb_votes(sfft_plan *plan, int enabled);
int64_t sfft_debug_comb_votes(const sfft_plan *plan, int signal, int32_t *hits, int64_t cap);

/* Diagnostics of a v3 plan's later execs.  serial_subtractions != 0: every
 * subtraction pass (src/computefourier-3.0.cc:465-482, applied in map
 * order) by one thread per bucket array, the reference's serial order,
 * instead of the owner-thread paths used above 16 pairs; the results must
 * be bit-identical either way (the tests compare them).  map_limit > 0:
 * the map of located frequencies holds at most that many entries (below
 * sfft_plan_output_capacity), to exercise the overflow handling; 0 keeps
 * the plan's capacity.  Returns 0, or -1 for a bad or non-v3 plan. */
int sfft_plan_debug_v3(sfft_plan *plan, int serial_subtractions, int map_limit);

/* Diagnostic of a v3 plan's later execs: enabled != 0 makes each signal's
 * transform workgroup run its four pre tasks (Mansour and Gauss buckets)
 * itself instead of taking them from the workgroups that normally run
 * them ahead -- the path a transform takes when it finds a task not yet
 * started.  Results are bit-identical either way (the tests compare them).
 * Returns 0, or -1 for a bad or non-v3 plan. */
int sfft_plan_debug_v3_inline_pre(sfft_plan *plan, int enabled);

/* The plan's filters (which: 0 = location / first Gauss filter, 1 =
 * estimation / second Gauss filter): its w taps into time (capacity
 * time_cap) and its response at offsets -half..half into resp (capacity
 * resp_cap, resp[d + half] = response at frequency d mod n; the only
 * offsets the exec reads).  Returns w, or -1. */
int sfft_plan_filter(const sfft_plan *plan, int which, complex_t *time, int time_cap,
                     complex_t *resp, int resp_cap, int *half);

/* Host-side filter design of the library (src/filters.cc:70-160), for
 * tests: w taps into time (capacity time_cap) and the n-point response into
 * freq (may be NULL).  Returns w, or -1 when time_cap < w. */
int sfft_design_filter(double lobefrac, double tolerance, int n, int b,
                       complex_t *time, int time_cap, complex_t *freq);

#ifdef __cplusplus
}
#endif
#endif
