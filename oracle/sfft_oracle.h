/*
 * sFFT CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference sFFT (ETH/MIT sFFT 0.1) hot path,
 * used by tests/, the smoke check and the CPU-baseline leg of bench.py as the
 * checker.  The product library (sfft_amd) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Reference
 * quirks are restated on purpose (they define what "parity" means):
 *   - k > 50 selects the K-parameter table with n instead of k
 *     (src/sfft.cc:316-321), so the hard-coded defaults apply;
 *   - estimate_values computes (ac+bd, ad-bc)/|G|^2, i.e. the conjugate of
 *     bucket/G, for v1/v2 (src/computefourier-1.0-2.0.cc:383-398);
 *   - the v2 located set is the comb prefill (cf12 505-512);
 *   - v3 returns an all-zero output when the first Mansour pass finds
 *     exactly k frequencies (src/computefourier-3.0.cc:859-860).
 */
#ifndef SFFT_ORACLE_H
#define SFFT_ORACLE_H

#include <complex.h>
#include <stdint.h>

typedef double complex or_cplx;

/* ---- numeric utilities (src/utils.cc) ---------------------------------- */
int or_gcd(int a, int b);                         /* utils.cc:41-46   */
int or_mod_inverse(int a, int n);                 /* utils.cc:85-101  */
int or_floor_to_pow2(double x);                   /* utils.cc:243-248 */
void or_find_largest_indices(int *out, int num, const double *samples,
                             int n);              /* utils.cc:130-158 */
void or_sort_ints(int *a, int size);              /* utils.cc:181-196 */

/* ---- FP64 DFT standing in for FFTW (src/fftw.cc:29-59) ------------------ */
/* out = sum_j in[j] e^{-+2 pi i jk/n}, unnormalised; out may alias in. */
void or_fft(or_cplx *out, const or_cplx *in, int n, int backward);

/* ---- filters (src/filters.cc) ------------------------------------------ */
int or_dolph_width(double lobefrac, double tolerance);   /* filters.cc:72-74 */
void or_make_dolphchebyshev_t(double lobefrac, double tolerance, int w,
                              or_cplx *x);               /* filters.cc:70-86 */
/* x: w window taps in, w filter taps out; freq: n outputs (filters.cc:109-160) */
void or_make_multiple_t(or_cplx *x, int w, int n, int b, or_cplx *freq);

/* ---- v1/v2 plan parameters (src/sfft.cc:298-372, src/parameters.cc) ---- */
typedef struct {
  int n, k, with_comb;
  double Bcst_loc, Bcst_est, Comb_cst, tol_loc, tol_est;
  int loc_loops, est_loops, thr_loops, comb_loops;
  double BB_loc, BB_est, lobefrac_loc, lobefrac_est;
  int b_loc, b_est;
  int B_loc, B_est, B_thresh, W_comb;
  int w_loc, w_est;
  long x_samp_size;
} or_v12_params;

void or_get_N_params(int N, int with_comb, double *Bcst_loc, double *Bcst_est,
                     double *Comb_cst, int *loc_loops, int *est_loops,
                     int *thr_loops, int *comb_loops, double *tol_loc,
                     double *tol_est);           /* parameters.cc:24-312 */
void or_get_K_params(int K, int with_comb, double *Bcst_loc, double *Bcst_est,
                     double *Comb_cst, int *loc_loops, int *est_loops,
                     int *thr_loops, int *comb_loops, double *tol_loc,
                     double *tol_est);           /* parameters.cc:314-513 */
int or_v12_params_get(int n, int k, int with_comb, or_v12_params *p);

typedef struct {
  or_v12_params p;
  or_cplx *loc_time, *loc_freq;   /* w_loc taps, n response */
  or_cplx *est_time, *est_freq;   /* w_est taps, n response */
} or_v12_plan;

or_v12_plan *or_v12_plan_make(int n, int k, int with_comb);
void or_v12_plan_free(or_v12_plan *P);
/* Replace the plan's filters by given ones (w_loc/w_est taps, n-point
 * responses), so that a test can run the oracle with another design's
 * filters and compare everything downstream bit for bit. */
void or_v12_plan_set_filters(or_v12_plan *P, const or_cplx *loc_time,
                             const or_cplx *loc_freq, const or_cplx *est_time,
                             const or_cplx *est_freq);

/* RNG draws of one v1/v2 exec, in reference order (cf12 465-474, 62):
 * loops x (a, ai) from random(), then comb_loops offsets from drand48(). */
void or_v12_draw(int n, int loops, int comb_loops, int W_comb, int *a,
                 int *ai, int *comb_off);

/* Optional intermediate results of one exec (any pointer may be NULL). */
typedef struct {
  or_cplx *x_sampt;      /* x_samp_size folded buckets before the FFTs     */
  or_cplx *x_samp;       /* x_samp_size buckets after the FFTs             */
  int *J;                /* loops * B_thresh selected buckets              */
  int *comb_approved;    /* comb_loops * B_thresh (deduped prefix)         */
  int num_comb;          /* out                                            */
  int *hits;             /* hits list, capacity hits_cap                   */
  long hits_cap;
  long hits_found;       /* out: total (prefill + appended)                */
  long hits_prefill;     /* out: prefill entries (v2)                      */
} or_v12_trace;

/* One sfft_exec of v1/v2 (src/sfft.cc:119-137 -> cf12 437-541).
 * out: n values, zero-filled here like sfft_exec. Returns 0 on success. */
int or_v12_exec(const or_v12_plan *P, const or_cplx *x, const int *a,
                const int *ai, const int *comb_off, or_cplx *out,
                or_v12_trace *tr);

/* sfft_exec_many (src/sfft.cc:139-147): OpenMP over signals; RNG draws are
 * passed in (num x loops, num x comb_loops) so the order is sequential. */
int or_v12_exec_many(const or_v12_plan *P, int num, const or_cplx *const *xs,
                     const int *a, const int *ai, const int *comb_off,
                     or_cplx *const *outs, int threads);

/* ---- v3 (src/sfft.cc:506-579, src/computefourier-3.0.cc) --------------- */
typedef struct {
  int n, k;
  int W_Man, B_g1, B_g2, w_g1, w_g2, b_g1, b_g2;
  double BB1, BB2, lobefrac_g1, lobefrac_g2, tol_g1, tol_g2;
} or_v3_params;

int or_v3_params_get(int n, int k, or_v3_params *p);

typedef struct {
  or_v3_params p;
  or_cplx *time1, *freq1;   /* w_g1 taps, n response */
  or_cplx *time2, *freq2;   /* w_g2 taps, n response */
} or_v3_plan;

or_v3_plan *or_v3_plan_make(int n, int k);
or_v3_plan *or_v3_plan_make_bare(int n, int k);  /* filters from set_filters */
void or_v3_plan_free(or_v3_plan *P);
/* Replace the plan's two filters (w_g1/w_g2 taps, n-point responses). */
void or_v3_plan_set_filters(or_v3_plan *P, const or_cplx *time1,
                            const or_cplx *freq1, const or_cplx *time2,
                            const or_cplx *freq2);

/* RNG draws of one v3 exec (cf3 800-810): a, ai, b, shift, offsets. */
void or_v3_draw(int n, int *draws6);

/* One sfft_exec of v3.  out: n values (zero-filled here).  Returns the
 * number of map entries; *early (if non-NULL) is set to 1 when the
 * reference's early return (cf3 859-860) left out all zero.
 * freqs/vals (capacity cap, may be NULL) receive the map in the order the
 * oracle's map iterates (ascending frequency). */
long or_v3_exec(const or_v3_plan *P, const or_cplx *x, const int *draws6,
                or_cplx *out, int *early, int *freqs, or_cplx *vals,
                long cap);

#endif
