/*
 * sFFT CPU oracle -- TEST INFRASTRUCTURE ONLY (see sfft_oracle.h).
 *
 * Plain C99 restatement of the reference sFFT hot path.  Only tests/, the
 * smoke check and bench.py's cpu_baseline leg use it, as the checker.
 * Compiled with -ffp-contract=off so that every product of the reference's
 * SSE code is rounded on its own, as _mm_mul_pd does.
 *
 * "cf12" is src/computefourier-1.0-2.0.cc and "cf3" is
 * src/computefourier-3.0.cc in the reference tree.
 */
#define _GNU_SOURCE
#include "sfft_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* numeric utilities                                                         */
/* ------------------------------------------------------------------------ */

/* src/utils.cc:41-46 (recursive Euclid; gcd(0,n) = n). */
int or_gcd(int a, int b)
{
  while (a % b != 0) {
    int t = a % b;
    a = b;
    b = t;
  }
  return b;
}

/* src/utils.cc:85-101 (extended Euclid, result in [0,n)). */
int or_mod_inverse(int a, int n)
{
  int i = n, v = 0, d = 1;
  while (a > 0) {
    int t = i / a, x = a;
    a = i % x;
    i = x;
    x = d;
    d = v - t * x;
    v = x;
  }
  v %= n;
  if (v < 0) v = (v + n) % n;
  return v;
}

/* src/utils.cc:243-248: largest power of two <= x, 0 when x < 1. */
int or_floor_to_pow2(double x)
{
  unsigned int ans;
  for (ans = 1; ans <= x; ans <<= 1) {
  }
  return (int)(ans / 2);
}

static int cmp_int(const void *a, const void *b)
{
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* src/utils.cc:181-196 sorts non-negative ints ascending (LSD radix). */
void or_sort_ints(int *a, int size) { qsort(a, (size_t)size, sizeof(int), cmp_int); }

/* Value of rank r (0-based, ascending) of v[0..n), v is scrambled.
 * Any exact selection returns what std::nth_element leaves at v+r. */
static double select_rank(double *v, int n, int r)
{
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    double piv = v[lo + (hi - lo) / 2];
    int i = lo, j = hi;
    while (i <= j) {
      while (v[i] < piv) i++;
      while (v[j] > piv) j--;
      if (i <= j) {
        double t = v[i];
        v[i] = v[j];
        v[j] = t;
        i++;
        j--;
      }
    }
    if (r <= j) hi = j;
    else if (r >= i) lo = i;
    else return v[r];
  }
  return v[r];
}

/* src/utils.cc:130-158: the num largest samples' indices, ascending.
 * cutoff = (num+1)-th largest; take > cutoff in index order, then
 * == cutoff in index order until num, then sort. */
void or_find_largest_indices(int *out, int num, const double *samples, int n)
{
  double *tmp = (double *)malloc(sizeof(double) * (size_t)n);
  memcpy(tmp, samples, sizeof(double) * (size_t)n);
  double cutoff = select_rank(tmp, n, n - num - 1);
  free(tmp);
  int count = 0;
  for (int i = 0; i < n; i++)
    if (samples[i] > cutoff) out[count++] = i;
  if (count < num) {
    for (int i = 0; i < n; i++) {
      if (samples[i] == cutoff) {
        out[count++] = i;
        if (count >= num) break;
      }
    }
    or_sort_ints(out, count);
  }
}

/* ------------------------------------------------------------------------ */
/* FP64 DFT (stands in for FFTW's unnormalised transforms, src/fftw.cc)     */
/* ------------------------------------------------------------------------ */

static void fft_pow2_inplace(or_cplx *x, int n, int backward)
{
  /* bit reversal */
  for (int i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      or_cplx t = x[i];
      x[i] = x[j];
      x[j] = t;
    }
  }
  if (n < 2) return;
  /* twiddles for the full size; stage of length len uses stride n/len */
  int half = n / 2;
  double *wr = (double *)malloc(sizeof(double) * (size_t)half);
  double *wi = (double *)malloc(sizeof(double) * (size_t)half);
  double sgn = backward ? 1.0 : -1.0;
  for (int k = 0; k < half; k++) {
    /* sincos() explicitly: compilers fuse sin+cos into it or not, and the
     * two differ in the last bit for rare arguments */
    double ang = 2.0 * M_PI * (double)k / (double)n, s, c;
    sincos(ang, &s, &c);
    wr[k] = c;
    wi[k] = sgn * s;
  }
  for (int len = 2; len <= n; len <<= 1) {
    int h = len >> 1, stride = n / len;
    for (int s = 0; s < n; s += len) {
      for (int j = 0; j < h; j++) {
        double cr = wr[j * stride], ci = wi[j * stride];
        double ur = creal(x[s + j]), ui = cimag(x[s + j]);
        double vr0 = creal(x[s + j + h]), vi0 = cimag(x[s + j + h]);
        double vr = vr0 * cr - vi0 * ci, vi = vr0 * ci + vi0 * cr;
        x[s + j] = CMPLX(ur + vr, ui + vi);
        x[s + j + h] = CMPLX(ur - vr, ui - vi);
      }
    }
  }
  free(wr);
  free(wi);
}

/* Bluestein chirp-z for sizes that are not powers of two. */
static void fft_bluestein(or_cplx *out, const or_cplx *in, int n, int backward)
{
  int m = 1;
  while (m < 2 * n - 1) m <<= 1;
  or_cplx *c = (or_cplx *)malloc(sizeof(or_cplx) * (size_t)n);
  or_cplx *A = (or_cplx *)calloc((size_t)m, sizeof(or_cplx));
  or_cplx *Bv = (or_cplx *)calloc((size_t)m, sizeof(or_cplx));
  double sgn = backward ? 1.0 : -1.0;
  long long two_n = 2LL * n;
  for (int j = 0; j < n; j++) {
    long long q = ((long long)j * j) % two_n;     /* exact j^2 mod 2n */
    double ang = M_PI * (double)q / (double)n;
    c[j] = CMPLX(cos(ang), sgn * sin(ang));        /* e^{-+ i pi j^2/n} */
  }
  for (int j = 0; j < n; j++) A[j] = in[j] * c[j];
  Bv[0] = conj(c[0]);
  for (int j = 1; j < n; j++) Bv[j] = Bv[m - j] = conj(c[j]);
  fft_pow2_inplace(A, m, 0);
  fft_pow2_inplace(Bv, m, 0);
  for (int j = 0; j < m; j++) A[j] *= Bv[j];
  fft_pow2_inplace(A, m, 1);
  double inv = 1.0 / (double)m;
  for (int k = 0; k < n; k++) out[k] = c[k] * A[k] * inv;
  free(c);
  free(A);
  free(Bv);
}

void or_fft(or_cplx *out, const or_cplx *in, int n, int backward)
{
  if (n <= 0) return;
  if ((n & (n - 1)) == 0) {
    if (out != in) memcpy(out, in, sizeof(or_cplx) * (size_t)n);
    fft_pow2_inplace(out, n, backward);
  } else {
    or_cplx *tmp = (or_cplx *)malloc(sizeof(or_cplx) * (size_t)n);
    fft_bluestein(tmp, in, n, backward);
    memcpy(out, tmp, sizeof(or_cplx) * (size_t)n);
    free(tmp);
  }
}

/* ------------------------------------------------------------------------ */
/* filters                                                                   */
/* ------------------------------------------------------------------------ */

/* src/filters.cc:62-68 */
static double cheb(double m, double x)
{
  if (fabs(x) <= 1) return cos(m * acos(x));
  return creal(ccosh(m * cacosh(x)));
}

/* src/filters.cc:72-74 */
int or_dolph_width(double lobefrac, double tolerance)
{
  int w = (int)((1 / M_PI) * (1 / lobefrac) * acosh(1. / tolerance));
  if (!(w % 2)) w--;
  return w;
}

/* src/filters.cc:70-86 with shift() from src/utils.cc:28-37. */
void or_make_dolphchebyshev_t(double lobefrac, double tolerance, int w,
                              or_cplx *x)
{
  (void)lobefrac;
  double t0 = cosh(acosh(1 / tolerance) / (w - 1));
  for (int i = 0; i < w; i++)
    x[i] = cheb(w - 1, t0 * cos(M_PI * i / w)) * tolerance;
  or_fft(x, x, w, 0);
  /* shift(x, w, w/2): x[i] <- x[i - w/2] cyclically */
  int r = w / 2;
  or_cplx *tmp = (or_cplx *)malloc(sizeof(or_cplx) * (size_t)w);
  for (int i = 0; i < w; i++) tmp[(i + r) % w] = x[i];
  for (int i = 0; i < w; i++) x[i] = creal(tmp[i]);
  free(tmp);
}

/* src/filters.cc:109-160 */
void or_make_multiple_t(or_cplx *x, int w, int n, int b, or_cplx *h)
{
  or_cplx *g = (or_cplx *)calloc((size_t)n, sizeof(or_cplx));
  memcpy(g, x + w / 2, sizeof(or_cplx) * (size_t)(w - (w / 2)));
  memcpy(g + n - w / 2, x, sizeof(or_cplx) * (size_t)(w / 2));
  or_fft(g, g, n, 0);
  or_cplx s = 0;
  for (int i = 0; i < b; i++) s += g[i];
  double max = 0;
  int offset = b / 2;
  for (int i = 0; i < n; i++) {
    h[(i + n + offset) % n] = s;
    double a = cabs(s);
    if (a > max) max = a;
    s = s + (g[(i + b) % n] - g[i]);
  }
  for (int i = 0; i < n; i++) h[i] /= max;
  or_cplx offsetc = 1;
  double th = -2 * M_PI * (double)(w / 2) / (double)n;
  or_cplx step = CMPLX(cos(th), sin(th));
  for (int i = 0; i < n; i++) {
    h[i] *= offsetc;
    offsetc *= step;
  }
  or_fft(g, h, n, 1);
  memmove(x, g, sizeof(or_cplx) * (size_t)w);
  free(g);
  for (int i = 0; i < w; i++) x[i] /= n;
}

/* ------------------------------------------------------------------------ */
/* parameter tables (src/parameters.cc)                                      */
/* ------------------------------------------------------------------------ */

typedef struct {
  int key;
  double bl, be, cc;
  int cl, el, ll, tl;
  double tlc, tes;
} prow;

#define ROW(K, BL, BE, CC, CL, EL, LL, TL, TOLL, TOLE) \
  { K, BL, BE, CC, CL, EL, LL, TL, TOLL, TOLE }

/* parameters.cc:33-157 (WITH_COMB) and 159-309 (without). Columns:
 * N, Bcst_loc, Bcst_est, Comb_cst, comb_loops, est_loops, loc_loops,
 * threshold_loops, tolerance_loc, tolerance_est. */
static const prow N_COMB[] = {
  ROW(8192, 2, 2, 32, 8, 16, 7, 6, 1e-8, 1e-8),
  ROW(16384, 4, 4, 32, 8, 10, 6, 5, 1e-8, 1e-8),
  ROW(32768, 4, 2, 64, 4, 8, 5, 4, 1e-8, 1e-8),
  ROW(65536, 4, 2, 128, 6, 10, 4, 2, 1e-8, 1e-8),
  ROW(131072, 1, 1, 8, 2, 12, 4, 3, 1e-6, 1e-8),
  ROW(262144, 1, 1, 8, 2, 14, 5, 4, 1e-6, 1e-8),
  ROW(524288, 0.5, 0.5, 8, 1, 10, 4, 3, 1e-6, 1e-8),
  ROW(1048576, 0.5, 0.5, 8, 2, 12, 4, 2, 1e-6, 1e-8),
  ROW(2097152, 0.5, 0.2, 8, 1, 10, 3, 2, 1e-6, 1e-8),
  ROW(4194304, 0.5, 0.2, 8, 1, 8, 3, 2, 1e-6, 1e-8),
  ROW(8388608, 0.5, 0.2, 8, 1, 8, 3, 2, 1e-6, 1e-8),
  ROW(16777216, 0.5, 0.2, 16, 1, 8, 3, 2, 1e-6, 1e-8),
};
static const prow N_PLAIN[] = {
  ROW(8192, 2, 2, 1, 1, 16, 7, 6, 1e-8, 1e-8),
  ROW(16384, 4, 4, 1, 1, 10, 6, 5, 1e-8, 1e-8),
  ROW(32768, 4, 2, 1, 1, 8, 5, 4, 1e-8, 1e-8),
  ROW(65536, 4, 2, 1, 1, 8, 5, 4, 1e-8, 1e-8),
  ROW(131072, 2, 1, 1, 1, 10, 5, 4, 1e-6, 1e-8),
  ROW(262144, 2, 0.5, 1, 1, 14, 4, 3, 1e-6, 1e-8),
  ROW(524288, 1, 0.5, 1, 1, 12, 5, 4, 1e-6, 1e-8),
  ROW(1048576, 2, 0.5, 1, 1, 12, 4, 3, 1e-6, 1e-8),
  ROW(2097152, 2, 0.2, 1, 1, 15, 3, 2, 1e-6, 1e-8),
  ROW(4194304, 4, 0.2, 1, 1, 10, 3, 2, 1e-6, 1e-8),
  ROW(8388608, 2, 0.2, 1, 1, 8, 3, 2, 1e-6, 1e-8),
  ROW(16777216, 4, 0.2, 1, 1, 8, 3, 2, 1e-6, 1e-8),
};
/* parameters.cc:323-412 (WITH_COMB) and 414-510 (without), keyed by K. */
static const prow K_COMB[] = {
  ROW(50, 0.5, 0.2, 16, 1, 10, 3, 2, 1e-6, 1.0e-8),
  ROW(100, 0.5, 0.2, 16, 1, 12, 4, 2, 1e-6, 1.0e-8),
  ROW(200, 0.5, 0.5, 32, 1, 8, 4, 3, 1e-6, 0.5e-8),
  ROW(500, 0.5, 0.5, 64, 1, 10, 4, 3, 1e-6, 0.5e-8),
  ROW(1000, 1, 1, 128, 3, 12, 4, 3, 1e-6, 0.5e-8),
  ROW(2000, 1, 1, 512, 3, 16, 4, 3, 1e-7, 0.2e-8),
  ROW(2500, 1, 1, 512, 3, 16, 4, 3, 1e-7, 0.2e-8),
  ROW(4000, 1, 2, 512, 3, 14, 8, 7, 1e-8, 0.5e-8),
};
static const prow K_PLAIN[] = {
  ROW(50, 4, 0.2, 1, 1, 10, 3, 2, 1e-6, 1.0e-8),
  ROW(100, 2, 0.2, 1, 1, 12, 3, 2, 1e-6, 1.0e-8),
  ROW(200, 4, 0.5, 1, 1, 10, 3, 2, 1e-6, 0.5e-8),
  ROW(500, 2, 1, 1, 1, 12, 4, 3, 1e-6, 0.5e-8),
  ROW(1000, 2, 1, 1, 1, 12, 5, 4, 1e-6, 1.0e-8),
  ROW(2000, 2, 1, 1, 1, 16, 5, 4, 1e-7, 0.5e-8),
  ROW(2500, 2, 1, 1, 1, 16, 5, 4, 1e-7, 0.5e-8),
  ROW(4000, 2, 2, 1, 1, 14, 6, 5, 1e-8, 1.0e-8),
};

static void apply_row(const prow *tab, int len, int key, double *Bcst_loc,
                      double *Bcst_est, double *Comb_cst, int *loc_loops,
                      int *est_loops, int *thr_loops, int *comb_loops,
                      double *tol_loc, double *tol_est)
{
  for (int i = 0; i < len; i++) {
    if (tab[i].key != key) continue;
    *Bcst_loc = tab[i].bl;
    *Bcst_est = tab[i].be;
    *Comb_cst = tab[i].cc;
    *comb_loops = tab[i].cl;
    *est_loops = tab[i].el;
    *loc_loops = tab[i].ll;
    *thr_loops = tab[i].tl;
    *tol_loc = tab[i].tlc;
    *tol_est = tab[i].tes;
    return;
  }
}

void or_get_N_params(int N, int with_comb, double *Bcst_loc, double *Bcst_est,
                     double *Comb_cst, int *loc_loops, int *est_loops,
                     int *thr_loops, int *comb_loops, double *tol_loc,
                     double *tol_est)
{
  if (with_comb)
    apply_row(N_COMB, 12, N, Bcst_loc, Bcst_est, Comb_cst, loc_loops, est_loops,
              thr_loops, comb_loops, tol_loc, tol_est);
  else
    apply_row(N_PLAIN, 12, N, Bcst_loc, Bcst_est, Comb_cst, loc_loops,
              est_loops, thr_loops, comb_loops, tol_loc, tol_est);
}

void or_get_K_params(int K, int with_comb, double *Bcst_loc, double *Bcst_est,
                     double *Comb_cst, int *loc_loops, int *est_loops,
                     int *thr_loops, int *comb_loops, double *tol_loc,
                     double *tol_est)
{
  if (with_comb)
    apply_row(K_COMB, 8, K, Bcst_loc, Bcst_est, Comb_cst, loc_loops, est_loops,
              thr_loops, comb_loops, tol_loc, tol_est);
  else
    apply_row(K_PLAIN, 8, K, Bcst_loc, Bcst_est, Comb_cst, loc_loops, est_loops,
              thr_loops, comb_loops, tol_loc, tol_est);
}

/* src/sfft.cc:298-372 (sizes only; filters in or_v12_plan_make). */
int or_v12_params_get(int n_in, int k_in, int with_comb, or_v12_params *p)
{
  memset(p, 0, sizeof(*p));
  unsigned n = (unsigned)n_in, k = (unsigned)k_in;
  p->Bcst_loc = 1;
  p->Bcst_est = 1;
  p->Comb_cst = 2;
  p->loc_loops = 4;
  p->est_loops = 16;
  p->thr_loops = 3;
  p->comb_loops = 1;
  p->tol_loc = 1.e-8;
  p->tol_est = 1.e-8;
  /* sfft.cc:316-327: the K table is called with n (quirk). */
  if (k > 50)
    or_get_K_params((int)n, with_comb, &p->Bcst_loc, &p->Bcst_est, &p->Comb_cst,
                    &p->loc_loops, &p->est_loops, &p->thr_loops,
                    &p->comb_loops, &p->tol_loc, &p->tol_est);
  else
    or_get_N_params((int)n, with_comb, &p->Bcst_loc, &p->Bcst_est, &p->Comb_cst,
                    &p->loc_loops, &p->est_loops, &p->thr_loops,
                    &p->comb_loops, &p->tol_loc, &p->tol_est);
  n = (unsigned)or_floor_to_pow2(n);
  p->n = (int)n;
  p->k = (int)k;
  p->with_comb = with_comb;
  p->BB_loc = (unsigned)(p->Bcst_loc * sqrt((double)(int)n * k / (log2(n))));
  p->BB_est = (unsigned)(p->Bcst_est * sqrt((double)(int)n * k / (log2(n))));
  p->lobefrac_loc = 0.5 / (p->BB_loc);
  p->lobefrac_est = 0.5 / (p->BB_est);
  p->b_loc = (int)(1.2 * 1.1 * ((double)n / p->BB_loc));
  p->b_est = (int)(1.4 * 1.1 * ((double)n / p->BB_est));
  p->B_loc = or_floor_to_pow2(p->BB_loc);
  p->B_thresh = 2 * (int)k;
  p->B_est = or_floor_to_pow2(p->BB_est);
  p->W_comb = or_floor_to_pow2(p->Comb_cst * n / p->B_loc);
  p->w_loc = or_dolph_width(p->lobefrac_loc, p->tol_loc);
  p->w_est = or_dolph_width(p->lobefrac_est, p->tol_est);
  p->x_samp_size = (long)p->loc_loops * p->B_loc + (long)p->est_loops * p->B_est;
  return 0;
}

or_v12_plan *or_v12_plan_make(int n, int k, int with_comb)
{
  or_v12_plan *P = (or_v12_plan *)calloc(1, sizeof(*P));
  or_v12_params_get(n, k, with_comb, &P->p);
  int N = P->p.n;
  P->loc_time = (or_cplx *)malloc(sizeof(or_cplx) * (size_t)P->p.w_loc);
  P->loc_freq = (or_cplx *)malloc(sizeof(or_cplx) * (size_t)N);
  or_make_dolphchebyshev_t(P->p.lobefrac_loc, P->p.tol_loc, P->p.w_loc,
                           P->loc_time);
  or_make_multiple_t(P->loc_time, P->p.w_loc, N, P->p.b_loc, P->loc_freq);
  P->est_time = (or_cplx *)malloc(sizeof(or_cplx) * (size_t)P->p.w_est);
  P->est_freq = (or_cplx *)malloc(sizeof(or_cplx) * (size_t)N);
  or_make_dolphchebyshev_t(P->p.lobefrac_est, P->p.tol_est, P->p.w_est,
                           P->est_time);
  or_make_multiple_t(P->est_time, P->p.w_est, N, P->p.b_est, P->est_freq);
  return P;
}

void or_v12_plan_set_filters(or_v12_plan *P, const or_cplx *loc_time,
                             const or_cplx *loc_freq, const or_cplx *est_time,
                             const or_cplx *est_freq)
{
  size_t N = (size_t)P->p.n;
  memcpy(P->loc_time, loc_time, sizeof(or_cplx) * (size_t)P->p.w_loc);
  memcpy(P->loc_freq, loc_freq, sizeof(or_cplx) * N);
  memcpy(P->est_time, est_time, sizeof(or_cplx) * (size_t)P->p.w_est);
  memcpy(P->est_freq, est_freq, sizeof(or_cplx) * N);
}

void or_v12_plan_free(or_v12_plan *P)
{
  if (!P) return;
  free(P->loc_time);
  free(P->loc_freq);
  free(P->est_time);
  free(P->est_freq);
  free(P);
}

/* cf12 465-474 and cf12 62. */
void or_v12_draw(int n, int loops, int comb_loops, int W_comb, int *a,
                 int *ai, int *comb_off)
{
  for (int i = 0; i < loops; i++) {
    a[i] = 0;
    while (or_gcd(a[i], n) != 1) a[i] = (int)(random() % n);
    ai[i] = or_mod_inverse(a[i], n);
  }
  int sigma = n / W_comb;
  for (int i = 0; i < comb_loops; i++)
    comb_off[i] = (int)(unsigned)floor(drand48() * sigma);
}

/* ------------------------------------------------------------------------ */
/* v1 / v2 exec                                                              */
/* ------------------------------------------------------------------------ */

static inline int timesmod(int x, int a, int n)
{
  return (int)(((long long)x * a) % n);
}

/* cf12 48-82: comb sub-sampling, W-point FFT, |.|^2, top num. */
static void comb_filt(const or_cplx *x, int n, int num, int W, int offset,
                      int *approved)
{
  int sigma = n / W;
  or_cplx *s = (or_cplx *)malloc(sizeof(or_cplx) * (size_t)W);
  double *mag = (double *)malloc(sizeof(double) * (size_t)W);
  for (int i = 0; i < W; i++) s[i] = x[offset + i * sigma];
  or_fft(s, s, W, 0);
  for (int i = 0; i < W; i++) {
    double re = creal(s[i]), im = cimag(s[i]);
    mag[i] = re * re + im * im;
  }
  or_find_largest_indices(approved, num, mag, W);
  free(s);
  free(mag);
}

/* cf12 91-116 */
static void filter_regular(const int *J, int n, int num, int B, int a,
                           int thr, int *score, int *hits, long *hits_found,
                           long cap)
{
  for (int i = 0; i < num; i++) {
    int low = ((int)ceil((J[i] - 0.5) * n / B) + n) % n;
    int high = ((int)ceil((J[i] + 0.5) * n / B) + n) % n;
    int loc = timesmod(low, a, n);
    for (int j = low; j != high; j = (j + 1) % n) {
      score[loc]++;
      if (score[loc] == thr) {
        if (*hits_found < cap) hits[*hits_found] = loc;
        (*hits_found)++;
      }
      loc = (loc + a) % n;
    }
  }
}

typedef struct {
  int first, second;
} ipair;

static int cmp_pair(const void *x, const void *y)
{
  const ipair *p = (const ipair *)x, *q = (const ipair *)y;
  if (p->first != q->first) return (p->first > q->first) - (p->first < q->first);
  return (p->second > q->second) - (p->second < q->second);
}

/* cf12 125-184 */
static void filter_comb(const int *J, int n, int num, int B, int a, int ai,
                        int thr, int *score, int *hits, long *hits_found,
                        long cap, const int *approved, int num_comb, int W)
{
  ipair *pa = (ipair *)malloc(sizeof(ipair) * (size_t)num_comb);
  for (int m = 0; m < num_comb; m++) {
    int prev = timesmod(approved[m], ai, W);
    pa[m].first = prev;
    pa[m].second = timesmod(prev, a, n);
  }
  qsort(pa, (size_t)num_comb, sizeof(ipair), cmp_pair);
  for (int i = 0; i < num; i++) {
    int low = ((int)ceil((J[i] - 0.5) * n / B) + n) % n;
    int high = ((int)ceil((J[i] + 0.5) * n / B) + n) % n;
    /* upper_bound(make_pair(low % W, -1)) */
    int key = low % W, index = 0;
    while (index < num_comb &&
           (pa[index].first < key || (pa[index].first == key &&
                                      pa[index].second <= -1)))
      index++;
    int location = low - (low % W);
    int locinv = timesmod(location, a, n);
    for (int j = index;; j++) {
      if (j == num_comb) {
        j -= num_comb;
        location = (location + W) % n;
        locinv = timesmod(location, a, n);
      }
      int approved_loc = location + pa[j].first;
      if ((low < high && (approved_loc >= high || approved_loc < low)) ||
          (low > high && (approved_loc >= high && approved_loc < low)))
        break;
      int loc = (locinv + pa[j].second) % n;
      score[loc]++;
      if (score[loc] == thr) {
        if (*hits_found < cap) hits[*hits_found] = loc;
        (*hits_found)++;
      }
    }
  }
  free(pa);
}

int or_v12_exec(const or_v12_plan *P, const or_cplx *x, const int *a,
                const int *ai, const int *comb_off, or_cplx *out,
                or_v12_trace *tr)
{
  const or_v12_params *p = &P->p;
  const int n = p->n, L = p->loc_loops + p->est_loops, Lloc = p->loc_loops;
  const int num = p->B_thresh, W = p->W_comb, thr = p->thr_loops;
  const long xs = p->x_samp_size;

  memset(out, 0, sizeof(or_cplx) * (size_t)n); /* sfft.cc:121-123 */

  int *approved = (int *)malloc(sizeof(int) * (size_t)(p->comb_loops * num + 1));
  int num_comb = num;
  /* hits can hold the prefill plus every appended hit */
  long cap = 2L * n + 16;
  int *hits = (int *)malloc(sizeof(int) * (size_t)cap);
  long hits_found = 0, prefill = 0;

  if (p->with_comb) { /* cf12 483-512 */
    for (int i = 0; i < p->comb_loops; i++)
      comb_filt(x, n, num, W, comb_off[i], approved + i * num);
    if (p->comb_loops > 1) {
      or_sort_ints(approved, p->comb_loops * num);
      int last = 0;
      for (int i = 1; i < p->comb_loops * num; i++)
        if (approved[i] != approved[last]) approved[++last] = approved[i];
      num_comb = last + 1;
    }
    prefill = (long)num_comb * (n / W);
    for (int j = 0; j < n / W; j++)
      for (int i = 0; i < num_comb; i++)
        hits[(long)j * num_comb + i] = j * W + approved[i];
    hits_found = prefill;
  }

  /* inner_loop_locate, cf12 195-327 */
  or_cplx *xt = (or_cplx *)calloc((size_t)xs, sizeof(or_cplx));
  or_cplx *xf = (or_cplx *)malloc(sizeof(or_cplx) * (size_t)xs);
  for (int j = 0; j < L; j++) {
    int loc = j < Lloc;
    const or_cplx *tap = loc ? P->loc_time : P->est_time;
    int w = loc ? p->w_loc : p->w_est;
    int B = loc ? p->B_loc : p->B_est;
    long off = (long)(j < Lloc ? j : Lloc) * p->B_loc +
               (long)(j - Lloc > 0 ? j - Lloc : 0) * p->B_est;
    double *acc = (double *)(xt + off);
    unsigned idx = 0; /* b = 0 */
    for (int i = 0; i < w; i++) {
      double xr = creal(x[idx]), xi = cimag(x[idx]);
      double c = creal(tap[i]), d = cimag(tap[i]);
      double re = xr * c - xi * d;  /* hadd(ac, -bd) */
      double im = xr * d + xi * c;  /* hadd(ad, bc)  */
      int s = i & (B - 1);
      acc[2 * s] = acc[2 * s] + re;
      acc[2 * s + 1] = acc[2 * s + 1] + im;
      idx = (idx + (unsigned)ai[j]) & (unsigned)(n - 1);
    }
  }
  if (tr && tr->x_sampt) memcpy(tr->x_sampt, xt, sizeof(or_cplx) * (size_t)xs);
  for (int j = 0; j < L; j++) {
    int B = j < Lloc ? p->B_loc : p->B_est;
    long off = (long)(j < Lloc ? j : Lloc) * p->B_loc +
               (long)(j - Lloc > 0 ? j - Lloc : 0) * p->B_est;
    or_fft(xf + off, xt + off, B, 0);
  }
  if (tr && tr->x_samp) memcpy(tr->x_samp, xf, sizeof(or_cplx) * (size_t)xs);

  double *samples = (double *)malloc(sizeof(double) * (size_t)xs);
  for (long j = 0; j < xs; j++) {
    double re = creal(xf[j]), im = cimag(xf[j]);
    samples[j] = re * re + im * im;
  }
  int *score = (int *)calloc((size_t)n, sizeof(int));
  int *J = (int *)malloc(sizeof(int) * (size_t)num);
  for (int j = 0; j < L; j++) {
    int B = j < Lloc ? p->B_loc : p->B_est;
    long off = (long)(j < Lloc ? j : Lloc) * p->B_loc +
               (long)(j - Lloc > 0 ? j - Lloc : 0) * p->B_est;
    or_find_largest_indices(J, num, samples + off, B);
    if (tr && tr->J) memcpy(tr->J + (long)j * num, J, sizeof(int) * (size_t)num);
    if (j < Lloc) {
      if (!p->with_comb)
        filter_regular(J, n, num, B, a[j], thr, score, hits, &hits_found, cap);
      else
        filter_comb(J, n, num, B, a[j], ai[j], thr, score, hits, &hits_found,
                    cap, approved, num_comb, W);
    }
  }
  free(score);
  free(J);
  free(samples);

  /* estimate_values, cf12 340-419 (the assert at 352 treated as NDEBUG) */
  double *vr = (double *)malloc(sizeof(double) * (size_t)L);
  double *vi = (double *)malloc(sizeof(double) * (size_t)L);
  long used = hits_found < cap ? hits_found : cap;
  for (long h = 0; h < used; h++) {
    for (int j = 0; j < L; j++) {
      long off = (long)(j < Lloc ? j : Lloc) * p->B_loc +
                 (long)(j - Lloc > 0 ? j - Lloc : 0) * p->B_est;
      int cur_B = j < Lloc ? p->B_loc : p->B_est;
      const or_cplx *F = j < Lloc ? P->loc_freq : P->est_freq;
      int m = n / cur_B;
      int pi = timesmod(ai[j], hits[h], n);
      int hashed_to = pi / m;
      int dist = pi % m;
      if (dist > m / 2) {
        hashed_to = (hashed_to + 1) % cur_B;
        dist -= m;
      }
      dist = (n - dist) % n;
      double A = creal(xf[off + hashed_to]), Bi = cimag(xf[off + hashed_to]);
      double c = creal(F[dist]), d = cimag(F[dist]);
      double den = c * c + d * d;
      vr[j] = (A * c + Bi * d) / den;
      vi[j] = (A * d - Bi * c) / den; /* reference sign (conjugate) */
    }
    int loc = (L - 1) / 2;
    double rv = select_rank(vr, L, loc);
    double iv = select_rank(vi, L, loc);
    out[hits[h]] = CMPLX(rv, iv);
  }
  free(vr);
  free(vi);

  if (tr) {
    if (tr->comb_approved)
      memcpy(tr->comb_approved, approved, sizeof(int) * (size_t)num_comb);
    tr->num_comb = p->with_comb ? num_comb : 0;
    if (tr->hits) {
      long c = used < tr->hits_cap ? used : tr->hits_cap;
      memcpy(tr->hits, hits, sizeof(int) * (size_t)c);
    }
    tr->hits_found = hits_found;
    tr->hits_prefill = prefill;
  }
  free(xt);
  free(xf);
  free(hits);
  free(approved);
  return 0;
}

int or_v12_exec_many(const or_v12_plan *P, int num, const or_cplx *const *xs,
                     const int *a, const int *ai, const int *comb_off,
                     or_cplx *const *outs, int threads)
{
  /* threads > 0: outs has num entries, one per signal.
   * threads < 0: -threads workers and outs has one entry per worker (a
   * timing mode that keeps the host footprint at one output per thread). */
  const int L = P->p.loc_loops + P->p.est_loops;
  const int cl = P->p.comb_loops > 0 ? P->p.comb_loops : 1;
  const int per_thread = threads < 0;
  const int nt = per_thread ? -threads : threads;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
  for (int s = 0; s < num; s++)
    or_v12_exec(P, xs[s], a + (long)s * L, ai + (long)s * L,
                comb_off + (long)s * cl,
                per_thread ? outs[omp_get_thread_num()] : outs[s], NULL);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* v3                                                                        */
/* ------------------------------------------------------------------------ */

/* src/sfft.cc:506-560 */
int or_v3_params_get(int n_in, int k_in, or_v3_params *p)
{
  memset(p, 0, sizeof(*p));
  unsigned n = (unsigned)or_floor_to_pow2((unsigned)n_in);
  unsigned k = (unsigned)k_in;
  p->n = (int)n;
  p->k = (int)k;
  p->W_Man = or_floor_to_pow2(10 * ((double)k));
  if ((unsigned)p->W_Man > n / 2) p->W_Man = (int)(n / 2);
  p->BB1 = (unsigned)(1 * ((double)k));
  p->B_g1 = or_floor_to_pow2(p->BB1);
  p->lobefrac_g1 = 0.5 / (p->BB1);
  p->b_g1 = (int)(1.00 * ((double)n / p->B_g1));
  p->tol_g1 = 1.e-8;
  p->w_g1 = or_dolph_width(p->lobefrac_g1, p->tol_g1);
  p->BB2 = (unsigned)(0.25 * ((double)k));
  p->B_g2 = or_floor_to_pow2(p->BB2);
  p->lobefrac_g2 = 0.5 / (p->BB2);
  p->b_g2 = (int)(1.00 * ((double)n / p->B_g2));
  p->tol_g2 = 1.e-8;
  p->w_g2 = or_dolph_width(p->lobefrac_g2, p->tol_g2);
  return 0;
}

or_v3_plan *or_v3_plan_make(int n, int k)
{
  or_v3_plan *P = (or_v3_plan *)calloc(1, sizeof(*P));
  or_v3_params_get(n, k, &P->p);
  int N = P->p.n;
  P->time1 = (or_cplx *)malloc(sizeof(or_cplx) * (size_t)P->p.w_g1);
  P->freq1 = (or_cplx *)malloc(sizeof(or_cplx) * (size_t)N);
  or_make_dolphchebyshev_t(P->p.lobefrac_g1, P->p.tol_g1, P->p.w_g1, P->time1);
  or_make_multiple_t(P->time1, P->p.w_g1, N, P->p.b_g1, P->freq1);
  P->time2 = (or_cplx *)malloc(sizeof(or_cplx) * (size_t)P->p.w_g2);
  P->freq2 = (or_cplx *)malloc(sizeof(or_cplx) * (size_t)N);
  or_make_dolphchebyshev_t(P->p.lobefrac_g2, P->p.tol_g2, P->p.w_g2, P->time2);
  or_make_multiple_t(P->time2, P->p.w_g2, N, P->p.b_g2, P->freq2);
  return P;
}

/* The same plan with its filters left for or_v3_plan_set_filters (no
 * filter design: the n-point transforms take seconds at n = 2^24). */
or_v3_plan *or_v3_plan_make_bare(int n, int k)
{
  or_v3_plan *P = (or_v3_plan *)calloc(1, sizeof(*P));
  or_v3_params_get(n, k, &P->p);
  int N = P->p.n;
  P->time1 = (or_cplx *)calloc((size_t)P->p.w_g1, sizeof(or_cplx));
  P->freq1 = (or_cplx *)calloc((size_t)N, sizeof(or_cplx));
  P->time2 = (or_cplx *)calloc((size_t)P->p.w_g2, sizeof(or_cplx));
  P->freq2 = (or_cplx *)calloc((size_t)N, sizeof(or_cplx));
  return P;
}

void or_v3_plan_set_filters(or_v3_plan *P, const or_cplx *time1,
                            const or_cplx *freq1, const or_cplx *time2,
                            const or_cplx *freq2)
{
  size_t N = (size_t)P->p.n;
  memcpy(P->time1, time1, sizeof(or_cplx) * (size_t)P->p.w_g1);
  memcpy(P->freq1, freq1, sizeof(or_cplx) * N);
  memcpy(P->time2, time2, sizeof(or_cplx) * (size_t)P->p.w_g2);
  memcpy(P->freq2, freq2, sizeof(or_cplx) * N);
}

void or_v3_plan_free(or_v3_plan *P)
{
  if (!P) return;
  free(P->time1);
  free(P->freq1);
  free(P->time2);
  free(P->freq2);
  free(P);
}

/* cf3 800-810.  draws6 = {a, ai, b, shift, init_offset, init_G_offset}. */
void or_v3_draw(int n, int *d)
{
  int a = 0;
  while (or_gcd(a, n) != 1) a = (int)(random() % n);
  int ai = or_mod_inverse(a, n);
  int b = (int)(random() % n);
  /* ((ai*b) % n + n) % n with a wrapping int product: (ai*b) mod n */
  int shift = (int)(((long long)ai * b) % n);
  int init_offset = (int)(unsigned)floor(drand48() * n);
  int init_G_offset = (int)(unsigned)floor(drand48() * n);
  d[0] = a;
  d[1] = ai;
  d[2] = b;
  d[3] = shift;
  d[4] = init_offset;
  d[5] = init_G_offset;
}

/* int products in the reference wrap modulo 2^32 (signed int overflow). */
static inline int wmul(int x, int y) { return (int)((unsigned)x * (unsigned)y); }

/* --- a tiny int -> complex map that iterates in ascending key order --- */
typedef struct {
  int *keys;
  or_cplx *vals;
  int *slot; /* open addressing: index into keys/vals, -1 empty */
  long size, cap, nslot;
} omap;

static void omap_init(omap *m, long cap)
{
  m->cap = cap < 16 ? 16 : cap;
  m->keys = (int *)malloc(sizeof(int) * (size_t)m->cap);
  m->vals = (or_cplx *)malloc(sizeof(or_cplx) * (size_t)m->cap);
  m->nslot = 1;
  while (m->nslot < 2 * m->cap) m->nslot <<= 1;
  m->slot = (int *)malloc(sizeof(int) * (size_t)m->nslot);
  for (long i = 0; i < m->nslot; i++) m->slot[i] = -1;
  m->size = 0;
}

static void omap_free(omap *m)
{
  free(m->keys);
  free(m->vals);
  free(m->slot);
}

/* operator[]: returns the index of key, inserting value 0 if absent. */
static long omap_at(omap *m, int key)
{
  unsigned long h = ((unsigned)key * 2654435761u) & (m->nslot - 1);
  while (m->slot[h] >= 0) {
    if (m->keys[m->slot[h]] == key) return m->slot[h];
    h = (h + 1) & (m->nslot - 1);
  }
  if (m->size == m->cap) {
    omap old = *m;
    omap_init(m, old.cap * 2);
    for (long i = 0; i < old.size; i++) {
      long j = omap_at(m, old.keys[i]);
      m->vals[j] = old.vals[i];
    }
    omap_free(&old);
    return omap_at(m, key);
  }
  long idx = m->size++;
  m->keys[idx] = key;
  m->vals[idx] = 0;
  m->slot[h] = (int)idx;
  return idx;
}

static int *omap_sorted_order(const omap *m)
{
  /* indices sorted by key (ascending) */
  int *ord = (int *)malloc(sizeof(int) * (size_t)(m->size + 1));
  long *tmp = (long *)malloc(sizeof(long) * (size_t)(m->size + 1));
  for (long i = 0; i < m->size; i++)
    tmp[i] = ((long)(unsigned)m->keys[i] << 32) | (long)i;
  /* sort the packed (key, index) pairs */
  for (long gap = m->size / 2; gap > 0; gap /= 2) /* shell sort: small sizes */
    for (long i = gap; i < m->size; i++) {
      long t = tmp[i], j = i;
      for (; j >= gap && tmp[j - gap] > t; j -= gap) tmp[j] = tmp[j - gap];
      tmp[j] = t;
    }
  for (long i = 0; i < m->size; i++) ord[i] = (int)(tmp[i] & 0xffffffffL);
  free(tmp);
  return ord;
}

/* cf3 96-127: two aliased sub-samplings, interleaved W-point FFTs. */
static void mansour_filt(const or_cplx *x, int n, int W, int g, or_cplx *S)
{
  int n1 = n - 1, sigma = n / W;
  int idx = g, idx1 = (g + 1) & n1;
  for (int i = 0; i < W; i++) {
    S[2 * i] = x[idx];
    S[2 * i + 1] = x[idx1];
    idx = (idx + sigma) & n1;
    idx1 = (idx1 + sigma) & n1;
  }
}

/* Two interleaved B-point forward FFTs (FFTW plan_many, stride 2 dist 1). */
static void fft_interleaved2(or_cplx *S, int B)
{
  or_cplx *t = (or_cplx *)malloc(sizeof(or_cplx) * (size_t)B);
  for (int l = 0; l < 2; l++) {
    for (int i = 0; i < B; i++) t[i] = S[2 * i + l];
    or_fft(t, t, B, 0);
    for (int i = 0; i < B; i++) S[2 * i + l] = t[i];
  }
  free(t);
}

/* Fold of cf3 130-212 / 215-297: chunks = floor(w/B) full chunks of B taps;
 * loop 0 uses src(t), loop 1 src(t+1). */
static void gauss_fold(or_cplx *S, int B, int w, const or_cplx *tap,
                       const or_cplx *src0, const or_cplx *src1)
{
  memset(S, 0, sizeof(or_cplx) * 2 * (size_t)B);
  int chunks = (2 * w) / (2 * B);
  for (int ch = 0; ch < chunks; ch++) {
    for (int r = 0; r < B; r++) {
      int t = ch * B + r;
      double c = creal(tap[t]), d = cimag(tap[t]);
      double a0 = creal(src0[t]), b0 = cimag(src0[t]);
      double a1 = creal(src1[t]), b1 = cimag(src1[t]);
      double re0 = c * a0 - d * b0, im0 = d * a0 + c * b0;
      double re1 = c * a1 - d * b1, im1 = d * a1 + c * b1;
      S[2 * r] = CMPLX(creal(S[2 * r]) + re0, cimag(S[2 * r]) + im0);
      S[2 * r + 1] = CMPLX(creal(S[2 * r + 1]) + re1, cimag(S[2 * r + 1]) + im1);
    }
  }
}

/* cf3 300-351 */
static void update_mansour(int key, or_cplx value, or_cplx *S, int n, int W,
                           int offset)
{
  const double PI2_OVER_N = 2 * M_PI / n;
  int h = key & (W - 1);
  double vr = creal(value), vi = cimag(value);
  double arg0 = PI2_OVER_N * key * offset, arg1 = PI2_OVER_N * key * (offset + 1);
  double s0 = sin(arg0), c0 = cos(arg0), s1 = sin(arg1), c1 = cos(arg1);
  double rr0 = W * (vr * c0 - vi * s0), ri0 = W * (vr * s0 + vi * c0);
  double rr1 = W * (vr * c1 - vi * s1), ri1 = W * (vr * s1 + vi * c1);
  S[2 * h] = CMPLX(creal(S[2 * h]) - rr0, cimag(S[2 * h]) - ri0);
  S[2 * h + 1] = CMPLX(creal(S[2 * h + 1]) - rr1, cimag(S[2 * h + 1]) - ri1);
}

/* cf3 354-463 */
static void update_gaussian(int key, or_cplx value, or_cplx *S,
                            const or_cplx *F, int n, int B, int g, int step)
{
  const double PI2_DIV_N = 2 * M_PI / n;
  const int n1 = n - 1;
  unsigned nb = (unsigned)n / (unsigned)B;
  int h = (int)(key * 1. / nb + 0.5) % B;
  double a1 = creal(value), b1 = cimag(value);
  int dist1 = (int)(((unsigned)h * nb - (unsigned)key + (unsigned)n) & (unsigned)n1);
  int dist2 = (int)(((unsigned)dist1 + (unsigned)n + nb) & (unsigned)n1);
  int dist3 = (int)(((unsigned)dist1 + (unsigned)n + (unsigned)n - nb) & (unsigned)n1);
  int ko = wmul(key, g) % n;
  int ko2 = wmul(key, (int)((unsigned)g + (unsigned)step)) % n;
  double arg0 = PI2_DIV_N * ko, arg1 = PI2_DIV_N * ko2;
  double b2 = sin(arg0), a2 = cos(arg0), b22 = sin(arg1), a22 = cos(arg1);
  double v1r = a1 * a2 - b1 * b2, v1i = a1 * b2 + b1 * a2;
  double v2r = a1 * a22 - b1 * b22, v2i = a1 * b22 + b1 * a22;
  int bk[3] = {h, (h + 1) % B, (h + B - 1) % B};
  int dd[3] = {dist1, dist2, dist3};
  for (int q = 0; q < 3; q++) {
    double fr = creal(F[dd[q]]), fi = cimag(F[dd[q]]);
    double r1r = v1r * fr - v1i * fi, r1i = v1r * fi + v1i * fr;
    double r2r = v2r * fr - v2i * fi, r2i = v2r * fi + v2i * fr;
    or_cplx *s0 = &S[2 * bk[q]], *s1 = &S[2 * bk[q] + 1];
    *s0 = CMPLX(creal(*s0) - r1r, cimag(*s0) - r1i);
    *s1 = CMPLX(creal(*s1) - r2r, cimag(*s1) - r2i);
  }
}

/* cf3 484-640 */
static int est_gauss(int B, const or_cplx *S, int g, int n, int a, int b,
                     int jump, const or_cplx *F, int *EF, or_cplx *EV)
{
  const unsigned n1 = (unsigned)n - 1, B1 = (unsigned)B - 1;
  const double PI2 = 2 * M_PI;
  const double N_OVER_PI2 = n / PI2;
  const double PI2_A_OFFSET_OVER_N = PI2 * a * g / n;
  const double BUCKETS_OVER_N = (double)B / n;
  const int N_OVER_BUCKETS = n / B;
  const double separation = 1. / jump;
  int found = 0;
  for (int i = 0; i < B; i += 2) {
    for (int j = 0; j < 2; j++) {
      int bk = i + j;
      double a0 = creal(S[2 * bk]), b0 = cimag(S[2 * bk]);
      double a1 = creal(S[2 * bk + 1]), b1 = cimag(S[2 * bk + 1]);
      double zb = (a0 * a0 + b0 * b0) + (a1 * a1 + b1 * b1);
      if (!(zb > 1e-8)) continue;
      double c0 = a0 * a0 + b0 * b0, c1 = a1 * a1 + b1 * b1;
      double d0 = atan2(b0, a0), d1 = atan2(b1, a1);
      double eb = c1 / c0 - 1;
      double err = eb * eb;
      err /= n;
      if (err < 1e-12 && c0 > 0.01) {
        double slope = (d1 - d0) * separation;
        int freq = (int)lrint(N_OVER_PI2 * slope) + n;
        freq = (int)((unsigned)freq & n1);
        unsigned hashed_to = (unsigned)lrint(freq * BUCKETS_OVER_N) & B1;
        if (hashed_to == (unsigned)bk) {
          double ph = d0 - PI2_A_OFFSET_OVER_N * freq;
          double mag = sqrt(c0);
          double sn = sin(ph), cs = cos(ph);
          double mr = mag * cs, mi = mag * sn;
          int dist = (int)((hashed_to * (unsigned)N_OVER_BUCKETS - (unsigned)freq +
                            (unsigned)n) & n1);
          double fr = creal(F[dist]), fi = cimag(F[dist]);
          double den = fr * fr + fi * fi;
          double vr = (mr * fr + mi * fi) / den;
          double vi = (mi * fr - mr * fi) / den;
          freq = (int)(((unsigned)(timesmod(freq, a, n) - b + n)) & n1);
          EF[found] = freq;
          EV[found] = CMPLX(vr, vi);
          found++;
        }
      }
    }
  }
  return found;
}

/* cf3 642-776 */
static int est_mansour(int W, const or_cplx *S, int g, int n, int *EF,
                       or_cplx *EV)
{
  const double PI2 = 2 * M_PI;
  const double N_OVER_PI2 = n / PI2;
  const double PI2_OVER_N = PI2 / n;
  const unsigned FREQ_MASK = (unsigned)(n - 1) & ~(unsigned)(W - 1);
  const double NORM = 1. / W;
  const double NORM2 = NORM * NORM;
  int found = 0;
  for (int i = 0; i < W; i += 2) {
    for (int j = 0; j < 2; j++) {
      int bk = i + j;
      double a0 = creal(S[2 * bk]), b0 = cimag(S[2 * bk]);
      double a1 = creal(S[2 * bk + 1]), b1 = cimag(S[2 * bk + 1]);
      double zb = (a0 * a0 + b0 * b0) * NORM2 + (a1 * a1 + b1 * b1) * NORM2;
      if (!(zb > 1e-8)) continue;
      double c0 = (a0 * a0 + b0 * b0) * NORM2;
      double c1 = (a1 * a1 + b1 * b1) * NORM2;
      double d0 = atan2(b0 * NORM, a0 * NORM), d1 = atan2(b1 * NORM, a1 * NORM);
      double median_abs = c0;
      double inv = 1. / median_abs;
      double bb = (c1 * inv) - 1;
      double err = bb * bb;
      if (err < n * 1e-10 && median_abs > 0.01) {
        double slope = d1 - d0;
        int f1 = (int)lrint(slope * N_OVER_PI2);
        unsigned f2 = (unsigned)f1 & FREQ_MASK;
        int f3 = (int)(f2 | (unsigned)bk);
        int fo = wmul(f3, g);
        double ph = d0 - PI2_OVER_N * fo;
        median_abs = sqrt(median_abs);
        double sn = sin(ph), cs = cos(ph);
        EF[found] = f3;
        EV[found] = CMPLX(median_abs * cs, median_abs * sn);
        found++;
      }
    }
  }
  return found;
}

long or_v3_exec(const or_v3_plan *P, const or_cplx *x, const int *d,
                or_cplx *out, int *early, int *freqs, or_cplx *vals, long cap)
{
  const or_v3_params *p = &P->p;
  const int n = p->n, k = p->k, W = p->W_Man, B1 = p->B_g1, B2 = p->B_g2;
  const int a = d[0], ai = d[1], b = d[2], shift = d[3], g0 = d[4], gG = d[5];
  memset(out, 0, sizeof(or_cplx) * (size_t)n);
  if (early) *early = 0;

  or_cplx *MS = (or_cplx *)malloc(sizeof(or_cplx) * 2 * (size_t)W);
  or_cplx *GS = (or_cplx *)malloc(sizeof(or_cplx) * 2 * (size_t)B1);
  or_cplx *GP = (or_cplx *)malloc(sizeof(or_cplx) * 2 * (size_t)B2);
  long ecap = 2L * W + 2L * B1 + 2L * B2 + 16;
  int *EF = (int *)malloc(sizeof(int) * (size_t)ecap);
  or_cplx *EV = (or_cplx *)malloc(sizeof(or_cplx) * (size_t)ecap);
  omap ans;
  omap_init(&ans, 5L * k + 16);
  long found = 0, nmap = 0;

  mansour_filt(x, n, W, g0, MS);
  fft_interleaved2(MS, W);
  int fn = est_mansour(W, MS, g0, n, EF, EV);
  for (int i = 0; i < fn; i++) {
    long e = omap_at(&ans, EF[i]);
    ans.vals[e] = EV[i];
    found++;
    MS[2 * (EF[i] % W)] = 0;
    MS[2 * (EF[i] % W) + 1] = 0;
  }
  if (found == k) {
    if (early) *early = 1;
    goto done;
  }

  { /* Gauss filter 1, cf3 871-925 */
    or_cplx *s0 = (or_cplx *)malloc(sizeof(or_cplx) * (size_t)p->w_g1 + 16);
    or_cplx *s1 = (or_cplx *)malloc(sizeof(or_cplx) * (size_t)p->w_g1 + 16);
    for (int t = 0; t < p->w_g1; t++) {
      s0[t] = x[(gG + t) & (n - 1)];
      s1[t] = x[(gG + 1 + t) & (n - 1)];
    }
    gauss_fold(GS, B1, p->w_g1, P->time1, s0, s1);
    free(s0);
    free(s1);
    fft_interleaved2(GS, B1);
    /* cf3 881-887 iterates a tr1::unordered_map, whose order no rebuild can
     * reproduce; the map is iterated in insertion order here (and on the
     * GPU), which changes only the rounding of these subtractions. */
    for (long q = 0; q < ans.size; q++)
      update_gaussian(ans.keys[q], ans.vals[q], GS, P->freq1, n, B1, gG, 1);
    fn = est_gauss(B1, GS, gG, n, 1, 0, 1, P->freq1, EF, EV);
    for (int i = 0; i < fn; i++) {
      long e = omap_at(&ans, EF[i]);
      double re = creal(ans.vals[e]), im = cimag(ans.vals[e]);
      if (re * re + im * im == 0) found++;
      ans.vals[e] = ans.vals[e] + EV[i];
      update_gaussian(EF[i], EV[i], GS, P->freq1, n, B1, gG, 1);
      update_mansour(EF[i], EV[i], MS, n, W, g0);
    }
  }

  { /* Gauss filter 2 (permuted), cf3 935-982 and permute_fourier cf3 66-87 */
    int w2 = p->w_g2;
    or_cplx *perm = (or_cplx *)malloc(sizeof(or_cplx) * (size_t)(w2 + 2));
    double th0 = 2 * M_PI * gG * b / n, th1 = 2 * M_PI * ai * b / n;
    double s0, c0, s1, c1;
    sincos(th0, &s0, &c0); /* explicit: see fft_pow2_inplace */
    sincos(th1, &s1, &c1);
    or_cplx sh = CMPLX(c0, s0);
    or_cplx st = CMPLX(c1, s1);
    int idx = gG % n;
    for (int i = 0; i < w2 + 1; i++) {
      double xr = creal(x[idx]), xi = cimag(x[idx]);
      double sr = creal(sh), si = cimag(sh);
      perm[i] = CMPLX(xr * sr - xi * si, xr * si + xi * sr);
      double tr = creal(st), ti = cimag(st);
      sh = CMPLX(sr * tr - si * ti, sr * ti + si * tr);
      idx = (idx + ai) % n;
    }
    gauss_fold(GP, B2, w2, P->time2, perm, perm + 1);
    free(perm);
    fft_interleaved2(GP, B2);
    for (long q = 0; q < ans.size; q++) /* insertion order, see above */
      update_gaussian((timesmod(ans.keys[q], ai, n) + shift) % n, ans.vals[q], GP,
                      P->freq2, n, B2, wmul(a, gG), 1);
    fn = est_gauss(B2, GP, gG, n, a, b, 1, P->freq2, EF, EV);
    for (int i = 0; i < fn; i++) {
      long e = omap_at(&ans, EF[i]);
      double re = creal(ans.vals[e]), im = cimag(ans.vals[e]);
      if (re * re + im * im == 0) found++;
      ans.vals[e] = ans.vals[e] + EV[i];
      update_gaussian((timesmod(EF[i], ai, n) + shift) % n, EV[i], GP, P->freq2,
                      n, B2, wmul(a, gG), 1);
      update_gaussian(EF[i], EV[i], GS, P->freq1, n, B1, gG, 1);
      update_mansour(EF[i], EV[i], MS, n, W, g0);
    }
  }

  { /* cf3 997-1076: round-robin until the three bucket counts repeat */
    int nana = 0, pm = 0, pg1 = 0, pg2 = 0;
    const int max_pass = 3 * 4096;
    while (nana < max_pass) {
      if (nana % 3 == 0) fn = est_mansour(W, MS, g0, n, EF, EV);
      else if (nana % 3 == 1) fn = est_gauss(B1, GS, gG, n, 1, 0, 1, P->freq1, EF, EV);
      else fn = est_gauss(B2, GP, gG, n, a, b, 1, P->freq2, EF, EV);
      for (int i = 0; i < fn; i++) {
        long e = omap_at(&ans, EF[i]);
        double re = creal(ans.vals[e]), im = cimag(ans.vals[e]);
        if (re * re + im * im == 0) found++;
        ans.vals[e] = ans.vals[e] + EV[i];
        update_gaussian((timesmod(EF[i], ai, n) + shift) % n, EV[i], GP,
                        P->freq2, n, B2, wmul(a, gG), 1);
        update_gaussian(EF[i], EV[i], GS, P->freq1, n, B1, gG, 1);
        update_mansour(EF[i], EV[i], MS, n, W, g0);
      }
      if (nana % 3 == 2) {
        int cm = 0, c1 = 0, c2 = 0;
        for (int j = 0; j < B1; j++) {
          double re = creal(GS[j]), im = cimag(GS[j]);
          if (re * re + im * im > 1e-6) c1++;
        }
        for (int j = 0; j < B2; j++) {
          double re = creal(GP[j]), im = cimag(GP[j]);
          if (re * re + im * im > 1e-6) c2++;
        }
        for (int j = 0; j < W; j++) {
          double re = creal(MS[2 * j]) / W, im = cimag(MS[2 * j]) / W;
          if (re * re + im * im > 1e-6) cm++;
        }
        if (pm == cm && pg1 == c1 && pg2 == c2) break;
        pm = cm;
        pg1 = c1;
        pg2 = c2;
      }
      nana++;
    }
  }

  {
    int *ord = omap_sorted_order(&ans);
    for (long q = 0; q < ans.size; q++) {
      out[ans.keys[ord[q]]] = ans.vals[ord[q]];
      if (freqs && q < cap) {
        freqs[q] = ans.keys[ord[q]];
        vals[q] = ans.vals[ord[q]];
      }
    }
    free(ord);
  }
done:
  nmap = ans.size;
  omap_free(&ans);
  free(MS);
  free(GS);
  free(GP);
  free(EF);
  free(EV);
  return nmap;
}
