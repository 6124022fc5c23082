/*
 * Model of the library's FP32 bucket form -- TEST INFRASTRUCTURE ONLY.
 *
 * Not a restatement of the reference (whose buckets are FP64, see
 * sfft_oracle.c): this is the summation order and rounding the FP32+refine
 * form defines for its buckets, written plainly so that tests can check the
 * GPU's FP32 buckets bit for bit.  Compiled with -ffp-contract=off; the FP32
 * complex product uses fmaf exactly where the device code does.
 *
 * Fold (the bucketization of src/computefourier-1.0-2.0.cc:222-261 in FP32):
 * bucket b of a loop with gather multiplier ai sums x[p] * tap[i] over the
 * taps i = b + kB < w, p = i*ai mod n.  The signal is viewed as n/B rows of B
 * samples (row of p = p / B); rows are cut into segments of seg_rows (all
 * rows when there are fewer); within a segment the taps whose sample lies in
 * it are added in ascending i with compensated (2Sum) FP32 sums of the FP32
 * products (float)x * tap32, and the segments' sums (sum + compensation)
 * are added in segment order.
 *
 * FFT (the bucket FFTs of cf12 270-275 in FP32): forward radix-2 DIT on
 * bit-reversed input with twiddles (float)cos, (float)-sin of 2 pi j / nmax
 * (nmax the plan's table size), one stage after another.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

static inline void cmulf(float ar, float ai, float br, float bi, float *cr, float *ci)
{
  *cr = fmaf(ar, br, -(ai * bi));
  *ci = fmaf(ar, bi, ai * br);
}

/* x: n complex doubles (interleaved); taps32: w complex floats; out: B
 * complex floats (pre-FFT buckets). */
void fm_fold32(const double *x, int n, int B, int w, const float *taps32, unsigned ai,
               int seg_rows, float *out)
{
  const int rows = n / B;
  const int seg = rows < seg_rows ? rows : seg_rows;
  const unsigned mask = (unsigned)n - 1;
  for (int b = 0; b < B; b++) {
    float tr = 0.f, ti = 0.f;
    for (int t0 = 0; t0 < rows; t0 += seg) {
      float sr = 0.f, si = 0.f, cr = 0.f, ci = 0.f;
      for (int i = b; i < w; i += B) {
        const unsigned p = ((unsigned)i * ai) & mask;
        const int t = (int)(p / (unsigned)B);
        if (t < t0 || t >= t0 + seg) continue;
        float pr, pi;
        cmulf((float)x[2 * (size_t)p], (float)x[2 * (size_t)p + 1], taps32[2 * i],
              taps32[2 * i + 1], &pr, &pi);
        float s = sr + pr, bp = s - sr;
        cr += (sr - (s - bp)) + (pr - bp);
        sr = s;
        s = si + pi;
        bp = s - si;
        ci += (si - (s - bp)) + (pi - bp);
        si = s;
      }
      const float qr = sr + cr, qi = si + ci;
      if (t0 == 0) {
        tr = qr;
        ti = qi;
      } else {
        tr = tr + qr;
        ti = ti + qi;
      }
    }
    out[2 * b] = tr;
    out[2 * b + 1] = ti;
  }
}

/* In-place forward FFT of N complex floats (natural order in and out). */
void fm_fft32(float *x, int N, int nmax)
{
  for (int i = 1, j = 0; i < N; i++) {
    int bit = N >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      float t = x[2 * i];
      x[2 * i] = x[2 * j];
      x[2 * j] = t;
      t = x[2 * i + 1];
      x[2 * i + 1] = x[2 * j + 1];
      x[2 * j + 1] = t;
    }
  }
  for (int len = 2; len <= N; len <<= 1) {
    const int h = len >> 1;
    const int tstride = nmax / len;
    for (int s = 0; s < N; s += len) {
      for (int j = 0; j < h; j++) {
        double sn, cs;
        sincos(2.0 * M_PI * (double)(j * tstride) / (double)nmax, &sn, &cs);
        const float wr = (float)cs, wi = (float)-sn;
        float vr, vi;
        cmulf(x[2 * (s + j + h)], x[2 * (s + j + h) + 1], wr, wi, &vr, &vi);
        const float ur = x[2 * (s + j)], ui = x[2 * (s + j) + 1];
        x[2 * (s + j)] = ur + vr;
        x[2 * (s + j) + 1] = ui + vi;
        x[2 * (s + j + h)] = ur - vr;
        x[2 * (s + j + h) + 1] = ui - vi;
      }
    }
  }
}
