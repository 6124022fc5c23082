/*
 * A C caller of include/sfft.h, in the call sequence of the reference's own
 * smoke test (test/test.c:13-44): plan v2 for n = 8192, k = 50 with
 * FFTW_ESTIMATE, seed random() with srand(17) and drand48() with srand48,
 * draw the k tone positions from drand48(), form the input as the inverse
 * DFT of x_f (x_f = 1 at the positions), run sfft_exec into a dense output,
 * free the plan.  Unlike test/test.c it allocates n complex values (the
 * reference passes n bytes to sfft_malloc) and writes the input and the
 * output as raw complex doubles to the two files named on the command line
 * (for the parity check).  The input is summed directly instead of with
 * fftw_dft, which the library does not export.  One exec on a zero input
 * runs before the streams are seeded, so that whatever the GPU runtime
 * initialises on first use happens before the draws that are checked; the
 * next random() and drand48() values after the exec are printed, so that a
 * checker can confirm the exec consumed exactly its draws.
 */
#define _DEFAULT_SOURCE /* srand48 / drand48 / M_PI under -std=c11 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "sfft.h"

#define FFTW_ESTIMATE 64

int main(int argc, char **argv)
{
  const int n = 8192, k = 50;
  const long seed48 = argc > 3 ? atol(argv[3]) : 4242;
  if (argc < 3) {
    fprintf(stderr, "usage: %s input.bin output.bin [srand48 seed]\n", argv[0]);
    return 2;
  }
  sfft_plan *plan = sfft_make_plan(n, k, SFFT_VERSION_2, FFTW_ESTIMATE);
  if (!plan) {
    fprintf(stderr, "sfft_make_plan: %s\n", sfft_last_error());
    return 1;
  }
  complex_t *input = (complex_t *)sfft_malloc(n * sizeof(complex_t));
  complex_t *output = (complex_t *)calloc(n, sizeof(complex_t));
  int *freq = (int *)malloc(k * sizeof(int));
  if (!input || !output || !freq) return 1;
  for (int t = 0; t < n; t++) input[t] = 0;
  sfft_exec(plan, input, output);  /* warm-up, before the streams are seeded */
  srand(17);
  srand48(seed48);
  for (int i = 0; i < k; i++) freq[i] = (int)(unsigned)floor(drand48() * n);
  for (int t = 0; t < n; t++) {
    complex_t s = 0;
    for (int i = 0; i < k; i++) {
      const double ang = 2.0 * M_PI * (double)(((long)freq[i] * t) % n) / (double)n;
      s += cos(ang) + I * sin(ang);
    }
    input[t] = s;
  }
  sfft_exec(plan, input, output);
  const long next_random = random();
  const double next_drand48 = drand48();
  const char *err = sfft_last_error();
  if (err && err[0]) {
    fprintf(stderr, "sfft_exec: %s\n", err);
    return 1;
  }
  int located = 0;
  for (int i = 0; i < n; i++) located += output[i] != 0;
  FILE *fi = fopen(argv[1], "wb"), *fo = fopen(argv[2], "wb");
  if (!fi || !fo) return 1;
  fwrite(input, sizeof(complex_t), n, fi);
  fwrite(output, sizeof(complex_t), n, fo);
  fclose(fi);
  fclose(fo);
  printf("next %ld %.17g\nlocated %d\n", next_random, next_drand48, located);
  sfft_free_plan(plan);
  sfft_free(input);
  free(output);
  free(freq);
  return 0;
}
