// Test infrastructure only: extern "C" entry points over the reference's own
// FFTW-free helpers (src/utils.cc, src/parameters.cc), compiled from the
// reference tree by oracle/Makefile into oracle/_ref/.  Used to pin the
// oracle's restatement of these helpers; nothing in the product loads it.
#include "utils.h"
#include "parameters.h"

extern "C" {
int ref_gcd(int a, int b) { return gcd(a, b); }
int ref_mod_inverse(int a, int n) { return mod_inverse(a, n); }
int ref_floor_to_pow2(double x) { return floor_to_pow2(x); }
void ref_find_largest_indices(int *out, int num, double *samples, int n,
                              double *tmp)
{
  find_largest_indices(out, num, samples, n, tmp);
}
void ref_radix_sort(int *a, int size) { radix_sort(a, size); }
void ref_get_params(int table_k, int key, int with_comb, double *d, int *i)
{
  // d = {Bcst_loc, Bcst_est, Comb_cst, tol_loc, tol_est}
  // i = {loc_loops, est_loops, threshold_loops, comb_loops}
  if (table_k)
    get_expermient_vs_K_parameters(key, with_comb != 0, d[0], d[1], d[2], i[0],
                                   i[1], i[2], i[3], d[3], d[4]);
  else
    get_expermient_vs_N_parameters(key, with_comb != 0, d[0], d[1], d[2], i[0],
                                   i[1], i[2], i[3], d[3], d[4]);
}
}
