// Plan-time parameter derivation for sFFT v1/v2/v3.
//
// The tuned constants are the reference's (src/parameters.cc:24-513) and the
// size formulas follow src/sfft.cc:298-372 (v1/v2) and :506-560 (v3),
// including the reference's choice of the K table with n as its key
// (src/sfft.cc:316-321), which makes every k > 50 use the defaults.
#include <cmath>
#include <cstring>

#include "internal.hpp"

namespace sfa {

namespace {

struct TunedRow {
  int key;  // N (n table) or K (k table)
  double Bcst_loc, Bcst_est, Comb_cst;
  int comb_loops, est_loops, loc_loops, thr_loops;
  double tol_loc, tol_est;
};

// src/parameters.cc:33-157 (comb) and :159-309 (plain), keyed by N.
const TunedRow kNComb[] = {
    {1 << 13, 2, 2, 32, 8, 16, 7, 6, 1e-8, 1e-8},
    {1 << 14, 4, 4, 32, 8, 10, 6, 5, 1e-8, 1e-8},
    {1 << 15, 4, 2, 64, 4, 8, 5, 4, 1e-8, 1e-8},
    {1 << 16, 4, 2, 128, 6, 10, 4, 2, 1e-8, 1e-8},
    {1 << 17, 1, 1, 8, 2, 12, 4, 3, 1e-6, 1e-8},
    {1 << 18, 1, 1, 8, 2, 14, 5, 4, 1e-6, 1e-8},
    {1 << 19, 0.5, 0.5, 8, 1, 10, 4, 3, 1e-6, 1e-8},
    {1 << 20, 0.5, 0.5, 8, 2, 12, 4, 2, 1e-6, 1e-8},
    {1 << 21, 0.5, 0.2, 8, 1, 10, 3, 2, 1e-6, 1e-8},
    {1 << 22, 0.5, 0.2, 8, 1, 8, 3, 2, 1e-6, 1e-8},
    {1 << 23, 0.5, 0.2, 8, 1, 8, 3, 2, 1e-6, 1e-8},
    {1 << 24, 0.5, 0.2, 16, 1, 8, 3, 2, 1e-6, 1e-8},
};
const TunedRow kNPlain[] = {
    {1 << 13, 2, 2, 1, 1, 16, 7, 6, 1e-8, 1e-8},
    {1 << 14, 4, 4, 1, 1, 10, 6, 5, 1e-8, 1e-8},
    {1 << 15, 4, 2, 1, 1, 8, 5, 4, 1e-8, 1e-8},
    {1 << 16, 4, 2, 1, 1, 8, 5, 4, 1e-8, 1e-8},
    {1 << 17, 2, 1, 1, 1, 10, 5, 4, 1e-6, 1e-8},
    {1 << 18, 2, 0.5, 1, 1, 14, 4, 3, 1e-6, 1e-8},
    {1 << 19, 1, 0.5, 1, 1, 12, 5, 4, 1e-6, 1e-8},
    {1 << 20, 2, 0.5, 1, 1, 12, 4, 3, 1e-6, 1e-8},
    {1 << 21, 2, 0.2, 1, 1, 15, 3, 2, 1e-6, 1e-8},
    {1 << 22, 4, 0.2, 1, 1, 10, 3, 2, 1e-6, 1e-8},
    {1 << 23, 2, 0.2, 1, 1, 8, 3, 2, 1e-6, 1e-8},
    {1 << 24, 4, 0.2, 1, 1, 8, 3, 2, 1e-6, 1e-8},
};
// src/parameters.cc:323-412 (comb) and :414-510 (plain), keyed by K.
const TunedRow kKComb[] = {
    {50, 0.5, 0.2, 16, 1, 10, 3, 2, 1e-6, 1.0e-8},
    {100, 0.5, 0.2, 16, 1, 12, 4, 2, 1e-6, 1.0e-8},
    {200, 0.5, 0.5, 32, 1, 8, 4, 3, 1e-6, 0.5e-8},
    {500, 0.5, 0.5, 64, 1, 10, 4, 3, 1e-6, 0.5e-8},
    {1000, 1, 1, 128, 3, 12, 4, 3, 1e-6, 0.5e-8},
    {2000, 1, 1, 512, 3, 16, 4, 3, 1e-7, 0.2e-8},
    {2500, 1, 1, 512, 3, 16, 4, 3, 1e-7, 0.2e-8},
    {4000, 1, 2, 512, 3, 14, 8, 7, 1e-8, 0.5e-8},
};
const TunedRow kKPlain[] = {
    {50, 4, 0.2, 1, 1, 10, 3, 2, 1e-6, 1.0e-8},
    {100, 2, 0.2, 1, 1, 12, 3, 2, 1e-6, 1.0e-8},
    {200, 4, 0.5, 1, 1, 10, 3, 2, 1e-6, 0.5e-8},
    {500, 2, 1, 1, 1, 12, 4, 3, 1e-6, 0.5e-8},
    {1000, 2, 1, 1, 1, 12, 5, 4, 1e-6, 1.0e-8},
    {2000, 2, 1, 1, 1, 16, 5, 4, 1e-7, 0.5e-8},
    {2500, 2, 1, 1, 1, 16, 5, 4, 1e-7, 0.5e-8},
    {4000, 2, 2, 1, 1, 14, 6, 5, 1e-8, 1.0e-8},
};

template <size_t N>
void apply_tuned(const TunedRow (&tab)[N], int key, V12Params &p)
{
  for (const TunedRow &r : tab) {
    if (r.key != key) continue;
    p.Bcst_loc = r.Bcst_loc;
    p.Bcst_est = r.Bcst_est;
    p.Comb_cst = r.Comb_cst;
    p.comb_loops = r.comb_loops;
    p.est_loops = r.est_loops;
    p.loc_loops = r.loc_loops;
    p.thr_loops = r.thr_loops;
    p.tol_loc = r.tol_loc;
    p.tol_est = r.tol_est;
    return;
  }
}

bool is_pow2(int64_t x) { return x > 0 && (x & (x - 1)) == 0; }

}  // namespace

int floor_to_pow2(double x)
{
  // Largest power of two <= x; 0 when x < 1 (src/utils.cc:243-248).
  if (!(x >= 1)) return 0;
  unsigned int v = 1;
  while ((double)(v << 1) <= x && v < (1u << 30)) v <<= 1;
  return (int)v;
}

int dolph_chebyshev_width(double lobefrac, double tolerance)
{
  // src/filters.cc:72-74, forced odd.
  int w = (int)((1 / M_PI) * (1 / lobefrac) * std::acosh(1. / tolerance));
  if (w % 2 == 0) w--;
  return w;
}

V12Params derive_v12(int n_in, int k_in, bool comb)
{
  V12Params p;
  p.comb = comb;
  const unsigned k = (unsigned)k_in;
  if (k > 50)
    apply_tuned(comb ? kKComb : kKPlain, n_in, p);  // keyed by n on purpose
  else
    apply_tuned(comb ? kNComb : kNPlain, n_in, p);
  const unsigned n = (unsigned)floor_to_pow2((double)(unsigned)n_in);
  p.n = (int)n;
  p.k = k_in;
  const double base = std::sqrt((double)(int)n * k / std::log2((double)n));
  p.BB_loc = (unsigned)(p.Bcst_loc * base);
  p.BB_est = (unsigned)(p.Bcst_est * base);
  p.b_loc = (int)(1.2 * 1.1 * ((double)n / p.BB_loc));
  p.b_est = (int)(1.4 * 1.1 * ((double)n / p.BB_est));
  p.B_loc = floor_to_pow2(p.BB_loc);
  p.B_est = floor_to_pow2(p.BB_est);
  p.B_thresh = 2 * (int)k;
  p.W_comb = p.B_loc > 0 ? floor_to_pow2(p.Comb_cst * n / p.B_loc) : 0;
  p.w_loc = p.BB_loc > 0 ? dolph_chebyshev_width(0.5 / p.BB_loc, p.tol_loc) : 0;
  p.w_est = p.BB_est > 0 ? dolph_chebyshev_width(0.5 / p.BB_est, p.tol_est) : 0;
  return p;
}

V3Params derive_v3(int n_in, int k_in)
{
  V3Params p;
  const unsigned n = (unsigned)floor_to_pow2((double)(unsigned)n_in);
  p.n = (int)n;
  p.k = k_in;
  unsigned wm = (unsigned)floor_to_pow2(10 * (double)k_in);
  p.W_Man = (int)(wm < n / 2 ? wm : n / 2);
  p.BB1 = (unsigned)(1 * (double)k_in);
  p.B_g1 = floor_to_pow2(p.BB1);
  p.b_g1 = p.B_g1 > 0 ? (int)(1.00 * ((double)n / p.B_g1)) : 0;
  p.w_g1 = p.BB1 > 0 ? dolph_chebyshev_width(0.5 / p.BB1, 1e-8) : 0;
  p.BB2 = (unsigned)(0.25 * (double)k_in);
  p.B_g2 = floor_to_pow2(p.BB2);
  p.b_g2 = p.B_g2 > 0 ? (int)(1.00 * ((double)n / p.B_g2)) : 0;
  p.w_g2 = p.BB2 > 0 ? dolph_chebyshev_width(0.5 / p.BB2, 1e-8) : 0;
  return p;
}

std::string check_v12(const V12Params &p)
{
  const int num = p.B_thresh;
  if (!is_pow2(p.n) || p.n < 16) return "n must be a power of two >= 16";
  if (p.n > (1 << 29)) return "n above 2^29 is not supported";
  if (p.k < 1) return "k must be >= 1";
  if (p.B_loc < 2 || p.B_est < 2) return "bucket count below 2 for this (n,k)";
  // find_largest_indices needs n >= num + 1 (src/utils.cc:134).
  if (p.B_loc < num + 1 || p.B_est < num + 1)
    return "B_loc/B_est < 2k+1: the reference's top-2k selection cannot run";
  if (p.comb && (p.W_comb < num + 1 || p.n % p.W_comb != 0))
    return "W_Comb < 2k+1: the reference's comb selection cannot run";
  if (p.loops() > kMaxLoops) return "more than 32 loops is not supported";
  // v1 vote bitmap: loc_loops * B_loc bits in 64 KiB of LDS
  if ((int64_t)p.loc_loops * ((p.B_loc + 31) / 32) * 4 > 65536)
    return "loc_loops * B_loc above 2^19 is not supported";
  if (p.comb && p.W_comb > (1 << 22)) return "comb width above 2^22 is not supported";
  if (p.B_loc > (1 << 22) || p.B_est > (1 << 22))
    return "bucket count above 2^22 is not supported";
  if (p.w_loc < 1 || p.w_est < 1) return "empty filter";
  if (p.w_loc > p.n || p.w_est > p.n) return "filter longer than n";
  return "";
}

std::string check_v3(const V3Params &p)
{
  if (!is_pow2(p.n) || p.n < 64) return "n must be a power of two >= 64";
  if (p.n > (1 << 29)) return "n above 2^29 is not supported";
  // estimate_freq_* walk buckets in pairs (cf3 537, 695): need >= 2.
  if (p.B_g2 < 2 || p.B_g1 < 2 || p.W_Man < 2) return "k must be >= 8 for v3";
  if (p.w_g1 + 2 >= p.n || p.w_g2 + 2 >= p.n) return "filter longer than n";
  return "";
}

}  // namespace sfa
