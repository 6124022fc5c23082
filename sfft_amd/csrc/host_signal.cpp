// Reference-compatible synthetic input: the harness generator of
// src/simulation.cc:95-112 / src/timing_many.cc:117-139 (k frequencies drawn
// with replacement as floor(drand48() * n), X[f] = 1, x = unnormalised
// inverse DFT) plus the complex Gaussian noise of AWGN (src/utils.cc:250-280),
// from a private drand48 stream seeded like srand48(seed) instead of the
// reference's time ^ pid seed.
#include <cmath>
#include <cstring>

#include "internal.hpp"

using namespace sfa;

extern "C" int sfft_generate_input(int n, int k, long drand48_seed, double std_noise,
                                   complex_t *x, complex_t *x_f, double *snr)
{
  set_error("");
  if (n <= 0 || (n & (n - 1)) != 0 || k < 0 || !x) {
    set_error("sfft_generate_input: n must be a power of two, k >= 0, x non-null");
    return -1;
  }
  Rng rng;
  rng.seed(1, drand48_seed);
  std::vector<cd> X(n, cd(0, 0));
  for (int i = 0; i < k; i++) X[(unsigned)std::floor(rng.next_drand48() * n)] = 1.0;
  if (x_f) std::memcpy(static_cast<void *>(x_f), X.data(), (size_t)n * sizeof(cd));
  host_fft(X.data(), n, true);  // fftw_dft(x, n, x_f, 1): e^{+2 pi i ft/n}
  double sig = 0, noise = 0;
  for (int h = 0; h < n; h++) {
    cd v = X[h];
    if (std_noise != 0) {
      sig += std::norm(v);
      const double u = rng.next_drand48();
      const double w = rng.next_drand48();
      const double r = std_noise * std::sqrt(-2 * std::log(u));
      v += cd(r * std::cos(2 * M_PI * w), r * std::sin(2 * M_PI * w));
      noise += -2 * std::log(u);
    }
    x[h] = v;
  }
  if (snr) *snr = std_noise == 0 ? 1e9 : sig / (noise * std_noise * std_noise);
  return 0;
}
