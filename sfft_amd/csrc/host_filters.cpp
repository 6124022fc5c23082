// Plan-time FP64 work on the host: an FFT for filter design and the
// Dolph-Chebyshev "multiple" filter of the reference (src/filters.cc:62-160).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.hpp"

namespace sfa {

namespace {

// Radix-2 decimation-in-frequency on a power-of-two size; output is
// bit-reversed and then permuted back.  Twiddles come from one table with
// angles reduced exactly (j < n/2).
void fft_pow2(cd *x, int n, bool backward)
{
  if (n < 2) return;
  const int half = n / 2;
  std::vector<cd> tw(half);
  const double sgn = backward ? 1.0 : -1.0;
#pragma omp parallel for if (half > 65536)
  for (int j = 0; j < half; j++) {
    const double ang = 2.0 * M_PI * (double)j / (double)n;
    tw[j] = cd(std::cos(ang), sgn * std::sin(ang));
  }
  for (int len = n; len >= 2; len >>= 1) {
    const int h = len / 2, stride = n / len;
    const int64_t groups = n / len;
#pragma omp parallel for if (n > 65536) schedule(static)
    for (int64_t t = 0; t < (int64_t)groups * h; t++) {
      const int64_t g = t / h;
      const int j = (int)(t % h);
      cd *a = x + g * len + j, *b = a + h;
      const cd u = *a, v = *b;
      *a = u + v;
      const cd d = u - v;
      const cd w = tw[(int64_t)j * stride];
      *b = cd(d.real() * w.real() - d.imag() * w.imag(),
              d.real() * w.imag() + d.imag() * w.real());
    }
  }
  int bits = 0;
  while ((1 << bits) < n) bits++;
  for (int i = 0; i < n; i++) {
    unsigned r = 0, v = (unsigned)i;
    for (int q = 0; q < bits; q++) {
      r = (r << 1) | (v & 1);
      v >>= 1;
    }
    if ((int)r > i) std::swap(x[i], x[r]);
  }
}

// Chirp-z (Bluestein) for other sizes.
void fft_any(cd *x, int n, bool backward)
{
  int m = 1;
  while (m < 2 * n - 1) m <<= 1;
  const double sgn = backward ? 1.0 : -1.0;
  std::vector<cd> chirp(n), a(m), b(m);
  const int64_t two_n = 2LL * n;
  for (int j = 0; j < n; j++) {
    const int64_t q = ((int64_t)j * j) % two_n;
    const double ang = M_PI * (double)q / (double)n;
    chirp[j] = cd(std::cos(ang), sgn * std::sin(ang));
  }
  for (int j = 0; j < n; j++) a[j] = x[j] * chirp[j];
  b[0] = std::conj(chirp[0]);
  for (int j = 1; j < n; j++) b[j] = b[m - j] = std::conj(chirp[j]);
  fft_pow2(a.data(), m, false);
  fft_pow2(b.data(), m, false);
  for (int j = 0; j < m; j++) a[j] *= b[j];
  fft_pow2(a.data(), m, true);
  const double inv = 1.0 / (double)m;
  for (int j = 0; j < n; j++) x[j] = chirp[j] * a[j] * inv;
}

// Chebyshev polynomial T_m(x) for any real x (src/filters.cc:62-68).
double chebyshev(double m, double x)
{
  if (std::fabs(x) <= 1) return std::cos(m * std::acos(x));
  return std::real(std::cosh(m * std::acosh(cd(x, 0.0))));
}

}  // namespace

void host_fft(cd *x, int n, bool backward)
{
  if (n <= 0) return;
  if ((n & (n - 1)) == 0)
    fft_pow2(x, n, backward);
  else
    fft_any(x, n, backward);
}

namespace {

using FftFn = bool (*)(cd *x, int n, bool backward);

bool host_fft_fn(cd *x, int n, bool backward)
{
  host_fft(x, n, backward);
  return true;
}

// n-point DFT on the GPU (Stockham passes), plan time only.
bool gpu_fft_fn(cd *x, int n, bool backward)
{
  double2 *d = nullptr, *t = nullptr;
  const size_t bytes = (size_t)n * sizeof(double2);
  bool ok = hipMalloc(&d, bytes) == hipSuccess && hipMalloc(&t, bytes) == hipSuccess &&
            hipMemcpy(d, x, bytes, hipMemcpyHostToDevice) == hipSuccess &&
            k::batch_fft64(d, t, n, 1, 1, n, n, backward, nullptr, 0, nullptr) == 0 &&
            hipMemcpy(x, d, bytes, hipMemcpyDeviceToHost) == hipSuccess;
  if (d) (void)hipFree(d);
  if (t) (void)hipFree(t);
  if (!ok) set_error("plan-time FFT on the GPU failed");
  return ok;
}

bool design(double lobefrac, double tolerance, int n, int b, FftFn nfft, HostFilter &f);

}  // namespace

HostFilter design_filter(double lobefrac, double tolerance, int n, int b)
{
  HostFilter f;
  design(lobefrac, tolerance, n, b, host_fft_fn, f);
  return f;
}

bool design_filter_gpu(double lobefrac, double tolerance, int n, int b, HostFilter &f)
{
  return design(lobefrac, tolerance, n, b, gpu_fft_fn, f);
}

namespace {

bool design(double lobefrac, double tolerance, int n, int b, FftFn nfft, HostFilter &f)
{
  // Dolph-Chebyshev window: w-point DFT of sampled T_{w-1}, centred, real
  // part (src/filters.cc:70-86).
  const int w = dolph_chebyshev_width(lobefrac, tolerance);
  std::vector<cd> win(w);
  const double t0 = std::cosh(std::acosh(1 / tolerance) / (w - 1));
  for (int i = 0; i < w; i++)
    win[i] = chebyshev(w - 1, t0 * std::cos(M_PI * i / w)) * tolerance;
  host_fft(win.data(), w, false);
  std::vector<cd> centred(w);
  for (int i = 0; i < w; i++) centred[(i + w / 2) % w] = cd(win[i].real(), 0.0);

  // Flatten the pass band over b bins (src/filters.cc:109-160): window placed
  // circularly about 0, n-point DFT, running sum of b bins re-centred,
  // normalised by its largest magnitude, phase ramp of w/2 samples, inverse
  // DFT, first w taps / n.
  std::vector<cd> g(n, cd(0, 0));
  std::copy(centred.begin() + w / 2, centred.end(), g.begin());
  std::copy(centred.begin(), centred.begin() + w / 2, g.end() - w / 2);
  if (!nfft(g.data(), n, false)) return false;
  f.freq.assign(n, cd(0, 0));
  cd run(0, 0);
  for (int i = 0; i < b; i++) run += g[i];
  double peak = 0;
  const int off = b / 2;
  for (int i = 0; i < n; i++) {
    f.freq[(i + n + off) % n] = run;
    peak = std::max(peak, std::abs(run));
    run = run + (g[(i + b) % n] - g[i]);
  }
  for (int i = 0; i < n; i++) f.freq[i] /= peak;
  const double th = -2 * M_PI * (double)(w / 2) / (double)n;
  const cd step(std::cos(th), std::sin(th));
  cd rot(1, 0);
  for (int i = 0; i < n; i++) {  // the ramp is iterated, as the reference does
    f.freq[i] *= rot;
    rot *= step;
  }
  std::vector<cd> t(f.freq);
  if (!nfft(t.data(), n, true)) return false;
  f.time.assign(t.begin(), t.begin() + w);
  for (int i = 0; i < w; i++) f.time[i] /= (double)n;
  return true;
}

}  // namespace
}  // namespace sfa
