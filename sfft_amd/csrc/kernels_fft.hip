// Batched complex FFTs of power-of-two size on gfx950.
//
//   * N <= kMaxLdsFft64 (FP64) / 2*kMaxLdsFft64 (FP32): one workgroup per
//     transform, in LDS (radix-2 DIT with the plan's twiddle table);
//   * larger N: Stockham autosort passes in global memory, radix 4 (and one
//     radix-2 pass for odd log2 N), twiddles from sincos in FP64.
//
// Transforms are addressed as base + outer*stride_outer + inner*stride_inner
// with contiguous elements, which covers the bucket arrays of a batch (one
// array per (signal, loop)) and plain contiguous batches.
#include <algorithm>

#include "device_common.hpp"
#include "internal.hpp"

namespace sfa {
namespace k {

using namespace sfa::dev;

namespace {

template <typename T2> struct Real;
template <> struct Real<double2> { using T = double; };
template <> struct Real<float2> { using T = float; };

template <typename T2>
__device__ __forceinline__ T2 mk(double x, double y)
{
  using T = typename Real<T2>::T;
  T2 r;
  r.x = (T)x;
  r.y = (T)y;
  return r;
}

__device__ __forceinline__ int64_t fft_base(int64_t t, int nb_inner, int64_t so, int64_t si)
{
  return (t / nb_inner) * so + (t % nb_inner) * si;
}

// One Stockham pass of radix R with current span Ns (1, R, R^2, ...):
// v[r] = in[j + r N/R] (times the twiddle e^{sign 2 pi i r (j mod Ns)/(Ns R)}),
// R-point DFT, out[(j/Ns) Ns R + j mod Ns + r Ns] = v[r].
template <typename T2, int R>
__global__ void __launch_bounds__(256) k_stockham(const T2 *in, T2 *out, int N, int Ns,
                                                  int64_t total, int nb_inner, int64_t so,
                                                  int64_t si, double sign)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int M = N / R;
  const int64_t b = t / M;
  const int j = (int)(t % M);
  const int64_t base = fft_base(b, nb_inner, so, si);
  const T2 *src = in + base;
  T2 *dst = out + base;
  T2 v[R];
#pragma unroll
  for (int r = 0; r < R; r++) v[r] = src[j + r * M];
  const int k = j % Ns;
  if (Ns > 1) {
#pragma unroll
    for (int r = 1; r < R; r++) {
      double s, c;
      sincos(sign * 2.0 * M_PI * (double)(r * k) / (double)(Ns * R), &s, &c);
      v[r] = cmul(v[r], mk<T2>(c, s));
    }
  }
  if (R == 2) {
    const T2 a = v[0], c = v[1];
    v[0] = cadd(a, c);
    v[1] = csub(a, c);
  } else {
    const T2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const T2 t2 = cadd(v[1], v[3]), t3 = csub(v[1], v[3]);
    // (sign i) * t3
    const T2 it3 = mk<T2>(-sign * (double)t3.y, sign * (double)t3.x);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, it3);
    v[3] = csub(t1, it3);
  }
  const int idx = (j / Ns) * Ns * R + k;
#pragma unroll
  for (int r = 0; r < R; r++) dst[idx + r * Ns] = v[r];
}

// One workgroup per transform, in LDS; forward only (the plan's table).
template <typename T2>
__global__ void __launch_bounds__(kBlock) k_lds_fft_batch(T2 *data, int logN, int nb_inner,
                                                          int64_t so, int64_t si, const T2 *tw,
                                                          int tws)
{
  extern __shared__ char smem_raw[];
  T2 *buf = reinterpret_cast<T2 *>(smem_raw);
  const int N = 1 << logN;
  T2 *io = data + fft_base(blockIdx.x, nb_inner, so, si);
  for (int m = threadIdx.x; m < N; m += blockDim.x) buf[bitrev(m, logN)] = io[m];
  __syncthreads();
  lds_fft(buf, logN, tw, tws);
  for (int m = threadIdx.x; m < N; m += blockDim.x) io[m] = buf[m];
}

template <typename T2>
__global__ void k_copy(const T2 *in, T2 *out, int N, int64_t total, int nb_inner, int64_t so,
                       int64_t si)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int64_t base = fft_base(t / N, nb_inner, so, si) + t % N;
  out[base] = in[base];
}

// Radix-2 DIT transforms larger than one workgroup's LDS, with the same
// butterflies and twiddles as lds_fft (so FP64 results are bit-identical to
// the in-LDS form): the first log2 C stages run in LDS on the C-point runs
// of the bit-reversed sequence (data -> tmp), the remaining stages as
// global butterfly passes (tmp in place; the last pass writes data).
//
// Run c of the bit-reversed sequence holds x[q*R + bitrev(c)] at LDS
// position bitrev(q), q < C, R = N/C runs.
template <typename T2>
__global__ void __launch_bounds__(kBlock) k_r2_runs(const T2 *data, T2 *tmp, int logN, int logC,
                                                  int nb_inner, int64_t so, int64_t si,
                                                  const T2 *tw, int tw_nmax)
{
  extern __shared__ char smem_raw[];
  T2 *buf = reinterpret_cast<T2 *>(smem_raw);
  const int C = 1 << logC;
  const int logR = logN - logC;
  const int64_t t = blockIdx.x >> logR;
  const int c = blockIdx.x & ((1 << logR) - 1);
  const int64_t base = fft_base(t, nb_inner, so, si);
  const T2 *in = data + base + bitrev(c, logR);
  for (int q = threadIdx.x; q < C; q += blockDim.x) buf[bitrev(q, logC)] = in[(int64_t)q << logR];
  __syncthreads();
  lds_fft(buf, logC, tw, tw_nmax >> logC);
  T2 *out = tmp + base + ((int64_t)c << logC);
  for (int q = threadIdx.x; q < C; q += blockDim.x) out[q] = buf[q];
}

// Stage s (span 2^s) of every transform: one butterfly per thread.
template <typename T2>
__global__ void __launch_bounds__(256) k_r2_pass(const T2 *src, T2 *dst, int logN, int s,
                                                 int64_t total, int nb_inner, int64_t so,
                                                 int64_t si, const T2 *tw, int tw_nmax)
{
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int64_t t = g >> (logN - 1);
  const int b = (int)(g & ((1 << (logN - 1)) - 1));
  const int64_t base = fft_base(t, nb_inner, so, si);
  const int h = 1 << (s - 1);
  const int j = b & (h - 1);
  const int i0 = ((b >> (s - 1)) << s) + j;
  const int i1 = i0 + h;
  const T2 w = tw[j * (tw_nmax >> s)];
  const T2 u = src[base + i0];
  const T2 v = cmul(src[base + i1], w);
  dst[base + i0] = cadd(u, v);
  dst[base + i1] = csub(u, v);
}

template <typename T2>
int radix2(T2 *data, T2 *tmp, int N, int C, int64_t nb_outer, int nb_inner, int64_t so,
           int64_t si, const T2 *tw, int tw_nmax, hipStream_t st)
{
  if (N <= 1) return 0;
  if (!tw || tw_nmax < N) return -1;
  int logN = 0, logC = 0;
  while ((1 << logN) < N) logN++;
  while ((1 << logC) < C) logC++;
  const int64_t nb = nb_outer * nb_inner;
  const int L = std::min(logN, logC);
  const size_t lds = ((size_t)1 << L) * sizeof(T2);
  if (lds > 65536) {
    raise_lds_limit(reinterpret_cast<const void *>(k_lds_fft_batch<T2>));
    raise_lds_limit(reinterpret_cast<const void *>(k_r2_runs<T2>));
  }
  if (logN <= logC) {
    if (nb >= (1LL << 31)) return -1;
    hipLaunchKernelGGL(k_lds_fft_batch<T2>, dim3((unsigned)nb), dim3(kBlock), lds, st, data,
                       logN, nb_inner, so, si, tw, tw_nmax / N);
    return hipGetLastError() == hipSuccess ? 0 : -1;
  }
  if (!tmp || (nb << (logN - logC)) >= (1LL << 31)) return -1;
  hipLaunchKernelGGL(k_r2_runs<T2>, dim3((unsigned)(nb << (logN - logC))), dim3(kBlock), lds,
                     st, data, tmp, logN, logC, nb_inner, so, si, tw, tw_nmax);
  const int64_t total = nb << (logN - 1);
  const dim3 grid((unsigned)((total + 255) / 256));
  for (int s = logC + 1; s <= logN; s++)
    hipLaunchKernelGGL(k_r2_pass<T2>, grid, dim3(256), 0, st, tmp, s == logN ? data : tmp, logN,
                       s, total, nb_inner, so, si, tw, tw_nmax);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename T2>
int stockham(T2 *data, T2 *tmp, int N, int64_t nb_outer, int nb_inner, int64_t so, int64_t si,
             bool backward, hipStream_t st)
{
  const int64_t batches = nb_outer * nb_inner;
  int passes = 0;
  for (int Ns = 1; Ns < N;) {
    Ns *= (Ns * 4 <= N) ? 4 : 2;
    passes++;
  }
  const double sign = backward ? 1.0 : -1.0;
  T2 *src = data, *dst = tmp;
  if (passes % 2 == 1) {  // make the last pass land in data
    const int64_t total = batches * N;
    hipLaunchKernelGGL(k_copy<T2>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                       data, tmp, N, total, nb_inner, so, si);
    src = tmp;
    dst = data;
  }
  for (int Ns = 1; Ns < N;) {
    const int R = (Ns * 4 <= N) ? 4 : 2;
    const int64_t total = batches * (N / R);
    const dim3 grid((unsigned)((total + 255) / 256));
    if (R == 4)
      hipLaunchKernelGGL((k_stockham<T2, 4>), grid, dim3(256), 0, st, src, dst, N, Ns, total,
                         nb_inner, so, si, sign);
    else
      hipLaunchKernelGGL((k_stockham<T2, 2>), grid, dim3(256), 0, st, src, dst, N, Ns, total,
                         nb_inner, so, si, sign);
    Ns *= R;
    std::swap(src, dst);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Uniform noise in [-1, 1) per component from a splitmix64 hash of the
// index (synthetic signals for the autotuner).
__global__ void k_fill_noise(double2 *x, int64_t total, uint64_t seed)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  uint64_t z = seed + 0x9e3779b97f4a7c15ull * (uint64_t)(i + 1);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  const double re = (double)(z >> 11) * 0x1.0p-53 * 2.0 - 1.0;
  const double im = (double)(z & 0x1fffffull) * 0x1.0p-21 * 2.0 - 1.0;
  x[i] = make_double2(re, im);
}

}  // namespace

int fill_noise(double2 *x, int64_t total, uint64_t seed, hipStream_t st)
{
  hipLaunchKernelGGL(k_fill_noise, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, x,
                     total, seed);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int radix2_fft64(double2 *data, double2 *tmp, int N, int64_t nb_outer, int nb_inner, int64_t so,
                 int64_t si, const double2 *tw, int tw_nmax, hipStream_t st)
{
  return radix2<double2>(data, tmp, N, kMaxLdsFft64, nb_outer, nb_inner, so, si, tw, tw_nmax,
                         st);
}

int radix2_fft32(float2 *data, float2 *tmp, int N, int64_t nb_outer, int nb_inner, int64_t so,
                 int64_t si, const float2 *tw, int tw_nmax, hipStream_t st)
{
  return radix2<float2>(data, tmp, N, 2 * kMaxLdsFft64, nb_outer, nb_inner, so, si, tw,
                        tw_nmax, st);
}

int batch_fft64(double2 *data, double2 *tmp, int N, int64_t nb_outer, int nb_inner,
                int64_t so, int64_t si, bool backward, const double2 *tw, int tw_nmax,
                hipStream_t st)
{
  if (N <= 1) return 0;
  if (!backward && N <= kMaxLdsFft64 && tw && tw_nmax >= N && nb_outer * nb_inner < (1LL << 31)) {
    int logN = 0;
    while ((1 << logN) < N) logN++;
    const size_t lds = (size_t)N * sizeof(double2);
    if (lds > 65536) raise_lds_limit(reinterpret_cast<const void *>(k_lds_fft_batch<double2>));
    hipLaunchKernelGGL(k_lds_fft_batch<double2>, dim3((unsigned)(nb_outer * nb_inner)),
                       dim3(kBlock), lds, st, data, logN, nb_inner, so, si, tw, tw_nmax / N);
    return hipGetLastError() == hipSuccess ? 0 : -1;
  }
  if (!tmp) return -1;
  return stockham<double2>(data, tmp, N, nb_outer, nb_inner, so, si, backward, st);
}

int batch_fft32(float2 *data, float2 *tmp, int N, int64_t nb_outer, int nb_inner, int64_t so,
                int64_t si, bool backward, const float2 *tw, int tw_nmax, hipStream_t st)
{
  if (N <= 1) return 0;
  if (!backward && N <= 2 * kMaxLdsFft64 && tw && tw_nmax >= N &&
      nb_outer * nb_inner < (1LL << 31)) {
    int logN = 0;
    while ((1 << logN) < N) logN++;
    const size_t lds = (size_t)N * sizeof(float2);
    if (lds > 65536) raise_lds_limit(reinterpret_cast<const void *>(k_lds_fft_batch<float2>));
    hipLaunchKernelGGL(k_lds_fft_batch<float2>, dim3((unsigned)(nb_outer * nb_inner)),
                       dim3(kBlock), lds, st, data, logN, nb_inner, so, si, tw, tw_nmax / N);
    return hipGetLastError() == hipSuccess ? 0 : -1;
  }
  if (!tmp) return -1;
  return stockham<float2>(data, tmp, N, nb_outer, nb_inner, so, si, backward, st);
}

}  // namespace k
}  // namespace sfa
