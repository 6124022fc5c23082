// Device helpers shared by the sFFT kernels: complex arithmetic, an
// in-LDS radix-2 FFT, block scans and the top-num selection with the
// reference's tie rule.  Compiled with -ffp-contract=off, so every FP64
// product and sum below is rounded on its own, as in the reference's SSE2
// code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfa {
namespace dev {

constexpr int kBlock = 256;

__device__ __forceinline__ double2 cadd(double2 a, double2 b)
{
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b)
{
  return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b)
{
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b)
{
  return make_float2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

__device__ __forceinline__ unsigned bitrev(unsigned v, int bits)
{
  return bits == 0 ? 0u : (__brev(v) >> (32 - bits));
}

// Stages s .. s+R-1 of an in-place radix-2 DIT FFT over N points in LDS, in
// registers: with h = 2^(s-1), the 2^R points i0 + m*h (m < 2^R) of a group
// only meet each other in those stages, so one thread loads them, applies the
// stages' butterflies in order (the same cmul / add / sub on the same
// twiddles as one stage per pass) and stores them.  Ends with a barrier.
template <typename T2, int R>
__device__ __forceinline__ void lds_fft_pass(T2 *buf, int N, int s, const T2 *__restrict__ tw,
                                             int tws)
{
  constexpr int M = 1 << R;
  const int h = 1 << (s - 1);
  const int groups = N >> R;
  for (int g = threadIdx.x; g < groups; g += blockDim.x) {
    const int j = g & (h - 1);
    const int i0 = ((g >> (s - 1)) << (s - 1 + R)) + j;  // group block * (h << R) + j
    T2 v[M];
#pragma unroll
    for (int m = 0; m < M; m++) v[m] = buf[i0 + m * h];
#pragma unroll
    for (int r = 0; r < R; r++) {
      // stage s + r: half-size H = h << r; the lower point i0 + m*h sits at
      // j + (m mod 2^r) * h within its butterfly block
      const int tstride = (N >> (s + r)) * tws;
#pragma unroll
      for (int m = 0; m < M; m++) {
        if (m & (1 << r)) continue;
        const int m1 = m + (1 << r);
        const T2 w = tw[(j + (m & ((1 << r) - 1)) * h) * tstride];
        const T2 u = v[m];
        const T2 t = cmul(v[m1], w);
        v[m] = cadd(u, t);
        v[m1] = csub(u, t);
      }
    }
#pragma unroll
    for (int m = 0; m < M; m++) buf[i0 + m * h] = v[m];
  }
  __syncthreads();
}

// In-place radix-2 decimation-in-time FFT of N = 2^logN points held in LDS in
// bit-reversed order; output in natural order.  tw[j * tws] = e^{-2 pi i j/N}
// where tws = Nmax / N for a table built for Nmax.  Three stages per LDS pass
// (lds_fft_pass); every butterfly is the radix-2 one, so results are those of
// one stage per pass, bit for bit.  Ends with a barrier.
template <typename T2>
__device__ void lds_fft(T2 *buf, int logN, const T2 *__restrict__ tw, int tws)
{
  const int N = 1 << logN;
  int s = 1;
  for (; logN - s + 1 >= 3; s += 3) lds_fft_pass<T2, 3>(buf, N, s, tw, tws);
  if (logN - s + 1 == 2) lds_fft_pass<T2, 2>(buf, N, s, tw, tws);
  else if (logN - s + 1 == 1) lds_fft_pass<T2, 1>(buf, N, s, tw, tws);
}

// Exclusive prefix sum over the block (blockDim.x a multiple of 64);
// tmp holds >= 16 ints of LDS.  Contains barriers.
__device__ __forceinline__ int block_exclusive_scan(int v, int *tmp, int &total)
{
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) tmp[wid] = x;
  __syncthreads();
  int off = 0, tot = 0;
  for (int q = 0; q < nw; q++) {
    const int t = tmp[q];
    off += (q < wid) ? t : 0;
    tot += t;
  }
  __syncthreads();
  total = tot;
  return off + x - v;
}

// Indices of the num largest of mag[0..N) (in LDS), written ascending to out
// -- the rule of src/utils.cc:130-158: cutoff = (num+1)-th largest value;
// every index with mag > cutoff, then indices with mag == cutoff in index
// order until num are taken; output sorted.  mag values are >= +0, so their
// IEEE bit patterns order like the values.  Requires N >= num + 1.
// The cutoff is found by byte-wise radix selection from the top byte; once
// at most 128 keys share the selected prefix, their exact rank is computed
// directly.  Uses 256+16 ints of LDS at sh.  Contains barriers.
__device__ void block_top_num(const double *mag, int N, int num, int *out,
                              int *sh)
{
  int *hist = sh;          // 256
  int *misc = sh + 256;    // 16
  const int T = blockDim.x, tid = threadIdx.x;
  // rank of the cutoff in ascending order among the keys matching prefix
  int rank = N - num - 1;
  uint64_t prefix = 0, hi_mask = 0;
  uint64_t cut = 0;
  bool direct = false;
  for (int pass = 7; pass >= 0; --pass) {
    const int shift = pass * 8;
    for (int d = tid; d < 256; d += T) hist[d] = 0;
    __syncthreads();
    for (int i = tid; i < N; i += T) {
      const uint64_t key = (uint64_t)__double_as_longlong(mag[i]);
      if ((key & hi_mask) == prefix) atomicAdd(&hist[(key >> shift) & 255], 1);
    }
    __syncthreads();
    if (tid < 64) {
      // lane l holds bins 4l..4l+3; the bin holding rank is in the first
      // lane whose inclusive count exceeds rank
      int h[4], sum = 0;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        h[q] = hist[4 * tid + q];
        sum += h[q];
      }
      int incl = sum;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (tid >= o) incl += y;
      }
      const unsigned long long ball = __ballot(rank < incl);
      if (tid == __ffsll(ball) - 1) {
        int cum = incl - sum, q = 0;
        while (q < 3 && rank >= cum + h[q]) cum += h[q++];
        misc[0] = 4 * tid + q;
        misc[1] = rank - cum;
        misc[2] = h[q];
      }
    }
    __syncthreads();
    prefix |= (uint64_t)misc[0] << shift;
    hi_mask |= 255ull << shift;
    rank = misc[1];
    const int cand = misc[2];
    __syncthreads();  // misc and hist are rewritten below
    if (pass > 0 && cand <= 128) {
      // the rank-th smallest of the cand keys matching prefix: the key with
      // (#keys < it) <= rank < (#keys <= it)
      // the 128 keys as low words in hist[0..128), high words in hist[128..256)
      unsigned *klo = reinterpret_cast<unsigned *>(hist), *khi = klo + 128;
      if (tid == 0) misc[3] = 0;
      __syncthreads();
      for (int i = tid; i < N; i += T) {
        const uint64_t key = (uint64_t)__double_as_longlong(mag[i]);
        if ((key & hi_mask) == prefix) {
          const int slot = atomicAdd(&misc[3], 1);
          klo[slot] = (unsigned)key;
          khi[slot] = (unsigned)(key >> 32);
        }
      }
      __syncthreads();
      if (tid < cand) {
        const uint64_t kj = (uint64_t)klo[tid] | (uint64_t)khi[tid] << 32;
        int lt = 0, le = 0;
        for (int q = 0; q < cand; q++) {
          const uint64_t kq = (uint64_t)klo[q] | (uint64_t)khi[q] << 32;
          lt += kq < kj;
          le += kq <= kj;
        }
        if (lt <= rank && rank < le) {  // every such thread holds the same key
          misc[4] = (int)(unsigned)(kj & 0xffffffffu);
          misc[5] = (int)(unsigned)(kj >> 32);
        }
      }
      __syncthreads();
      cut = (uint64_t)(unsigned)misc[4] | (uint64_t)(unsigned)misc[5] << 32;
      __syncthreads();
      direct = true;
      break;
    }
  }
  if (!direct) cut = prefix;
  // contiguous chunk per thread so that the compaction keeps index order
  const int C = (N + T - 1) / T;
  const int lo = tid * C, hi = min(N, lo + C);
  int ngt = 0, neq = 0;
  for (int i = lo; i < hi; i++) {
    const uint64_t key = (uint64_t)__double_as_longlong(mag[i]);
    ngt += key > cut;
    neq += key == cut;
  }
  int tot_gt, tot_eq;
  (void)block_exclusive_scan(ngt, misc, tot_gt);
  const int eq_before = block_exclusive_scan(neq, misc, tot_eq);
  const int need = num - tot_gt;  // ties to take, lowest indices first
  int nsel = 0, er = eq_before;
  for (int i = lo; i < hi; i++) {
    const uint64_t key = (uint64_t)__double_as_longlong(mag[i]);
    if (key > cut) nsel++;
    else if (key == cut) {
      nsel += er < need;
      er++;
    }
  }
  int tot_sel;
  int pos = block_exclusive_scan(nsel, misc, tot_sel);
  er = eq_before;
  for (int i = lo; i < hi; i++) {
    const uint64_t key = (uint64_t)__double_as_longlong(mag[i]);
    bool take = false;
    if (key > cut) take = true;
    else if (key == cut) {
      take = er < need;
      er++;
    }
    if (take && pos < num) out[pos++] = i;
  }
  __syncthreads();
}

}  // namespace dev
}  // namespace sfa
