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

// In-place radix-2 decimation-in-time FFT of N = 2^logN points held in LDS in
// bit-reversed order; output in natural order.  tw[j * tws] = e^{-2 pi i j/N}
// where tws = Nmax / N for a table built for Nmax.  Ends with a barrier.
template <typename T2>
__device__ void lds_fft(T2 *buf, int logN, const T2 *__restrict__ tw, int tws)
{
  const int N = 1 << logN;
  for (int s = 1; s <= logN; ++s) {
    const int h = 1 << (s - 1);
    const int tstride = (N >> s) * tws;
    for (int b = threadIdx.x; b < (N >> 1); b += blockDim.x) {
      const int j = b & (h - 1);
      const int i0 = ((b >> (s - 1)) << s) + j;
      const int i1 = i0 + h;
      const T2 w = tw[j * tstride];
      const T2 u = buf[i0];
      const T2 t = cmul(buf[i1], w);
      buf[i0] = cadd(u, t);
      buf[i1] = csub(u, t);
    }
    __syncthreads();
  }
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
// Uses 256+16 ints of LDS at sh.  Contains barriers.
__device__ void block_top_num(const double *mag, int N, int num, int *out,
                              int *sh)
{
  int *hist = sh;          // 256
  int *misc = sh + 256;    // 16
  const int T = blockDim.x, tid = threadIdx.x;
  // rank of the cutoff in ascending order
  int rank = N - num - 1;
  uint64_t prefix = 0;
  for (int pass = 7; pass >= 0; --pass) {
    const int shift = pass * 8;
    const uint64_t hi_mask = pass == 7 ? 0ull : (~0ull << (shift + 8));
    for (int d = tid; d < 256; d += T) hist[d] = 0;
    __syncthreads();
    for (int i = tid; i < N; i += T) {
      const uint64_t key = (uint64_t)__double_as_longlong(mag[i]);
      if ((key & hi_mask) == prefix) atomicAdd(&hist[(key >> shift) & 255], 1);
    }
    __syncthreads();
    if (tid == 0) {
      int cum = 0, d = 0;
      for (; d < 256; d++) {
        if (rank < cum + hist[d]) break;
        cum += hist[d];
      }
      misc[0] = d;
      misc[1] = rank - cum;
    }
    __syncthreads();
    prefix |= (uint64_t)misc[0] << shift;
    rank = misc[1];
    __syncthreads();
  }
  const uint64_t cut = prefix;
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
