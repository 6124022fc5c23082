// sFFT v3 (exactly sparse) for gfx950: one workgroup per signal runs the
// whole of src/computefourier-3.0.cc:784-1088 -- Mansour filter, two
// Dolph-Chebyshev "Gauss" filters (the second permuted), and the round-robin
// estimate/subtract loop.
//
// Parallel within a signal: the filters (gathers, folds, FFTs in LDS), the
// per-bucket estimation (atan2 / sincos), the result-map updates and the
// subtraction values.  The subtractions themselves reach every bucket in the
// reference's order (found-list order: short passes by one thread per bucket
// array, longer ones by the thread owning the bucket), so every bucket sees
// the same sequence of roundings as in the oracle.  The
// result map is iterated in insertion order (the reference's
// tr1::unordered_map order is unreproducible; see the oracle).
#include <algorithm>
#include <climits>

#include "device_common.hpp"
#include "internal.hpp"

namespace sfa {
namespace k {

using namespace sfa::dev;

namespace {

constexpr int kV3Block = 512;
static_assert((kV3Block & (kV3Block - 1)) == 0, "owner tests take residues by masking");
// Subtraction passes of up to this many pairs run on one thread per bucket
// array; longer ones on bucket-owner threads.
constexpr int kSerialApply = 16;
// Bucket indices an owner thread loads ahead in the subtraction scans.
constexpr int kApplyAhead = 32;
// Steps of a sweep over the Mansour buckets whose loads are issued together.
constexpr int kSweepAhead = 4;
// Loads the pre-kernel's gathers and folds issue together per thread.
constexpr int kPreAhead = 8;
// Pairs per LDS chunk of the Gauss subtractions (fewer fitting than
// kMinStageChunk: indices and values read from memory instead).
constexpr int kStageChunk = 1024, kMinStageChunk = 32;

__device__ __forceinline__ int ilog2v(int v) { return 31 - __clz(v); }

// Sum over the 64 lanes of a wave (every lane gets it).
__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// int products in the reference wrap modulo 2^32 (cf3 381-382, 475, 743)
__device__ __forceinline__ int wmul(int x, int y) { return (int)((unsigned)x * (unsigned)y); }

__device__ __forceinline__ int timesmod(int x, int a, int n)
{
  return (int)(((long long)x * a) % n);
}

// Filter response at signed offset d from a window stored as
// win[d + half], d in [-half, half].
__device__ __forceinline__ double2 win_at(const double2 *win, int half, unsigned dist, int n)
{
  int d = (int)dist;
  if (d > n / 2) d -= n;
  return win[d + half];
}

struct Ctx {
  const V3Launch *a;
  int s, n, W, B1, B2;
  int aa, ai, b, shift, g0, gG;
  double2 *GS0, *GS1, *GP0, *GP1;  // LDS
  double2 *MS;                      // global, interleaved [2i + loop]
  int *keys;
  double2 *vals;
  int *hidx;
  int *cf;
  double2 *cv;
  int *ef;
  double2 *ev;
  double2 *rem;
  int *remi;
  // Mansour buckets written since the last Mansour estimate (a bit per
  // bucket, LDS; null when it does not fit) and the running count of
  // buckets above the loop's power threshold (LDS)
  unsigned *dirty;
  int *cm;
  double2 *stage;  // LDS for the subtractions, room bytes
  size_t room;
};

// The round-robin loop's test on a Mansour bucket (loop-0 value z): power
// of z / W above 1e-6 (cf3 1054-1060).  W is a power of two, so z * (1 / W)
// rounds like z / W.
__device__ __forceinline__ int ms_counted(double2 z, double invW)
{
  const double re = z.x * invW, im = z.y * invW;
  return re * re + im * im > 1e-6;
}

// MS bucket h less (r0, r1), its dirty bit set and the count updated.
__device__ __forceinline__ void ms_sub(const Ctx &c, int h, double2 r0, double2 r1)
{
  const double2 old = c.MS[2 * h];
  const double2 now = csub(old, r0);
  c.MS[2 * h] = now;
  c.MS[2 * h + 1] = csub(c.MS[2 * h + 1], r1);
  if (c.dirty) {
    const double invW = 1.0 / c.W;
    atomicOr(&c.dirty[h >> 5], 1u << (h & 31));
    const int d = ms_counted(now, invW) - ms_counted(old, invW);
    if (d) atomicAdd(c.cm, d);
  }
}

// ---------------------------------------------------------------- map
__device__ __forceinline__ unsigned hslot(int key, int hcap)
{
  return ((unsigned)key * 2654435761u) & (unsigned)(hcap - 1);
}

__device__ int map_find(const Ctx &c, int key)
{
  const int H = c.a->Hcap;
  unsigned h = hslot(key, H);
  for (int probe = 0; probe < H; probe++) {
    const int e = c.hidx[h];
    if (e < 0) return -1;
    if (c.keys[e] == key) return e;
    h = (h + 1) & (unsigned)(H - 1);
  }
  return -1;
}

__device__ void map_insert_slot(const Ctx &c, int key, int e)
{
  const int H = c.a->Hcap;
  unsigned h = hslot(key, H);
  for (int probe = 0; probe < H; probe++) {
    if (atomicCAS(&c.hidx[h], -1, e) == -1) return;
    h = (h + 1) & (unsigned)(H - 1);
  }
}

// ---------------------------------------------------------------- FFT
// Two B-point forward FFTs of a (x0, x1) pair held in LDS in natural order.
__device__ void fft_pair(double2 *x0, double2 *x1, int B, const V3Launch &a)
{
  const int logB = ilog2v(B);
  // bit-reverse in place
  for (int i = threadIdx.x; i < B; i += blockDim.x) {
    const int r = (int)bitrev((unsigned)i, logB);
    if (r > i) {
      double2 t = x0[i];
      x0[i] = x0[r];
      x0[r] = t;
      t = x1[i];
      x1[i] = x1[r];
      x1[r] = t;
    }
  }
  __syncthreads();
  lds_fft(x0, logB, a.tw64, a.tw_nmax / B);
  lds_fft(x1, logB, a.tw64, a.tw_nmax / B);
}

// ---------------------------------------------------------------- removals
// update_gaussian_loops2 (cf3 354-463): what to subtract from buckets
// h, h+1, h-1 of both loops for (key, value); r[0..5] = loop0 x3, loop1 x3.
// update_gauss_loops2 (cf3 353-463) for one located (key, value) in a
// Gauss filter of B buckets at offsets g, g + step: the centre bucket h, the
// three window distances, then (gauss_values) the removal values r[0..2]
// (offset g) and r[3..5] (g + step) for buckets h, h+1, h-1.  Split so the
// window loads of several removals can be issued together.
struct GaussIdx {
  int h;
  unsigned dd[3];
};

__device__ __forceinline__ GaussIdx gauss_index(const Ctx &c, int key, int B)
{
  const int n = c.n;
  const unsigned n1 = (unsigned)n - 1;
  const unsigned nb = (unsigned)n / (unsigned)B;
  GaussIdx x;
  x.h = (int)(key * 1. / nb + 0.5) % B;
  const unsigned dist1 = ((unsigned)x.h * nb - (unsigned)key + (unsigned)n) & n1;
  x.dd[0] = dist1;
  x.dd[1] = (dist1 + (unsigned)n + nb) & n1;
  x.dd[2] = (dist1 + (unsigned)n + (unsigned)n - nb) & n1;
  return x;
}

__device__ __forceinline__ void gauss_values(const Ctx &c, int key, double2 value, int g,
                                             int step, const double2 F[3], double2 *r)
{
  const int n = c.n;
  const double PI2_DIV_N = 2 * M_PI / n;
  const double a1 = value.x, b1 = value.y;
  const int ko = wmul(key, g) % n;
  const int ko2 = wmul(key, (int)((unsigned)g + (unsigned)step)) % n;
  double b2, a2, b22, a22;
  sincos(PI2_DIV_N * ko, &b2, &a2);
  sincos(PI2_DIV_N * ko2, &b22, &a22);
  const double v1r = a1 * a2 - b1 * b2, v1i = a1 * b2 + b1 * a2;
  const double v2r = a1 * a22 - b1 * b22, v2i = a1 * b22 + b1 * a22;
#pragma unroll
  for (int q = 0; q < 3; q++) {
    r[q] = make_double2(v1r * F[q].x - v1i * F[q].y, v1r * F[q].y + v1i * F[q].x);
    r[3 + q] = make_double2(v2r * F[q].x - v2i * F[q].y, v2r * F[q].y + v2i * F[q].x);
  }
}

__device__ void gauss_removal(const Ctx &c, int key, double2 value, int B,
                              const double2 *win, int half, int g, int step,
                              int &h, double2 *r)
{
  const GaussIdx x = gauss_index(c, key, B);
  h = x.h;
  double2 F[3];
#pragma unroll
  for (int q = 0; q < 3; q++) F[q] = win_at(win, half, x.dd[q], c.n);
  gauss_values(c, key, value, g, step, F, r);
}

__device__ void gauss_apply_pair(double2 *S0, double2 *S1, int B, int h, const double2 *r)
{
  const int bk[3] = {h, (h + 1) % B, (h + B - 1) % B};
  for (int q = 0; q < 3; q++) S0[bk[q]] = csub(S0[bk[q]], r[q]);
  for (int q = 0; q < 3; q++) S1[bk[q]] = csub(S1[bk[q]], r[3 + q]);
}

// A few pairs (removal values in LDS: 14 per pair, this filter's at roff;
// bucket index at 3 i + slot) applied by bucket owners: thread t takes its
// buckets t, t + blockDim, ... and walks the pairs in order, applying the
// parts that land on the bucket (h, h+1, h-1 within a pair) -- the order
// in which gauss_apply_pair applies them pair after pair.
__device__ __forceinline__ void gauss_apply_few(double2 *S0, double2 *S1, int B, const int *si,
                                                int slot, const double2 *sv, int roff, int cnt)
{
  const int bm = B - 1;  // B a power of two
  for (int tb = threadIdx.x; tb < B; tb += blockDim.x) {
    double2 a0 = make_double2(0, 0), a1 = make_double2(0, 0);
    bool any = false;
    for (int i = 0; i < cnt; i++) {
      const int h = si[i * 3 + slot];
      if ((unsigned)((tb - h + 1) & bm) > 2u) continue;
      if (!any) {
        a0 = S0[tb];
        a1 = S1[tb];
        any = true;
      }
      const double2 *r = sv + i * 14 + roff;
      if (h == tb) {
        a0 = csub(a0, r[0]);
        a1 = csub(a1, r[3]);
      }
      if (((h + 1) & bm) == tb) {
        a0 = csub(a0, r[1]);
        a1 = csub(a1, r[4]);
      }
      if (((h + bm) & bm) == tb) {
        a0 = csub(a0, r[2]);
        a1 = csub(a1, r[5]);
      }
    }
    if (any) {
      S0[tb] = a0;
      S1[tb] = a1;
    }
  }
}

// Pair i subtracts r[0..2] from buckets h, h+1, h-1 of S0 and r[3..5] from
// those of S1 (in that order; with B <= 2 a bucket can be several of them).
// Thread t owns the buckets b = t mod blockDim and takes the pairs in pair
// order, kApplyAhead bucket indices loaded ahead (each batch costs one
// memory latency).
__device__ void gauss_apply_owned(double2 *__restrict__ S0, double2 *__restrict__ S1, int B,
                                  const int *__restrict__ remi, int slot,
                                  const double2 *__restrict__ rem, int roff, int cnt)
{
  constexpr int T = kV3Block;
  const int t = threadIdx.x;
  if (t >= B) return;
  // B and T are powers of two: the buckets h-1, h, h+1 of a pair can only
  // include one of this thread's (b mod T == t) when (h - t + 1) mod
  // min(B, T) <= 2; only those pairs take the exact test below
  const int M = min(B, T);
  for (int i0 = 0; i0 < cnt; i0 += kApplyAhead) {
    int hh[kApplyAhead];
#pragma unroll
    for (int u = 0; u < kApplyAhead; u++)
      hh[u] = i0 + u < cnt ? remi[(i0 + u) * 3 + slot] : -1;
#pragma unroll
    for (int u = 0; u < kApplyAhead; u++) {
      const int h = hh[u];
      if (h < 0) break;
      if ((unsigned)((h - t + 1) & (M - 1)) > 2u) continue;
      const int bk[3] = {h, (h + 1) % B, (h + B - 1) % B};
#pragma unroll
      for (int q = 0; q < 3; q++) {
        if (bk[q] % T != t) continue;
        const double2 *r = rem + (int64_t)(i0 + u) * 14 + roff;
        S0[bk[q]] = csub(S0[bk[q]], r[q]);
        S1[bk[q]] = csub(S1[bk[q]], r[3 + q]);
      }
    }
  }
}

// LDS bytes the staged application needs per pair of a chunk: 6 removal
// values, the bucket index, three sort entries.
constexpr size_t kStagePairBytes = 6 * sizeof(double2) + sizeof(int) + 3 * sizeof(int);

// As gauss_apply_owned, a chunk of pairs at a time through LDS: the block
// copies the chunk's bucket indices and removal values next to the bucket
// arrays (one memory latency per chunk) and sorts the chunk's 3m entries
// e = 3 u + q (pair u, q = 0, 1, 2 for its buckets h, h+1, h-1) by bucket:
// counts, an exclusive scan, atomic placement, then each bucket's short
// list re-sorted by e.  The owner of a bucket applies its entries in e
// order, which is pair order and h, h+1, h-1 within a pair (cf3 465-482):
// the subtractions a bucket sees and their order are unchanged.  stage
// holds chunk * kStagePairBytes + (B + 16) * 4 bytes of LDS.  Called by
// every thread of the block (barriers inside).
__device__ void gauss_apply_staged(double2 *__restrict__ S0, double2 *__restrict__ S1, int B,
                                   const int *__restrict__ remi, int slot,
                                   const double2 *__restrict__ rem, int roff, int cnt,
                                   double2 *__restrict__ stage, int chunk)
{
  constexpr int T = kV3Block;
  const int t = threadIdx.x;
  const int bm = B - 1;  // B a power of two
  double2 *__restrict__ sv = stage;
  int *__restrict__ sh_idx = reinterpret_cast<int *>(sv + 6 * (size_t)chunk);
  int *__restrict__ list = sh_idx + chunk;
  int *__restrict__ pos = list + 3 * chunk;  // B bucket counts, then ends
  int *__restrict__ sh = pos + B;            // scan scratch
  const int C = (B + T - 1) / T;             // buckets per thread in the scan
  for (int base = 0; base < cnt; base += chunk) {
    const int m = min(chunk, cnt - base);
    __syncthreads();  // the previous chunk is applied
    for (int j = t; j < m; j += T) sh_idx[j] = remi[(int64_t)(base + j) * 3 + slot];
    for (int j = t; j < 6 * m; j += T) {
      const int u = j / 6, q = j - 6 * u;
      sv[j] = rem[(int64_t)(base + u) * 14 + roff + q];
    }
    for (int b = t; b < B; b += T) pos[b] = 0;
    __syncthreads();
    for (int e = t; e < 3 * m; e += T) {
      const int u = e / 3, q = e - 3 * u;
      const int h = sh_idx[u];
      const int b = q == 0 ? h : (q == 1 ? (h + 1) & bm : (h + bm) & bm);
      atomicAdd(&pos[b], 1);
    }
    __syncthreads();
    {
      const int lo = t * C, hi = min(B, lo + C);
      int sum = 0;
      for (int b = lo; b < hi; b++) sum += pos[b];
      int total;
      int run = block_exclusive_scan(sum, sh, total);
      for (int b = lo; b < hi; b++) {
        const int v = pos[b];
        pos[b] = run;  // start of bucket b's entries
        run += v;
      }
    }
    __syncthreads();
    for (int e = t; e < 3 * m; e += T) {
      const int u = e / 3, q = e - 3 * u;
      const int h = sh_idx[u];
      const int b = q == 0 ? h : (q == 1 ? (h + 1) & bm : (h + bm) & bm);
      list[atomicAdd(&pos[b], 1)] = e;  // pos[b] ends at the end of bucket b
    }
    __syncthreads();
    for (int tb = t; tb < B; tb += T) {
      const int lo = tb ? pos[tb - 1] : 0, hi = pos[tb];
      for (int i = lo + 1; i < hi; i++) {  // insertion sort by e (short lists)
        const int v = list[i];
        int j = i - 1;
        while (j >= lo && list[j] > v) {
          list[j + 1] = list[j];
          j--;
        }
        list[j + 1] = v;
      }
      double2 a0 = S0[tb], a1 = S1[tb];
      for (int i = lo; i < hi; i++) {
        const int e = list[i];
        const int u = e / 3, q = e - 3 * u;
        a0 = csub(a0, sv[6 * u + q]);
        a1 = csub(a1, sv[6 * u + 3 + q]);
      }
      S0[tb] = a0;
      S1[tb] = a1;
    }
  }
  __syncthreads();
}

// update_mansour_loops2 (cf3 300-351)
__device__ void mansour_removal(const Ctx &c, int key, double2 value, int &h, double2 *r)
{
  const double PI2_OVER_N = 2 * M_PI / c.n;
  const int W = c.W;
  h = key & (W - 1);
  const double vr = value.x, vi = value.y;
  double s0, c0, s1, c1;
  sincos(PI2_OVER_N * key * c.g0, &s0, &c0);
  sincos(PI2_OVER_N * key * (c.g0 + 1), &s1, &c1);
  r[0] = make_double2(W * (vr * c0 - vi * s0), W * (vr * s0 + vi * c0));
  r[1] = make_double2(W * (vr * c1 - vi * s1), W * (vr * s1 + vi * c1));
}

// Subtraction values of cnt (key, value) pairs, then their application in
// pair order.  which: bit 0 Gauss-1, bit 1 Gauss-2, bit 2 Mansour.
// Slots per pair in rem: [0..5] G1, [6..11] G2, [12..13] Mansour.
// The removal values of cnt (key, value) pairs for the filters in which
// (bit 0 Gauss-1, 1 Gauss-2, 2 Mansour) into r (14 per pair) and ri (3
// bucket indices per pair).  One work item per (filter, pair); each
// filter's pairs start on a wave boundary, so a wave runs one filter's
// code and the filters' sincos and window loads proceed in parallel.
__device__ __forceinline__ void removals(const Ctx &c, const int *keys, const double2 *vals,
                                         int cnt, int which, double2 *r0, int *ri0)
{
  const V3Launch &a = *c.a;
  int fs[3], nf = 0;
  for (int f = 0; f < 3; f++)
    if ((which >> f) & 1) fs[nf++] = f;
  const int stride = (cnt + 63) & ~63;
  for (int t = threadIdx.x; t < nf * stride; t += blockDim.x) {
    const int q = t / stride, i = t - q * stride;
    if (i >= cnt) continue;
    const int f = fs[q];
    const int key = keys[i];
    const double2 v = vals[i];
    double2 *r = r0 + (int64_t)i * 14;
    int *ri = ri0 + (int64_t)i * 3;
    if (f == 0) gauss_removal(c, key, v, c.B1, a.win1, a.win1_half, c.gG, 1, ri[0], r);
    else if (f == 1)
      gauss_removal(c, (timesmod(key, c.ai, c.n) + c.shift) % c.n, v, c.B2, a.win2,
                    a.win2_half, wmul(c.aa, c.gG), 1, ri[1], r + 6);
    else mansour_removal(c, key, v, ri[2], r + 12);
  }
}

__device__ __forceinline__ void subtract(const Ctx &c, const int *keys, const double2 *vals,
                                         int cnt, int which)
{
  const V3Launch &a = *c.a;
  // LDS staging for the applications
  const size_t room = c.room;
  double2 *const sv = c.stage;
  if (cnt <= kSerialApply && !a.serial_all &&
      room >= (size_t)kSerialApply * (14 * sizeof(double2) + 3 * sizeof(int))) {
    // Few pairs: the removal values stay in LDS; one thread per Gauss
    // bucket array applies them in pair order, and each Mansour bucket
    // (in HBM) is updated by its owner thread h mod kV3Block in pair order
    // -- every bucket sees the same subtractions in the same order.
    int *si = reinterpret_cast<int *>(sv + 14 * kSerialApply);
    removals(c, keys, vals, cnt, which, sv, si);
    __syncthreads();
    if (which & 1) gauss_apply_few(c.GS0, c.GS1, c.B1, si, 0, sv, 0, cnt);
    if (which & 2) gauss_apply_few(c.GP0, c.GP1, c.B2, si, 1, sv, 6, cnt);
    if (which & 4) {
      for (int i = 0; i < cnt; i++) {
        const int h = si[i * 3 + 2];
        if ((h & (kV3Block - 1)) != (int)threadIdx.x) continue;
        ms_sub(c, h, sv[i * 14 + 12], sv[i * 14 + 13]);
      }
    }
    __threadfence_block();
    __syncthreads();
    return;
  }
  removals(c, keys, vals, cnt, which, c.rem, c.remi);
  __threadfence_block();
  __syncthreads();
  // One thread per bucket, each taking the pairs in pair order: a bucket
  // sees the same subtractions in the same order as when one thread applies
  // every pair in turn (cf3 465-482), so the results are the same.
  if (cnt <= kSerialApply || c.a->serial_all) {
    // few pairs: one thread per bucket array applies them all
    if (threadIdx.x == 0 && (which & 1)) {
      for (int i = 0; i < cnt; i++) gauss_apply_pair(c.GS0, c.GS1, c.B1, c.remi[i * 3], c.rem + (int64_t)i * 14);
    } else if (threadIdx.x == 64 && (which & 2)) {
      for (int i = 0; i < cnt; i++)
        gauss_apply_pair(c.GP0, c.GP1, c.B2, c.remi[i * 3 + 1], c.rem + (int64_t)i * 14 + 6);
    } else if (threadIdx.x == 128 && (which & 4)) {
      for (int i = 0; i < cnt; i++) {
        const int h = c.remi[i * 3 + 2];
        const double2 *r = c.rem + (int64_t)i * 14 + 12;
        ms_sub(c, h, r[0], r[1]);
      }
    }
    __threadfence_block();
    __syncthreads();
    return;
  }
  {
    // staged in LDS (kStagePairBytes per pair plus the bucket counts)
    const size_t fixed = (size_t)(max(c.B1, c.B2) + 16) * sizeof(int);
    const int chunk =
        room > fixed ? (int)min((size_t)kStageChunk, (room - fixed) / kStagePairBytes) : 0;
    if (chunk >= kMinStageChunk) {
      if (which & 1) gauss_apply_staged(c.GS0, c.GS1, c.B1, c.remi, 0, c.rem, 0, cnt, sv, chunk);
      if (which & 2) gauss_apply_staged(c.GP0, c.GP1, c.B2, c.remi, 1, c.rem, 6, cnt, sv, chunk);
    } else {
      if (which & 1) gauss_apply_owned(c.GS0, c.GS1, c.B1, c.remi, 0, c.rem, 0, cnt);
      if (which & 2) gauss_apply_owned(c.GP0, c.GP1, c.B2, c.remi, 1, c.rem, 6, cnt);
    }
  }
  if (which & 4) {
    constexpr int T = kV3Block;
    const int t = threadIdx.x;
    const int *__restrict__ remi = c.remi;
    const double2 *__restrict__ rem = c.rem;
    for (int i0 = 0; i0 < cnt; i0 += kApplyAhead) {
      int hh[kApplyAhead];
#pragma unroll
      for (int u = 0; u < kApplyAhead; u++)
        hh[u] = i0 + u < cnt ? remi[(i0 + u) * 3 + 2] : -1;
#pragma unroll
      for (int u = 0; u < kApplyAhead; u++) {
        const int h = hh[u];
        if (h < 0) break;
        if ((h & (T - 1)) != t) continue;  // h = key mod W >= 0
        const double2 *r = rem + (int64_t)(i0 + u) * 14 + 12;
        ms_sub(c, h, r[0], r[1]);
      }
    }
  }
  __threadfence_block();
  __syncthreads();
}

// ---------------------------------------------------------------- estimation
// An estimate's accepted buckets go to ef/ev in ascending bucket order (the
// order the reference appends them in, hence the map's insertion order),
// without a staging slot per bucket.

// One step of kV3Block buckets (bucket = step start + threadIdx.x) whose
// results are in registers: the accepted ones are written at base + their
// rank among the step's accepted buckets.  Returns the step's count
// (uniform).  Uses NW ints at sh; contains barriers.
__device__ __forceinline__ int emit_step(const Ctx &c, int base, bool ok, int out, double2 val,
                                         int *sh)
{
  constexpr int NW = kV3Block / 64;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint64_t bal = __ballot(ok);
  if (lane == 0) sh[wid] = __popcll(bal);
  __syncthreads();
  int off = base, tot = 0;
#pragma unroll
  for (int q = 0; q < NW; q++) {
    const int t = sh[q];
    off += q < wid ? t : 0;
    tot += t;
  }
  __syncthreads();
  if (ok) {
    const int pos = off + __popcll(bal & ((1ull << lane) - 1));
    c.ef[pos] = out;
    c.ev[pos] = val;
  }
  return tot;
}

// estimate_freq_gauss_loops2 (cf3 484-640) on (S0, S1), B buckets.
__device__ __forceinline__ int est_gauss(const Ctx &c, const double2 *S0, const double2 *S1, int B, int a_,
                         int b_, const double2 *win, int half, int *sh)
{
  const int n = c.n;
  const unsigned n1 = (unsigned)n - 1, B1m = (unsigned)B - 1;
  const double PI2 = 2 * M_PI;
  const double N_OVER_PI2 = n / PI2;
  const double PI2_A_OFFSET_OVER_N = PI2 * a_ * c.gG / n;
  const double BUCKETS_OVER_N = (double)B / n;
  const int N_OVER_BUCKETS = n / B;
  int found = 0;
  for (int s0 = 0; s0 < B; s0 += blockDim.x) {
    const int bk = s0 + threadIdx.x;
    int out = -1;
    double2 val = make_double2(0, 0);
    if (bk < B) {
      const double a0 = S0[bk].x, b0 = S0[bk].y, a1 = S1[bk].x, b1 = S1[bk].y;
      const double zb = (a0 * a0 + b0 * b0) + (a1 * a1 + b1 * b1);
      if (zb > 1e-8) {
        const double c0 = a0 * a0 + b0 * b0, c1 = a1 * a1 + b1 * b1;
        const double d0 = atan2(b0, a0), d1 = atan2(b1, a1);
        const double eb = c1 / c0 - 1;
        double err = eb * eb;
        err /= n;
        if (err < 1e-12 && c0 > 0.01) {
          const double slope = (d1 - d0) * 1.0;
          int freq = (int)rint(N_OVER_PI2 * slope) + n;
          freq = (int)((unsigned)freq & n1);
          const unsigned hashed = (unsigned)(long long)rint(freq * BUCKETS_OVER_N) & B1m;
          if (hashed == (unsigned)bk) {
            const double ph = d0 - PI2_A_OFFSET_OVER_N * freq;
            const double mag = sqrt(c0);
            double sn, cs;
            sincos(ph, &sn, &cs);
            const double mr = mag * cs, mi = mag * sn;
            const unsigned dist =
                (hashed * (unsigned)N_OVER_BUCKETS - (unsigned)freq + (unsigned)n) & n1;
            const double2 F = win_at(win, half, dist, n);
            const double den = F.x * F.x + F.y * F.y;
            val = make_double2((mr * F.x + mi * F.y) / den, (mi * F.x - mr * F.y) / den);
            out = (int)(((unsigned)(timesmod(freq, a_, n) - b_ + n)) & n1);
          }
        }
      }
    }
    found += emit_step(c, found, out >= 0, out, val, sh);
  }
  __threadfence_block();
  __syncthreads();
  return found;
}

// estimate_freq_mansour_loops2 (cf3 642-776) for bucket bk: the acceptance
// test, and with Full the frequency and the value.  The test needs only the
// bucket powers, so a sweep can evaluate it alone and the full estimate
// later for the accepted buckets (the same expressions, so the same bits).
struct MansourConst {
  double N_OVER_PI2, PI2_OVER_N, NORM, NORM2, err_max;
  unsigned FREQ_MASK;
};

template <bool Full>
__device__ __forceinline__ bool mansour_bucket(const Ctx &c, const MansourConst &m, int bk,
                                               double2 s0, double2 s1, int &out, double2 &val)
{
  const double a0 = s0.x, b0 = s0.y, a1 = s1.x, b1 = s1.y;
  const double zb = (a0 * a0 + b0 * b0) * m.NORM2 + (a1 * a1 + b1 * b1) * m.NORM2;
  if (!(zb > 1e-8)) return false;
  const double c0 = (a0 * a0 + b0 * b0) * m.NORM2;
  const double c1 = (a1 * a1 + b1 * b1) * m.NORM2;
  double median_abs = c0;
  const double inv = 1. / median_abs;
  const double bb = (c1 * inv) - 1;
  const double err = bb * bb;
  if (!(err < m.err_max && median_abs > 0.01)) return false;
  if (Full) {
    const double d0 = atan2(b0 * m.NORM, a0 * m.NORM), d1 = atan2(b1 * m.NORM, a1 * m.NORM);
    const double slope = d1 - d0;
    const int f1 = (int)rint(slope * m.N_OVER_PI2);
    const unsigned f2 = (unsigned)f1 & m.FREQ_MASK;
    const int f3 = (int)(f2 | (unsigned)bk);
    const int fo = wmul(f3, c.g0);
    const double ph = d0 - m.PI2_OVER_N * fo;
    median_abs = sqrt(median_abs);
    double sn, cs;
    sincos(ph, &sn, &cs);
    out = f3;
    val = make_double2(median_abs * cs, median_abs * sn);
  }
  return true;
}

// The W Mansour buckets swept in segments of 64 steps of kV3Block buckets
// (bucket = segment start + step kV3Block + threadIdx.x).  The first sweep
// evaluates only the acceptance test, on kSweepAhead steps' buckets loaded
// together (one memory latency for them): each thread keeps its accepted
// steps as bits, each wave writes the count of a step to rank[step NW +
// wave]; a
// block scan over rank (which is bucket order) gives every (step, wave) its
// first slot; the second sweep estimates the accepted buckets and writes
// each at that slot plus its accepted lanes below.  rank holds kV3Block
// ints of LDS.
__device__ __forceinline__ int est_mansour(const Ctx &c, int *rank, int *sh)
{
  constexpr int T = kV3Block, NW = T / 64, SEG = 64;
  const int n = c.n, W = c.W;
  const double PI2 = 2 * M_PI;
  MansourConst m;
  m.N_OVER_PI2 = n / PI2;
  m.PI2_OVER_N = PI2 / n;
  m.FREQ_MASK = ((unsigned)n - 1) & ~((unsigned)W - 1);
  m.NORM = 1. / W;
  m.NORM2 = m.NORM * m.NORM;
  m.err_max = n * 1e-10;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int found = 0;
  for (int s0 = 0; s0 < W; s0 += SEG * T) {
    const int steps = min(SEG, (W - s0 + T - 1) / T);
    uint64_t mine = 0;
    for (int st0 = 0; st0 < steps; st0 += kSweepAhead) {
      double2 z0[kSweepAhead], z1[kSweepAhead];
#pragma unroll
      for (int u = 0; u < kSweepAhead; u++) {
        // clamped: past the end the loads repeat bucket W - 1 unused
        const int bk = min(s0 + (st0 + u) * T + (int)threadIdx.x, W - 1);
        z0[u] = c.MS[2 * bk];
        z1[u] = c.MS[2 * bk + 1];
      }
#pragma unroll
      for (int u = 0; u < kSweepAhead; u++) {
        const int st = st0 + u;
        if (st >= steps) break;
        const int bk = s0 + st * T + threadIdx.x;
        int o;
        double2 v;
        const bool ok = bk < W && mansour_bucket<false>(c, m, bk, z0[u], z1[u], o, v);
        mine |= (uint64_t)ok << st;
        const uint64_t bal = __ballot(ok);
        if (lane == 0) rank[st * NW + wid] = __popcll(bal);
      }
    }
    __syncthreads();
    const int nr = steps * NW;
    int total;
    const int first = block_exclusive_scan(threadIdx.x < nr ? rank[threadIdx.x] : 0, sh, total);
    if (threadIdx.x < nr) rank[threadIdx.x] = found + first;
    __syncthreads();
    for (int st = 0; st < steps; st++) {
      const bool ok = (mine >> st) & 1;
      const uint64_t bal = __ballot(ok);
      if (ok) {
        const int bk = s0 + st * T + threadIdx.x;
        int o = -1;
        double2 v = make_double2(0, 0);
        (void)mansour_bucket<true>(c, m, bk, c.MS[2 * bk], c.MS[2 * bk + 1], o, v);
        const int pos = rank[st * NW + wid] + __popcll(bal & ((1ull << lane) - 1));
        c.ef[pos] = o;
        c.ev[pos] = v;
      }
    }
    found += total;
    __syncthreads();  // rank is rewritten by the next segment
  }
  __threadfence_block();
  __syncthreads();
  return found;
}

// est_mansour over the buckets written since the previous Mansour estimate
// only: a bucket left unwritten holds the values that estimate rejected (an
// accepted bucket is always written next, by the subtraction of what it
// found), so it is rejected again and the result is the full sweep's.
// Thread t takes bit word w = t (+ kV3Block ...): buckets 32 w .. 32 w + 31
// in ascending order, so a block scan over the threads' accepted counts
// ranks the accepted buckets in bucket order.  Clears the bits it reads.
__device__ __forceinline__ int est_mansour_dirty(const Ctx &c, int *sh)
{
  const int n = c.n, W = c.W;
  const double PI2 = 2 * M_PI;
  MansourConst m;
  m.N_OVER_PI2 = n / PI2;
  m.PI2_OVER_N = PI2 / n;
  m.FREQ_MASK = ((unsigned)n - 1) & ~((unsigned)W - 1);
  m.NORM = 1. / W;
  m.NORM2 = m.NORM * m.NORM;
  m.err_max = n * 1e-10;
  const int words = (W + 31) >> 5;
  int found = 0;
  for (int w0 = 0; w0 < words; w0 += blockDim.x) {
    const int wd = w0 + threadIdx.x;
    unsigned bits = 0, acc = 0;
    if (wd < words) {
      bits = c.dirty[wd];
      c.dirty[wd] = 0u;
    }
    for (unsigned b = bits; b; b &= b - 1) {
      const int bk = wd * 32 + __ffs(b) - 1;
      int o;
      double2 v;
      if (mansour_bucket<false>(c, m, bk, c.MS[2 * bk], c.MS[2 * bk + 1], o, v))
        acc |= b & (~b + 1u);
    }
    int total;
    int pos = found + block_exclusive_scan(__popc(acc), sh, total);
    for (; acc; acc &= acc - 1) {
      const int bk = wd * 32 + __ffs(acc) - 1;
      int o = -1;
      double2 v = make_double2(0, 0);
      (void)mansour_bucket<true>(c, m, bk, c.MS[2 * bk], c.MS[2 * bk + 1], o, v);
      c.ef[pos] = o;
      c.ev[pos] = v;
      pos++;
    }
    found += total;
  }
  __threadfence_block();
  __syncthreads();
  return found;
}

// Adds the F found (ef, ev) to the map (new keys appended in found order).
// Returns false when the map would overflow.
__device__ __forceinline__ bool map_add_found(const Ctx &c, int F, int *sh, int *s_size)
{
  const int T = blockDim.x;
  const int C = (F + T - 1) / T;
  const int lo = threadIdx.x * C, hi = min(F, lo + C);
  int nnew = 0;
  for (int i = lo; i < hi; i++) {
    const int e = map_find(c, c.ef[i]);
    c.remi[i * 3] = e;  // remember (reused as scratch before removals)
    nnew += e < 0;
  }
  int total;
  int pos = block_exclusive_scan(nnew, sh, total);
  const int base = *s_size;
  if (base + total > c.a->map_limit) return false;
  for (int i = lo; i < hi; i++) {
    int e = c.remi[i * 3];
    if (e < 0) {
      e = base + pos++;
      c.keys[e] = c.ef[i];
      c.vals[e] = make_double2(0.0, 0.0);
      map_insert_slot(c, c.ef[i], e);
    }
    // ans[f] = ans[f] + v (distinct keys within one pass)
    c.vals[e] = cadd(c.vals[e], c.ev[i]);
  }
  __threadfence_block();
  __syncthreads();
  if (threadIdx.x == 0) *s_size = base + total;
  __syncthreads();
  return true;
}

}  // namespace

// The whole of one signal's transform, in one workgroup.
// ---------------------------------------------------------------- refine
// The refine form re-estimates the values of the located frequencies after
// the reference's loop, keeping the located set.  The Mansour filter is an
// exact aliasing: at offset g its bucket j is
//   M_g[j] = W * sum over f = j (mod W) of X[f] e^{2 pi i f g / n},
// so fresh buckets at the offsets g0 + l t (l < kRefOff, t = ref_t) give,
// per bucket, kRefOff equations in the values of the frequencies of that
// residue class.  Members f = j + W q enter through the nodes
// e^{2 pi i q t / (n/W)}; with t odd near 0.618 n/W, members a few W apart
// (where the reference's false positives sit) get well separated nodes,
// which consecutive offsets (t = 1) would crowd within 2 pi m / (n/W).
// With the located members of bucket j (at most kRefMax) as unknowns the
// least-squares solution (QR) is exact up to rounding; every phase is reduced in
// integers, f (g0 + l t) mod n, before its sincos, instead of the reference's
// FP64 product 2 pi a g f / n (cf3 497, 591-593), whose rounding reaches a
// quarter radian at n = 2^24.  A bucket whose residual stays large holds
// unlocated tones: after annihilating the located members the sequence d_l
// is a sum of r exponentials w_u^l, w_u = e^{2 pi i f_u t / n}; for r = 1
// w_u = d_{l+1} / d_l, for r = 2 the roots of the linear predictor
// d_{l+2} + c1 d_{l+1} + c0 d_l = 0 (a Prony step).  Each w_u is snapped to
// the residue class and added as an unknown; if the residual is still large
// the bucket keeps the reference's values (status bit 8).
constexpr int kRefOff = kV3RefineOffsets;  // Mansour offsets: equations per bucket
constexpr int kRefMax = 7;  // unknowns per bucket (located + unlocated)

__device__ __forceinline__ double2 cis_mod(long long r, int n)
{
  double sn, cs;
  sincos(2 * M_PI * (double)r / (double)n, &sn, &cs);
  return make_double2(cs, sn);
}

__device__ __forceinline__ double2 cconjmul(double2 a, double2 b)  // conj(a) * b
{
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

__device__ __forceinline__ double2 csqrt_d(double2 z)
{
  const double r = hypot(z.x, z.y);
  const double re = sqrt(0.5 * (r + z.x)), im = sqrt(0.5 * (r - z.x));
  return make_double2(re, z.y < 0 ? -im : im);
}

// The frequency of residue class j whose node e^{2 pi i f t / n} is nearest
// the angle ang: f t = j t + W (q t mod n/W) (mod n), so the nearest
// multiple of W of (n ang / 2 pi - j t) mod n gives q t mod n/W, and tinv
// undoes the stride.
__device__ __forceinline__ long long ref_snap(double ang, int j, int n, int W, int t, int tinv)
{
  const long long fe = llrint(ang * n / (2 * M_PI));
  long long r = (fe - (long long)j * t) % n;
  if (r < 0) r += n;
  const long long np1 = n / W - 1;
  const long long qt = llrint((double)r / W) & np1;
  return j + (long long)W * ((qt * tinv) & np1);
}

// Row l of signal s's refine buckets: the offset g0 + l t (rows 0, 2, 3,
// ... of the Mansour pre-pass above the LDS size, whose row 1 is the
// reference's g0 + 1; else ref_ms, written by the refine pre tasks).
__device__ __forceinline__ const double2 *ref_row(const V3Launch &a, int s, int l)
{
  return a.ms_pre ? a.ms_pre + ((int64_t)s * a.ms_loops + (l ? l + 1 : 0)) * a.W
                  : a.ref_ms + ((int64_t)s * kRefOff + l) * a.W;
}

// ---------------------------------------------------------------- refine solve
// Each non-empty residue class is solved by a group of kRefOff lanes, lane
// l holding offset l: its bucket value, its entries of the columns, its row
// of R.  Every sum over the offsets is taken in offset order from zero,
// term by term, by every lane of the group (gsum*), so every lane holds
// the same sums -- those of one thread looping over the offsets -- and the
// group's control flow stays uniform.
__device__ __forceinline__ double gsum(double v, int lim = kRefOff)
{
  double acc = 0;
#pragma unroll
  for (int l = 0; l < kRefOff; l++) {
    const double t = __shfl(v, l, kRefOff);
    if (l < lim) acc += t;
  }
  return acc;
}

__device__ __forceinline__ double2 gsum2(double2 v, int lim = kRefOff)
{
  double2 acc = make_double2(0, 0);
#pragma unroll
  for (int l = 0; l < kRefOff; l++) {
    const double tx = __shfl(v.x, l, kRefOff), ty = __shfl(v.y, l, kRefOff);
    if (l < lim) acc = cadd(acc, make_double2(tx, ty));
  }
  return acc;
}

// 0 - t_0 - t_1 - ... over lanes l < lim (a subtracted sum)
__device__ __forceinline__ double2 gsub2(double2 v, int lim)
{
  double2 acc = make_double2(0, 0);
#pragma unroll
  for (int l = 0; l < kRefOff; l++) {
    const double tx = __shfl(v.x, l, kRefOff), ty = __shfl(v.y, l, kRefOff);
    if (l < lim) acc = csub(acc, make_double2(tx, ty));
  }
  return acc;
}

__device__ __forceinline__ double2 gshfl2(double2 v, int src)
{
  return make_double2(__shfl(v.x, src, kRefOff), __shfl(v.y, src, kRefOff));
}

// Least squares over the kRefOff equations: y_l ~ sum_m X[m] e[m]_l, pc <=
// kRefMax unknowns, by modified Gram-Schmidt QR of the columns (samples of
// unit exponentials: QR keeps their conditioning instead of squaring it).
// e[m] is this lane's entry of column m, y its bucket value.  Returns the
// squared residual (1e300 when the columns are dependent); X on every lane.
__device__ __forceinline__ double ref_solve_lanes(const double2 (&e)[kRefMax], double2 y,
                                                  int pc, double2 (&X)[kRefMax], int lane)
{
  double2 Q[kRefMax], Rrow[kRefMax], z[kRefMax];  // Rrow: row `lane` of R
#pragma unroll
  for (int m = 0; m < kRefMax; m++) {
    Rrow[m] = make_double2(0, 0);
    Q[m] = make_double2(0, 0);
    z[m] = make_double2(0, 0);
  }
#pragma unroll
  for (int m = 0; m < kRefMax; m++) {
    if (m >= pc) break;
    Q[m] = e[m];
#pragma unroll
    for (int q = 0; q < m; q++) {
      const double2 d = gsum2(cconjmul(Q[q], Q[m]));
      if (lane == q) Rrow[m] = d;
      Q[m] = csub(Q[m], cmul(d, Q[q]));
    }
    const double nn = gsum(Q[m].x * Q[m].x + Q[m].y * Q[m].y);
    if (!(nn > 1e-24)) return 1e300;
    const double inv = 1.0 / sqrt(nn);
    if (lane == m) Rrow[m] = make_double2(sqrt(nn), 0);
    Q[m] = make_double2(Q[m].x * inv, Q[m].y * inv);
  }
#pragma unroll
  for (int m = 0; m < kRefMax; m++)
    if (m < pc) z[m] = gsum2(cconjmul(Q[m], y));
  // back substitution: row m lives on lane m, X[m] is then broadcast
#pragma unroll
  for (int m = kRefMax - 1; m >= 0; m--) {
    if (m >= pc) continue;
    double2 acc = z[m];
#pragma unroll
    for (int q = m + 1; q < kRefMax; q++)
      if (q < pc) acc = csub(acc, cmul(Rrow[q], X[q]));
    const double2 xm = make_double2(acc.x / Rrow[m].x, acc.y / Rrow[m].x);
    X[m] = gshfl2(xm, m);
  }
  double2 r = y;
#pragma unroll
  for (int m = 0; m < kRefMax; m++)
    if (m < pc) r = csub(r, cmul(X[m], e[m]));
  return gsum(r.x * r.x + r.y * r.y);
}

// The refine solve of class j of signal s by a lane group (lane = this
// lane's offset): members from the lists head/next (entries of keys/vals),
// values written to vals.  Returns 1 when the bucket keeps the reference's
// values.
__device__ __forceinline__ int refine_class(const V3Launch &a, int s, int g0, int j,
                                            const int *head, const int *next, const int *keys,
                                            double2 *vals, int lane)
{
  const int n = a.n, W = a.W, t = a.ref_t, tinv = a.ref_tinv;
  const unsigned mask = (unsigned)n - 1;
  int mem[kRefMax];
  long long fk[kRefMax];
#pragma unroll
  for (int m = 0; m < kRefMax; m++) mem[m] = 0;  // entry 0 exists (p >= 1 below)
  int p = 0;
  bool over = false;
  for (int e = head[j]; e >= 0; e = next[e]) {
    if (p == kRefMax) {
      over = true;
      break;
    }
#pragma unroll
    for (int m = 0; m < kRefMax; m++)
      if (m == p) mem[m] = e;
    p++;
  }
  if (p == 0) return 0;
  if (over || p >= kRefOff) return 1;  // no spare equation: keep the reference values
  // members by frequency (distinct keys: the order insertion sorting gives)
#pragma unroll
  for (int m = 0; m < kRefMax; m++) fk[m] = m < p ? (long long)keys[mem[m]] : LLONG_MAX;
#pragma unroll
  for (int round = 0; round < kRefMax; round++)
#pragma unroll
    for (int i = round & 1; i + 1 < kRefMax; i += 2)
      if (fk[i] > fk[i + 1]) {
        const long long tf = fk[i];
        fk[i] = fk[i + 1];
        fk[i + 1] = tf;
        const int tm = mem[i];
        mem[i] = mem[i + 1];
        mem[i + 1] = tm;
      }
  const double2 mv = ref_row(a, s, lane)[j];
  const double2 y = make_double2(mv.x / W, mv.y / W);
  const double scale = gsum(y.x * y.x + y.y * y.y);
  double2 e[kRefMax], w[kRefMax], X[kRefMax];
#pragma unroll
  for (int m = 0; m < kRefMax; m++) {
    e[m] = make_double2(0, 0);
    w[m] = make_double2(0, 0);
    X[m] = make_double2(0, 0);
    if (m < p) {
      e[m] = cis_mod((fk[m] * ((long long)g0 + (long long)lane * t)) & mask, n);
      w[m] = cis_mod((fk[m] * t) & mask, n);
    }
  }
  const double tol = 1e-20 * (scale > 1e-8 ? scale : 1e-8);
  double res = ref_solve_lanes(e, y, p, X, lane);
  if (res > tol) {
    // annihilate the members' nodes: d_l <- d_{l+1} - w_m d_l, once per member
    double2 d = y;
    int len = kRefOff;
#pragma unroll
    for (int m = 0; m < kRefMax; m++) {
      if (m >= p) break;
      const double2 dn = gshfl2(d, lane + 1 < kRefOff ? lane + 1 : lane);
      if (lane + 1 < len) d = csub(dn, cmul(w[m], d));
      len--;
    }
    double2 X2[kRefMax];
    if (p + 1 <= kRefMax && len >= 2) {  // one unlocated tone: w_u = d_{l+1} / d_l
      const double2 d1 = gshfl2(d, lane + 1 < kRefOff ? lane + 1 : lane);
      const double2 num = gsum2(cconjmul(d, d1), len - 1);
      const double den = gsum(d.x * d.x + d.y * d.y, len - 1);
      if (den > 0) {
        const long long u = ref_snap(atan2(num.y, num.x), j, n, W, t, tinv);
#pragma unroll
        for (int m = 0; m < kRefMax; m++)
          if (m == p) e[m] = cis_mod((u * ((long long)g0 + (long long)lane * t)) & mask, n);
        const double res2 = ref_solve_lanes(e, y, p + 1, X2, lane);
        if (res2 <= tol) {
          res = res2;
#pragma unroll
          for (int m = 0; m < kRefMax; m++)
            if (m < p) X[m] = X2[m];
        }
      }
    }
    if (res > tol && p + 2 <= kRefMax && len >= 4) {
      // two unlocated tones: least-squares predictor over l < len - 2,
      // normal equations of the columns (d_{l+1}, d_l) against -d_{l+2}
      const int lim = len - 2;
      const double2 d1 = gshfl2(d, lane + 1 < kRefOff ? lane + 1 : lane);
      const double2 d2 = gshfl2(d, lane + 2 < kRefOff ? lane + 2 : lane);
      const double g11 = gsum(d1.x * d1.x + d1.y * d1.y, lim);
      const double g22 = gsum(d.x * d.x + d.y * d.y, lim);
      const double2 g12 = gsum2(cconjmul(d1, d), lim);
      const double2 b1 = gsub2(cconjmul(d1, d2), lim);
      const double2 b2 = gsub2(cconjmul(d, d2), lim);
      const double det = g11 * g22 - (g12.x * g12.x + g12.y * g12.y);
      if (det > 0) {
        const double2 g21 = make_double2(g12.x, -g12.y);
        const double2 n1 = csub(make_double2(g22 * b1.x, g22 * b1.y), cmul(g12, b2));
        const double2 n0 = csub(make_double2(g11 * b2.x, g11 * b2.y), cmul(g21, b1));
        const double2 c1 = make_double2(n1.x / det, n1.y / det);
        const double2 c0 = make_double2(n0.x / det, n0.y / det);
        const double2 sq = csqrt_d(csub(cmul(c1, c1), make_double2(4 * c0.x, 4 * c0.y)));
        const double2 z1 = make_double2(0.5 * (sq.x - c1.x), 0.5 * (sq.y - c1.y));
        const double2 z2 = make_double2(-0.5 * (sq.x + c1.x), -0.5 * (sq.y + c1.y));
        const long long u1 = ref_snap(atan2(z1.y, z1.x), j, n, W, t, tinv);
        const long long u2 = ref_snap(atan2(z2.y, z2.x), j, n, W, t, tinv);
        if (u1 != u2) {
#pragma unroll
          for (int m = 0; m < kRefMax; m++) {
            if (m == p) e[m] = cis_mod((u1 * ((long long)g0 + (long long)lane * t)) & mask, n);
            if (m == p + 1) e[m] = cis_mod((u2 * ((long long)g0 + (long long)lane * t)) & mask, n);
          }
          const double res2 = ref_solve_lanes(e, y, p + 2, X2, lane);
          if (res2 <= tol) {
            res = res2;
#pragma unroll
            for (int m = 0; m < kRefMax; m++)
              if (m < p) X[m] = X2[m];
          }
        }
      }
    }
  }
  if (res > tol) return 1;
  if (lane == 0) {
#pragma unroll
    for (int m = 0; m < kRefMax; m++)
      if (m < p) vals[mem[m]] = X[m];
  }
  return 0;
}

__device__ void write_result(const V3Launch &a, int s, const int *keys, const double2 *vals,
                             int M, int flags, double2 *lds);
__device__ __forceinline__ void pre_task(const V3Launch &a, const int s, const int task,
                                         double2 *lds, const size_t bytes, double2 *scr,
                                         const size_t scr_bytes);

// Claim word of a pre task (per block and task 0..3): free, being computed
// by the workgroup that claimed it, done.
constexpr int kClaimFree = 0, kClaimBusy = 1, kClaimDone = 2;
// Polls before a waiting transform gives up on a busy task (status bit 16).
// A task is only ever busy in a running workgroup, which finishes it, so
// this is a guard against a defect, not a path of normal runs.
constexpr int kClaimSpins = 1 << 20;

// Marks a finished pre task done.  The workgroup barrier orders every
// thread's stores before lane 0's device-scope release (the L2 written
// back once), then the flag.  All threads; contains a barrier.
__device__ __forceinline__ void publish_task(int *cl)
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_store(cl, kClaimDone, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// The results of pre tasks t0..t1 of signal s, before the transform reads
// them.  Lane 0 claims every task no workgroup has started; the block runs
// those here in turn (with pre_task's LDS arguments; no one else reads
// them, so they stay claimed, not published); lane 0 then waits for the
// others with plain polls and acquires once (one cache invalidation).
// False if a wait gave up.  All threads; contains barriers.
__device__ __forceinline__ bool need_tasks(const V3Launch &a, int s, int *cl, int t0, int t1,
                                           double2 *lds, size_t bytes, double2 *scr,
                                           size_t scr_bytes, int *sh)
{
  if (threadIdx.x == 0) {
    int mine = 0;
    for (int t = t0; t <= t1; t++)
      if (atomicCAS(cl + t, kClaimFree, kClaimBusy) == kClaimFree) mine |= 1 << t;
    *sh = mine;
  }
  __syncthreads();
  const int mine = *sh;
  __syncthreads();
  for (int t = t0; t <= t1; t++)
    if ((mine >> t) & 1) {
      pre_task(a, s, t, lds, bytes, scr, scr_bytes);
      __syncthreads();
    }
  if (threadIdx.x == 0) {
    int ok = 1;
    bool others = false;
    for (int t = t0; t <= t1; t++) {
      if ((mine >> t) & 1) continue;
      others = true;
      int spins = 0;
      while (__hip_atomic_load(cl + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != kClaimDone)
        if (++spins >= kClaimSpins) {
          ok = 0;
          break;
        } else {
          __builtin_amdgcn_s_sleep(2);
        }
    }
    if (others) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    *sh = ok;
  }
  __syncthreads();
  const bool ok = *sh;
  __syncthreads();  // sh is reused
  return ok;
}

__device__ __forceinline__ void v3_signal(const V3Launch &a, const int s, int *claims)
{
  extern __shared__ double2 lds[];
  __shared__ int sh[16];
  __shared__ int s_rank[kV3Block];
  __shared__ int s_size, s_flag, s_cnt[3], s_cm, s_claim;
  // the Mansour buckets (pre tasks 0, 1) first: a task run here uses the
  // LDS from its start
  // (and Gauss-1, task 2, which is done by then too: one wait for all three)
  bool waits_ok = need_tasks(a, s, claims, 0, 2, lds, a.lds_bytes, nullptr, 0, &s_claim);
  Ctx c;
  c.a = &a;
  c.s = s;
  c.n = a.n;
  c.W = a.W;
  c.B1 = a.B1;
  c.B2 = a.B2;
  const int *prm = a.params + (int64_t)s * 6;
  c.aa = prm[0];
  c.ai = prm[1];
  c.b = prm[2];
  c.shift = prm[3];
  c.g0 = prm[4];
  c.gG = prm[5];
  c.GS0 = lds;
  c.GS1 = lds + a.B1;
  c.GP0 = lds + 2 * a.B1;
  c.GP1 = lds + 2 * a.B1 + a.B2;
  c.MS = a.ms + (int64_t)s * 2 * a.W;
  c.keys = a.map_keys + (int64_t)s * a.Mcap;
  c.vals = a.map_vals + (int64_t)s * a.Mcap;
  c.hidx = a.map_hidx + (int64_t)s * a.Hcap;
  c.cf = a.cf + (int64_t)s * a.Fcap;
  c.cv = a.cv + (int64_t)s * a.Fcap;
  c.ef = a.ef + (int64_t)s * a.Fcap;
  c.ev = a.ev + (int64_t)s * a.Fcap;
  c.rem = a.rem + (int64_t)s * a.Rcap * 14;
  c.remi = a.remi + (int64_t)s * a.Rcap * 3;
  const int W = a.W;
  {
    // LDS past the four Gauss bucket arrays: the dirty bits (when they
    // leave room for the subtractions' smallest staging), then staging
    const size_t gauss = (2 * (size_t)a.B1 + 2 * (size_t)a.B2) * sizeof(double2);
    const size_t bits = (((size_t)W + 31) / 32 * sizeof(unsigned) + 15) / 16 * 16;
    const size_t rest = a.lds_bytes > gauss ? a.lds_bytes - gauss : 0;
    const bool track = rest >= bits + (size_t)kMinStageChunk * 112 + 4 * ((size_t)max(a.B1, a.B2) + 16);
    c.dirty = track ? reinterpret_cast<unsigned *>(lds + 2 * a.B1 + 2 * a.B2) : nullptr;
    c.cm = &s_cm;
    c.stage = lds + 2 * a.B1 + 2 * a.B2 + (track ? bits / sizeof(double2) : 0);
    c.room = rest - (track ? bits : 0);
    if (track)
      for (int i = threadIdx.x; i < (W + 31) / 32; i += blockDim.x) c.dirty[i] = 0u;
  }
  for (int i = threadIdx.x; i < a.Hcap; i += blockDim.x) c.hidx[i] = -1;
  if (threadIdx.x == 0) {
    s_size = 0;
    s_flag = 0;
  }
  __threadfence_block();
  __syncthreads();

  // ---- Mansour filter (cf3 96-127): buckets from pre tasks 0, 1
  int F = est_mansour(c, s_rank, sh);
  if (!map_add_found(c, F, sh, &s_size)) {
    if (threadIdx.x == 0) s_flag |= 2;
  }
  for (int i = threadIdx.x; i < F; i += blockDim.x) {
    const int h = c.ef[i] % W;  // one found per bucket
    const double2 old = c.MS[2 * h];
    c.MS[2 * h] = make_double2(0, 0);
    c.MS[2 * h + 1] = make_double2(0, 0);
    if (c.dirty) {
      atomicOr(&c.dirty[h >> 5], 1u << (h & 31));
      if (ms_counted(old, 1.0 / W)) atomicAdd(c.cm, -1);
    }
  }
  __threadfence_block();
  __syncthreads();
  if (F == a.k) {  // cf3 859-860: the reference returns with out untouched
    if (threadIdx.x == 0) {
      a.count[s] = 0;
      a.status[s] = 1;
    }
    return;
  }

  // ---- Gauss filter 1 (cf3 130-212): buckets from pre task 2
  {
    const double2 *gp = a.gpre + (int64_t)s * (2 * a.B1 + 2 * a.B2);
    for (int r = threadIdx.x; r < a.B1; r += blockDim.x) {
      c.GS0[r] = gp[r];
      c.GS1[r] = gp[a.B1 + r];
    }
    __syncthreads();
  }
  subtract(c, c.keys, c.vals, s_size, 1);  // the map so far, insertion order
  F = est_gauss(c, c.GS0, c.GS1, a.B1, 1, 0, a.win1, a.win1_half, sh);
  if (!map_add_found(c, F, sh, &s_size) && threadIdx.x == 0) s_flag |= 2;
  __syncthreads();
  subtract(c, c.ef, c.ev, F, 1 | 4);

  // ---- Gauss filter 2 (cf3 215-297): buckets from pre task 3 (run here:
  // in its own bucket arrays' LDS, phases and samples in the staging LDS)
  waits_ok &= need_tasks(a, s, claims, 3, 3, c.GP0, 2 * (size_t)a.B2 * sizeof(double2), c.stage,
                         c.room, &s_claim);
  {
    const double2 *gp = a.gpre + (int64_t)s * (2 * a.B1 + 2 * a.B2) + 2 * a.B1;
    for (int r = threadIdx.x; r < a.B2; r += blockDim.x) {
      c.GP0[r] = gp[r];
      c.GP1[r] = gp[a.B2 + r];
    }
    __syncthreads();
  }
  subtract(c, c.keys, c.vals, s_size, 2);
  F = est_gauss(c, c.GP0, c.GP1, a.B2, c.aa, c.b, a.win2, a.win2_half, sh);
  if (!map_add_found(c, F, sh, &s_size) && threadIdx.x == 0) s_flag |= 2;
  __syncthreads();
  subtract(c, c.ef, c.ev, F, 7);

  // ---- round robin until the three bucket counts repeat (cf3 997-1076)
  int prev[3] = {0, 0, 0};
  int pass = 0;
  bool cm_valid = false;  // s_cm holds the Mansour count (uniform)
  for (; pass < a.max_passes; pass++) {
    const int sel = pass % 3;
    if (sel == 0) F = c.dirty ? est_mansour_dirty(c, sh) : est_mansour(c, s_rank, sh);
    else if (sel == 1) F = est_gauss(c, c.GS0, c.GS1, a.B1, 1, 0, a.win1, a.win1_half, sh);
    else F = est_gauss(c, c.GP0, c.GP1, a.B2, c.aa, c.b, a.win2, a.win2_half, sh);
    if (!map_add_found(c, F, sh, &s_size)) {
      if (threadIdx.x == 0) s_flag |= 2;
      __syncthreads();
      break;
    }
    subtract(c, c.ef, c.ev, F, 7);
    if (sel == 2) {
      if (threadIdx.x < 3) s_cnt[threadIdx.x] = 0;
      __syncthreads();
      int c1 = 0, c2 = 0, cm = 0;
      for (int j = threadIdx.x; j < a.B1; j += blockDim.x) {
        // the reference counts elements 0..B1-1 of its interleaved array
        const double2 z = (j & 1) ? c.GS1[j >> 1] : c.GS0[j >> 1];
        c1 += z.x * z.x + z.y * z.y > 1e-6;
      }
      for (int j = threadIdx.x; j < a.B2; j += blockDim.x) {
        const double2 z = (j & 1) ? c.GP1[j >> 1] : c.GP0[j >> 1];
        c2 += z.x * z.x + z.y * z.y > 1e-6;
      }
      // the Mansour count is kept up to date by every bucket write once it
      // has been counted in full (the first test)
      if (!cm_valid) {
        const double invW = 1.0 / W;
        for (int j0 = threadIdx.x; j0 < W; j0 += kSweepAhead * kV3Block) {
          double2 z[kSweepAhead];  // loaded together, clamped past the end
#pragma unroll
          for (int u = 0; u < kSweepAhead; u++) z[u] = c.MS[2 * min(j0 + u * kV3Block, W - 1)];
#pragma unroll
          for (int u = 0; u < kSweepAhead; u++) cm += j0 + u * kV3Block < W && ms_counted(z[u], invW);
        }
      }
      c1 = wave_sum(c1);
      c2 = wave_sum(c2);
      if (!cm_valid) cm = wave_sum(cm);
      if ((threadIdx.x & 63) == 0) {
        if (!cm_valid) atomicAdd(&s_cnt[0], cm);
        atomicAdd(&s_cnt[1], c1);
        atomicAdd(&s_cnt[2], c2);
      }
      __syncthreads();
      const int now[3] = {cm_valid ? s_cm : s_cnt[0], s_cnt[1], s_cnt[2]};
      __syncthreads();
      if (!cm_valid && c.dirty) {
        if (threadIdx.x == 0) s_cm = now[0];
        cm_valid = true;
      }
      if (now[0] == prev[0] && now[1] == prev[1] && now[2] == prev[2]) break;
      prev[0] = now[0];
      prev[1] = now[1];
      prev[2] = now[2];
    }
  }
  if (pass >= a.max_passes && threadIdx.x == 0) s_flag |= 4;
  if (!waits_ok && threadIdx.x == 0) s_flag |= 16;
  __syncthreads();
  if (a.refine) {  // the refine kernels refine the map and write the result
    if (threadIdx.x == 0) {
      a.count[s] = s_size;
      a.status[s] = s_flag;
    }
    return;
  }
  write_result(a, s, c.keys, c.vals, s_size, s_flag, lds);
}

// Ascending bitonic sort of P (a power of two) distinct u64 keys in LDS by
// the whole block.  Each wave takes a contiguous share of the P/2
// compare-exchange pairs of a stage, q = wave ppw + r; a pair (i, i + j) of
// a stage with j <= ppw then lies in the wave's own span of 2 ppw keys, so
// consecutive such stages need only the wave's own ordering (its LDS
// operations complete in order) and block barriers surround only the
// stages with a wider j.  Contains barriers.
__device__ void block_bitonic_u64(uint64_t *v, int P)
{
  constexpr int NW = kV3Block / 64;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int half = P >> 1;
  const int ppw = half >= NW ? half / NW : half;  // pairs per wave in use
  const int nwu = ppw ? half / ppw : 0;           // waves in use
  bool prev_local = false;
  for (int k2 = 2; k2 <= P; k2 <<= 1) {
    for (int j = k2 >> 1; j > 0; j >>= 1) {
      const bool local = j <= ppw;
      if (local && prev_local) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      } else {
        __syncthreads();
      }
      prev_local = local;
      if (wid >= nwu) continue;
      const int lj = ilog2v(j);
      for (int r = lane; r < ppw; r += 64) {
        const int q = wid * ppw + r;
        const int i = ((q >> lj) << (lj + 1)) | (q & (j - 1));
        const int p = i | j;
        const uint64_t x = v[i], y = v[p];
        if ((x > y) == ((i & k2) == 0)) {
          v[i] = y;
          v[p] = x;
        }
      }
    }
  }
  __syncthreads();
}

// The result of signal s: the map (M entries in insertion order) sorted by
// frequency into the sparse outputs and scattered into the dense one; count
// and status flags.  All threads; contains barriers.
__device__ void write_result(const V3Launch &a, int s, const int *keys, const double2 *vals,
                             int M, int flags, double2 *lds)
{
  int P = 1;
  while (P < M) P <<= 1;
  uint64_t *srt = reinterpret_cast<uint64_t *>(lds);
  const bool sortable = (size_t)P * 8 <= a.lds_bytes;
  if (sortable) {
    for (int i = threadIdx.x; i < P; i += blockDim.x)
      srt[i] = i < M ? (((uint64_t)(unsigned)keys[i] << 32) | (unsigned)i) : ~0ull;
    block_bitonic_u64(srt, P);
  }
  for (int t = threadIdx.x; t < M; t += blockDim.x) {
    const int e = sortable ? (int)(srt[t] & 0xffffffffu) : t;
    const int key = keys[e];
    const double2 v = vals[e];
    if (a.o_freq && t < a.o_cap) {
      int32_t *of = a.o_freq[s];
      double2 *ov = a.o_val[s];
      if (of) of[t] = key;
      if (ov) ov[t] = v;
    }
    if (a.o_dense && a.o_dense[s]) a.o_dense[s][key] = v;
  }
  if (threadIdx.x == 0) {
    a.count[s] = M;
    a.status[s] = flags;
  }
}

// permute_fourier's phases (cf3 66-87): phase[i] = shift step^i for
// i <= w by repeated complex multiplication, the reference's rounding.
__device__ __forceinline__ void phase_chain(double2 *phase, double2 sh, const double2 st, int w)
{
  // unrolled: the loop control and stores stay off the product chain
#pragma unroll 8
  for (int i = 0; i <= w; i++) {
    phase[i] = sh;
    sh = make_double2(sh.x * st.x - sh.y * st.y, sh.x * st.y + sh.y * st.x);
  }
}

// lds[bitrev(i)] = x[(off + i sigma) & mask] for i < W (the input of an
// in-place W-point FFT), kPreAhead loads per thread issued together.
__device__ __forceinline__ void gather_strided(double2 *lds, const double2 *x, unsigned off,
                                               int sigma, int W, int logW, unsigned mask)
{
  const int T = blockDim.x;
  for (int i0 = threadIdx.x; i0 < W; i0 += kPreAhead * T) {
    double vr[kPreAhead], vi[kPreAhead];  // (scalars: a double2 array went to scratch)
#pragma unroll
    for (int u = 0; u < kPreAhead; u++) {
      const double2 v = x[(off + (unsigned)min(i0 + u * T, W - 1) * (unsigned)sigma) & mask];
      vr[u] = v.x;
      vi[u] = v.y;
    }
#pragma unroll
    for (int u = 0; u < kPreAhead; u++)  // past the end: sample W - 1 again, same value
      lds[bitrev((unsigned)min(i0 + u * T, W - 1), logW)] = make_double2(vr[u], vi[u]);
  }
}

// The start of every signal's transform that does not depend on its map of
// located frequencies, batch-wide, four workgroups per signal (the batch's
// other CUs would idle while one workgroup per signal did it in turn):
// block ntask j + t of (group, signal) blocks[j] computes
//   t = 0, 1: Mansour loop t (cf3 96-127): the sub-samples x[g0 + t + i n/W]
//             and their W-point FFT into the signal's interleaved buckets
//             (with the pre-pass above W = 8192, block 0 copies them);
//   t = 2:    Gauss filter 1 (cf3 130-212), both offsets, folded and FFT'd;
//   t = 3:    Gauss filter 2 (cf3 215-297): permute_fourier's phase chain
//             (cf3 66-87), the permuted fold and the FFTs;
//   t = 4 + l (ntask = 4 + kRefOff when a group refines): the refine form's
//             buckets at offset g0 + l ref_t.
// Same operations in the same order whichever workgroup runs them, so the
// buckets are the same bits.  The Gauss buckets go to gpre (GS0 | GS1 |
// GP0 | GP1 per signal).  lds: the LDS the task may use (bytes of it;
// task 3 puts its two bucket arrays there and its phases and samples in
// scr, scr_bytes of LDS).
__device__ __forceinline__ void pre_task(const V3Launch &a, const int s, const int task,
                                         double2 *lds, const size_t bytes, double2 *scr,
                                         const size_t scr_bytes)
{
  const int n = a.n, W = a.W;
  const unsigned mask = (unsigned)n - 1;
  const int *prm = a.params + (int64_t)s * 6;
  const int ai = prm[1], b = prm[2], g0 = prm[4], gG = prm[5];
  const double2 *x = a.x[s];
  double2 *gp = a.gpre + (int64_t)s * (2 * a.B1 + 2 * a.B2);
  if (task >= 4) {
    // the refine form's buckets at g0 + l t, l = task - 4 (from the
    // Mansour pre-pass instead above the LDS size)
    if (!a.refine || a.ms_pre) return;
    const int l = task - 4, sigma = n / W, logW = ilog2v(W);
    const unsigned off = (unsigned)g0 + (unsigned)l * (unsigned)a.ref_t;
    gather_strided(lds, x, off, sigma, W, logW, mask);
    __syncthreads();
    lds_fft(lds, logW, a.tw64, a.tw_nmax / W);
    double2 *RM = a.ref_ms + ((int64_t)s * kRefOff + l) * W;
    for (int i = threadIdx.x; i < W; i += blockDim.x) RM[i] = lds[i];
    return;
  }
  if (task < 2) {
    double2 *MS = a.ms + (int64_t)s * 2 * W;
    if (a.ms_pre) {
      if (task == 0) {
        const double2 *pre = a.ms_pre + (int64_t)s * a.ms_loops * W;
        for (int i = threadIdx.x; i < 2 * W; i += blockDim.x) MS[2 * (i % W) + i / W] = pre[i];
      }
      return;
    }
    const int l = task, sigma = n / W, logW = ilog2v(W);
    gather_strided(lds, x, (unsigned)g0 + (unsigned)l, sigma, W, logW, mask);
    __syncthreads();
    lds_fft(lds, logW, a.tw64, a.tw_nmax / W);
    for (int i = threadIdx.x; i < W; i += blockDim.x) MS[2 * i + l] = lds[i];
    return;
  }
  if (task == 2) {
    const int B = a.B1, chunks = a.w1 / B;
    double2 *GS0 = lds, *GS1 = lds + B;
    for (int r = threadIdx.x; r < B; r += blockDim.x) {
      double2 s0 = make_double2(0, 0), s1 = make_double2(0, 0);
      // kPreAhead chunks' loads issued together (clamped to the last
      // chunk), then summed in chunk order
      for (int c0 = 0; c0 < chunks; c0 += kPreAhead) {
        double2 tp[kPreAhead], x0[kPreAhead], x1[kPreAhead];
#pragma unroll
        for (int u = 0; u < kPreAhead; u++) {
          const int t = min(c0 + u, chunks - 1) * B + r;
          tp[u] = a.taps1[t];
          x0[u] = x[((unsigned)gG + (unsigned)t) & mask];
          x1[u] = x[((unsigned)gG + 1u + (unsigned)t) & mask];
        }
#pragma unroll
        for (int u = 0; u < kPreAhead; u++) {
          if (c0 + u >= chunks) break;
          s0.x = s0.x + (tp[u].x * x0[u].x - tp[u].y * x0[u].y);
          s0.y = s0.y + (tp[u].y * x0[u].x + tp[u].x * x0[u].y);
          s1.x = s1.x + (tp[u].x * x1[u].x - tp[u].y * x1[u].y);
          s1.y = s1.y + (tp[u].y * x1[u].x + tp[u].x * x1[u].y);
        }
      }
      GS0[r] = s0;
      GS1[r] = s1;
    }
    __syncthreads();
    fft_pair(GS0, GS1, B, a);
    __syncthreads();
    for (int r = threadIdx.x; r < B; r += blockDim.x) {
      gp[r] = GS0[r];
      gp[B + r] = GS1[r];
    }
    return;
  }
  // task 3: permute_fourier forms the phase of sample i as shift * step^i
  // by repeated complex multiplication; one lane replays that chain so the
  // samples carry the reference's rounding (into LDS past the bucket arrays
  // when it fits: a lone lane's global stores are several times slower)
  // (the bucket arrays at lds, the phases and samples in scr)
  const int B = a.B2, chunks = a.w2 / B;
  double2 *GP0 = lds, *GP1 = lds + B;
  const bool ph_lds = ((size_t)a.w2 + 1) * sizeof(double2) <= scr_bytes;
  double2 *const ph_s = scr;                             // LDS (typed stores)
  double2 *const ph_g = a.phase + (int64_t)s * a.Pcap;  // HBM otherwise
  // while lane 0 runs the chain, waves 1.. gather the permuted samples
  // x[(gG + i ai) mod n], i <= w2 (cf3 78-85; gG, ai >= 0 and n a power of
  // two, so the mod is a mask) into LDS past the phases, where they fit
  const bool staged = 2 * ((size_t)a.w2 + 1) * sizeof(double2) <= scr_bytes;
  double2 *xs = scr + a.w2 + 1;  // (an LDS address either way)
  if (staged && threadIdx.x >= 64) {
    const int T1 = blockDim.x - 64;
    for (int i0 = threadIdx.x - 64; i0 <= a.w2; i0 += kPreAhead * T1) {
      double vr[kPreAhead], vi[kPreAhead];  // (scalars: a double2 array went to scratch)
#pragma unroll
      for (int u = 0; u < kPreAhead; u++) {
        const unsigned long long i = (unsigned long long)min(i0 + u * T1, a.w2);
        const double2 v = x[((unsigned long long)gG + i * (unsigned long long)ai) & (unsigned long long)mask];
        vr[u] = v.x;
        vi[u] = v.y;
      }
#pragma unroll
      for (int u = 0; u < kPreAhead; u++)  // past the end: sample w2 again, same value
        xs[min(i0 + u * T1, a.w2)] = make_double2(vr[u], vi[u]);
    }
  }
  if (threadIdx.x == 0) {
    double s0, c0, s1, c1;
    sincos(2 * M_PI * gG * b / n, &s0, &c0);
    sincos(2 * M_PI * ai * b / n, &s1, &c1);
    const double2 sh = make_double2(c0, s0);
    const double2 st = make_double2(c1, s1);
    // separate calls so the LDS stores compile as LDS stores (a pointer
    // that may be either takes the slower generic path)
    if (ph_lds) phase_chain(ph_s, sh, st, a.w2);
    else phase_chain(ph_g, sh, st, a.w2);
  }
  __threadfence_block();
  __syncthreads();
  for (int r = threadIdx.x; r < B; r += blockDim.x) {
    double2 s0 = make_double2(0, 0), s1 = make_double2(0, 0);
    for (int ch = 0; ch < chunks; ch++) {
      const int t = ch * B + r;
      const double2 tp = a.taps2[t];
      double2 p[2];
#pragma unroll
      for (int q = 0; q < 2; q++) {
        // sample (gG + (t+q) ai) mod n (cf3 78-85)
        const double2 xv =
            staged ? xs[t + q] : x[(unsigned)(((long long)gG + (long long)(t + q) * ai) % n)];
        const double2 sv = ph_lds ? ph_s[t + q] : ph_g[t + q];
        p[q] = make_double2(xv.x * sv.x - xv.y * sv.y, xv.x * sv.y + xv.y * sv.x);
      }
      s0.x = s0.x + (tp.x * p[0].x - tp.y * p[0].y);
      s0.y = s0.y + (tp.y * p[0].x + tp.x * p[0].y);
      s1.x = s1.x + (tp.x * p[1].x - tp.y * p[1].y);
      s1.y = s1.y + (tp.y * p[1].x + tp.x * p[1].y);
    }
    GP0[r] = s0;
    GP1[r] = s1;
  }
  __syncthreads();
  fft_pair(GP0, GP1, B, a);
  __syncthreads();
  for (int r = threadIdx.x; r < B; r += blockDim.x) {
    gp[2 * a.B1 + r] = GP0[r];
    gp[2 * a.B1 + B + r] = GP1[r];
  }
}

// The whole batch in one launch, signals of several plans (different n, k)
// alike; blocks[j] = (plan group, signal).  Workgroups, in dispatch order:
//   4 j + t (t < 4)    pre task t of block j, if no one has claimed it;
//   4 nb + j           the signal's transform (v3_signal), which takes each
//                      pre task's result once done, or claims and runs a
//                      task no workgroup has started;
//   5 nb + nref j + l  refine task 4 + l of block j (nref = kRefOff when a
//                      group refines, else 0), read only by later kernels.
// The transforms start while the Gauss-2 phase chains still run.
__global__ void __launch_bounds__(kV3Block) k_v3_run(const V3Launch *__restrict__ groups,
                                                     const int2 *__restrict__ blocks, int nb,
                                                     int nref, int *__restrict__ claims)
{
  // restrict: the descriptors are read-only here, so their fields need not
  // be reloaded after every store inside the signal's loops
  extern __shared__ double2 lds[];
  __shared__ int s_go;
  int w = (int)blockIdx.x;
  if (w < 4 * nb) {
    const int j = w >> 2, task = w & 3;
    const int2 bb = blocks[j];
    const V3Launch &a = groups[bb.x];
    int *cl = claims + w;
    if (a.inline_pre) return;  // diagnostic: the transform runs the task
    if (threadIdx.x == 0) s_go = atomicCAS(cl, kClaimFree, kClaimBusy) == kClaimFree;
    __syncthreads();
    if (!s_go) return;
    const size_t gp = 2 * (size_t)a.B2 * sizeof(double2);
    pre_task(a, bb.y, task, lds, a.lds_bytes, lds + 2 * a.B2,
             a.lds_bytes > gp ? a.lds_bytes - gp : 0);
    publish_task(cl);
    return;
  }
  w -= 4 * nb;
  if (w < nb) {
    const int2 bb = blocks[w];
    v3_signal(groups[bb.x], bb.y, claims + 4 * w);
    return;
  }
  w -= nb;
  const int2 bb = blocks[w / nref];
  pre_task(groups[bb.x], bb.y, 4 + w % nref, lds, groups[bb.x].lds_bytes, nullptr, 0);
}

// The refine form after k_v3_run, in three kernels over the same
// (group, signal) blocks: k_v3_refine lists the located entries of every
// residue class (linked lists through head, W entries in cf, and next, M
// entries in remi) and the non-empty classes; k_v3_refine_solve solves
// those, kRefOff lanes each (nsub workgroups per signal, the later ones
// mostly idle); k_v3_refine_write sorts and writes the
// result.  Groups without refine, and signals with the reference's early
// return, wrote their result in k_v3_run and are skipped.
__global__ void __launch_bounds__(kV3Block) k_v3_refine(const V3Launch *__restrict__ groups,
                                                        const int2 *__restrict__ blocks)
{
  const int2 bb = blocks[blockIdx.x];
  const V3Launch &a = groups[bb.x];
  const int s = bb.y;
  if (!a.refine || (a.status[s] & 1)) return;
  const int M = (int)a.count[s], W = a.W;
  const int *keys = a.map_keys + (int64_t)s * a.Mcap;
  int *head = a.cf + (int64_t)s * a.Fcap, *next = a.remi + (int64_t)s * a.Rcap * 3;
  // the non-empty classes, for the solve (in cv, free after the loop):
  // nz[0] their count, nz[1..] the classes in no particular order
  int *nz = reinterpret_cast<int *>(a.cv + (int64_t)s * a.Fcap);
  __shared__ int s_nz;
  for (int j = threadIdx.x; j < W; j += blockDim.x) head[j] = -1;
  if (threadIdx.x == 0) s_nz = 0;
  __threadfence_block();
  __syncthreads();
  for (int e = threadIdx.x; e < M; e += blockDim.x) {
    const int j = keys[e] & (W - 1);
    const int old = atomicExch(&head[j], e);
    next[e] = old;
    if (old < 0) nz[1 + atomicAdd(&s_nz, 1)] = j;  // first entry of class j
  }
  __syncthreads();
  if (threadIdx.x == 0) nz[0] = s_nz;
}

__global__ void __launch_bounds__(kV3Block) k_v3_refine_solve(const V3Launch *__restrict__ groups,
                                                              const int2 *__restrict__ blocks,
                                                              int nsub)
{
  const int2 bb = blocks[blockIdx.x / (unsigned)nsub];
  const int gid = (int)(blockIdx.x % (unsigned)nsub) * blockDim.x + threadIdx.x;
  const int jj = gid / kRefOff, lane = gid % kRefOff;  // a group of kRefOff lanes a class
  const V3Launch &a = groups[bb.x];
  const int s = bb.y;
  if (!a.refine || (a.status[s] & 1)) return;  // no barriers here
  // one group per non-empty class (listed by k_v3_refine); a group returns
  // or continues as a whole, so its shuffles see all its lanes
  const int *nz = reinterpret_cast<const int *>(a.cv + (int64_t)s * a.Fcap);
  if (jj >= nz[0]) return;
  const int j = nz[1 + jj];
  const int g0 = a.params[(int64_t)s * 6 + 4];
  if (refine_class(a, s, g0, j, a.cf + (int64_t)s * a.Fcap, a.remi + (int64_t)s * a.Rcap * 3,
                   a.map_keys + (int64_t)s * a.Mcap, a.map_vals + (int64_t)s * a.Mcap, lane) &&
      lane == 0)
    atomicOr(&a.status[s], 8);
}

__global__ void __launch_bounds__(kV3Block) k_v3_refine_write(const V3Launch *__restrict__ groups,
                                                              const int2 *__restrict__ blocks)
{
  extern __shared__ double2 lds[];
  const int2 bb = blocks[blockIdx.x];
  const V3Launch &a = groups[bb.x];
  const int s = bb.y;
  if (!a.refine) return;
  const int st = a.status[s];
  if (st & 1) return;  // the reference's early return: no result
  write_result(a, s, a.map_keys + (int64_t)s * a.Mcap, a.map_vals + (int64_t)s * a.Mcap,
               (int)a.count[s], st, lds);
}

// Sub-samples of the Mansour loops for the pre-pass: ms_pre[(s*L+l)*W + i]
// = x[(init_offset + o_l + i n/W) mod n] (cf3 96-127), L = ms_loops, o_0 = 0
// and o_1 = 1 the reference's two, o_l = (l - 1) ref_t after them (the
// refine form's strided offsets).
__global__ void k_v3_mansour_gather(V3Launch a)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int L = a.ms_loops;
  if (t >= (int64_t)a.S * L * a.W) return;
  const int i = (int)(t % a.W);
  const int l = (int)((t / a.W) % L);
  const int s = (int)(t / (L * (int64_t)a.W));
  const unsigned g0 = (unsigned)a.params[(int64_t)s * 6 + 4];
  const unsigned sigma = (unsigned)(a.n / a.W);
  const unsigned off = g0 + (l < 2 ? (unsigned)l : (unsigned)(l - 1) * (unsigned)a.ref_t);
  a.ms_pre[t] = a.x[s][(off + (unsigned)i * sigma) & ((unsigned)a.n - 1)];
}

// Mansour pre-pass of one plan group whose W exceeds the LDS size.
static int mansour_prepass(const V3Launch &a, hipStream_t st)
{
  if (!a.ms_pre) return 0;
  const int64_t total = (int64_t)a.S * a.ms_loops * a.W;
  hipLaunchKernelGGL(k_v3_mansour_gather, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     st, a);
  if (hipGetLastError() != hipSuccess ||
      radix2_fft64(a.ms_pre, a.ms_tmp, a.W, (int64_t)a.S * a.ms_loops, 1, a.W, a.W, a.tw64,
                   a.tw_nmax, st) != 0) {
    set_error("v3: Mansour pre-pass launch failed");
    return -1;
  }
  return 0;
}

static void allow_v3_lds(const void *kernel, size_t bytes)
{
  if (bytes > 65536) raise_lds_limit(kernel);
}

int v3_multi(const V3Launch *h_groups, int G, const V3Launch *d_groups, const int2 *d_blocks,
             int nblocks, int *d_claims, hipStream_t st)
{
  size_t lds = 0;
  for (int g = 0; g < G; g++) {
    if (mansour_prepass(h_groups[g], st) != 0) return -1;
    lds = std::max(lds, h_groups[g].lds_bytes);
  }
  bool refine = false;
  int nsub = 1;
  for (int g = 0; g < G; g++)
    if (h_groups[g].refine) {
      refine = true;
      nsub = std::max(nsub, (h_groups[g].W * kRefOff + kV3Block - 1) / kV3Block);
    }
  const int nref = refine ? kRefOff : 0;
  allow_v3_lds(reinterpret_cast<const void *>(k_v3_run), lds);
  hipLaunchKernelGGL(k_v3_run, dim3((5 + nref) * nblocks), dim3(kV3Block), lds, st, d_groups,
                     d_blocks, nblocks, nref, d_claims);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("v3 run: ") + hipGetErrorString(e));
    return -1;
  }
  if (refine) {
    // groups without refine write their result in k_v3_run; their blocks
    // return at once here
    hipLaunchKernelGGL(k_v3_refine, dim3(nblocks), dim3(kV3Block), 0, st, d_groups, d_blocks);
    hipLaunchKernelGGL(k_v3_refine_solve, dim3(nblocks * nsub), dim3(kV3Block), 0, st, d_groups,
                       d_blocks, nsub);
    allow_v3_lds(reinterpret_cast<const void *>(k_v3_refine_write), lds);
    hipLaunchKernelGGL(k_v3_refine_write, dim3(nblocks), dim3(kV3Block), lds, st, d_groups,
                       d_blocks);
    e = hipGetLastError();
    if (e != hipSuccess) {
      set_error(std::string("v3 refine: ") + hipGetErrorString(e));
      return -1;
    }
  }
  return 0;
}

}  // namespace k
}  // namespace sfa
