// sFFT v1/v2 kernels for gfx950.
//
//   comb          v2 heavy-hitter prefilter: strided W-point sub-sample, FFT
//                 in LDS, top-2k residues (src/computefourier-1.0-2.0.cc:48-82)
//   comb_merge    sort+dedupe of several comb loops' residues (cf12 492-502)
//   fold_fft      permuted-filter bucketization folded into B buckets, then
//                 the B-point FFT in LDS (cf12 222-275)
//   select        top-2k buckets of each location loop (cf12 278-302)
//   vote          v1 exact-integer location votes (cf12 91-116) without an
//                 n-sized score array (count / scan / write over chunks)
//   estimate      per located frequency, the lower median over loops of
//                 bucket / filter (cf12 340-419)
//
// Signals of a batch are independent; blockIdx encodes (signal, loop) or
// (signal, chunk).  All index arithmetic is exact: for n = 2^p the low 32
// bits of i * a already give (i * a) mod n.
#include <algorithm>
#include <mutex>
#include <string>

#include "device_common.hpp"
#include "internal.hpp"

namespace sfa {
namespace k {

using namespace sfa::dev;

namespace {

__device__ __forceinline__ int ilog2(int v) { return 31 - __clz(v); }

__device__ __forceinline__ int64_t loop_offset(int l, int Lloc, int Bl, int Be)
{
  return (int64_t)min(l, Lloc) * Bl + (int64_t)max(0, l - Lloc) * Be;
}

inline int64_t loop_offset_host(int l, int Lloc, int Bl, int Be)
{
  return (int64_t)std::min(l, Lloc) * Bl + (int64_t)std::max(0, l - Lloc) * Be;
}

// ---------------------------------------------------------------- comb
// One block per (signal, comb loop).  LDS: W double2.  TB threads: kBlock
// for batches; more for a few signals, whose comb is on the latency path.
template <int TB>
__global__ void __launch_bounds__(TB) k_comb(V12Launch a)
{
  extern __shared__ double2 smem64[];
  __shared__ int sh[256 + 16];
  const int cl = a.comb_loops;
  const int s = blockIdx.x / cl, c = blockIdx.x % cl;
  const int stride = 2 * a.L + max(cl, 1);
  const int off = a.params[(int64_t)s * stride + 2 * a.L + c];
  const int W = a.W, logW = ilog2(W);
  const int sigma = a.n / W;
  const double2 *x = a.x[s];
  for (int i = threadIdx.x; i < W; i += blockDim.x)
    smem64[bitrev(i, logW)] = x[(int64_t)off + (int64_t)i * sigma];
  __syncthreads();
  lds_fft(smem64, logW, a.tw64, a.tw_nmax / W);
  // |.|^2 into the same LDS, as doubles (staged through registers: W is at
  // most kMaxLdsComb)
  constexpr int kPer = (kMaxLdsComb + TB - 1) / TB;
  double v[kPer];
#pragma unroll
  for (int q = 0; q < kPer; q++) {
    const int i = threadIdx.x + q * TB;
    if (i < W) {
      const double2 z = smem64[i];
      v[q] = z.x * z.x + z.y * z.y;
    }
  }
  __syncthreads();
  double *mag = reinterpret_cast<double *>(smem64);
#pragma unroll
  for (int q = 0; q < kPer; q++) {
    const int i = threadIdx.x + q * TB;
    if (i < W) mag[i] = v[q];
  }
  __syncthreads();
  int *out = a.approved + ((int64_t)s * cl + c) * a.num;
  block_top_num(mag, W, a.num, out, sh);
  if (cl == 1 && threadIdx.x == 0) {
    a.num_comb[s] = a.num;
    a.count[s] = (int64_t)a.num * (a.n / W);
  }
}

// W above the LDS limit: the strided sub-samples go to comb_buf, the W-point
// FFTs run as LDS radix-2 runs plus global passes, then k_comb_select_global.
__global__ void k_comb_gather(V12Launch a)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int cl = a.comb_loops, W = a.W;
  if (t >= (int64_t)a.S * cl * W) return;
  const int i = (int)(t % W);
  const int64_t sc = t / W;
  const int s = (int)(sc / cl), c = (int)(sc % cl);
  const int stride = 2 * a.L + max(cl, 1);
  const int off = a.params[(int64_t)s * stride + 2 * a.L + c];
  a.comb_buf[t] = a.x[s][(int64_t)off + (int64_t)i * (a.n / W)];
}

__global__ void __launch_bounds__(kBlock) k_comb_select_global(V12Launch a)
{
  __shared__ int sh[256 + 16];
  const int cl = a.comb_loops, W = a.W;
  const int s = blockIdx.x / cl, c = blockIdx.x % cl;
  const double2 *z = a.comb_buf + (int64_t)blockIdx.x * W;
  double *mag = a.mag_buf + (int64_t)blockIdx.x * W;
  for (int i = threadIdx.x; i < W; i += blockDim.x) mag[i] = z[i].x * z[i].x + z[i].y * z[i].y;
  __threadfence_block();
  __syncthreads();
  block_top_num(mag, W, a.num, a.approved + ((int64_t)s * cl + c) * a.num, sh);
  if (cl == 1 && threadIdx.x == 0) {
    a.num_comb[s] = a.num;
    a.count[s] = (int64_t)a.num * (a.n / W);
  }
}

// Several comb loops: union of residues, ascending (bitmap over W in LDS).
__global__ void __launch_bounds__(kBlock) k_comb_merge(V12Launch a)
{
  extern __shared__ unsigned bits[];
  __shared__ int sh[16];
  const int s = blockIdx.x;
  const int W = a.W, words = (W + 31) / 32;
  const int tot = a.comb_loops * a.num;
  int *ap = a.approved + (int64_t)s * tot;
  for (int i = threadIdx.x; i < words; i += blockDim.x) bits[i] = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < tot; i += blockDim.x) {
    const int r = ap[i];
    atomicOr(&bits[r >> 5], 1u << (r & 31));
  }
  __syncthreads();
  const int C = (words + blockDim.x - 1) / blockDim.x;
  const int lo = threadIdx.x * C, hi = min(words, lo + C);
  int cnt = 0;
  for (int wd = lo; wd < hi; wd++) cnt += __popc(bits[wd]);
  int total;
  int pos = block_exclusive_scan(cnt, sh, total);
  // write after everyone has read the unmerged lists
  __syncthreads();
  for (int wd = lo; wd < hi; wd++) {
    unsigned b = bits[wd];
    while (b) {
      const int bit = __ffs(b) - 1;
      ap[pos++] = wd * 32 + bit;
      b &= b - 1;
    }
  }
  if (threadIdx.x == 0) {
    a.num_comb[s] = total;
    a.count[s] = (int64_t)total * (a.n / W);
  }
}

// ---------------------------------------------------------------- fold + FFT
// One block per (signal, loop).  Each thread owns buckets m = tid + q*T and
// sums, in ascending tap order like the reference, x[(i*ai) mod n]*tap[i]
// over the taps i = m + r*B < w.  LDS: B double2.  FP64 only (the FP32
// form folds in k_fold_chunks).
constexpr int kQ = 8;  // buckets per thread in flight (independent gathers)

__global__ void __launch_bounds__(kBlock) k_fold_fft64(V12Launch a, int l0, int nl)
{
  extern __shared__ double2 smem64[];
  const int s = blockIdx.x / nl, l = l0 + blockIdx.x % nl;
  const bool isloc = l < a.Lloc;
  const int B = isloc ? a.B_loc : a.B_est;
  const int w = isloc ? a.w_loc : a.w_est;
  const double2 *__restrict__ taps = isloc ? a.taps_loc : a.taps_est;
  const int stride = 2 * a.L + max(a.comb_loops, 1);
  const unsigned ai = (unsigned)a.params[(int64_t)s * stride + a.L + l];
  const double2 *__restrict__ x = a.x[s];
  const unsigned mask = (unsigned)a.n - 1;
  const int logB = ilog2(B);
  const int T = blockDim.x;
  const int R = (w + B - 1) / B;
  for (int mb = threadIdx.x; mb < B; mb += T * kQ) {
    double2 acc[kQ];
#pragma unroll
    for (int q = 0; q < kQ; q++) acc[q] = make_double2(0.0, 0.0);
    for (int r = 0; r < R; r++) {
      double2 xv[kQ], tv[kQ];
#pragma unroll
      for (int q = 0; q < kQ; q++) {
        const int m = mb + q * T;
        const int i = m + r * B;
        if (m < B && i < w) {
          xv[q] = x[((unsigned)i * ai) & mask];
          tv[q] = taps[i];
        }
      }
#pragma unroll
      for (int q = 0; q < kQ; q++) {
        const int m = mb + q * T;
        const int i = m + r * B;
        if (m < B && i < w) {
          const double re = xv[q].x * tv[q].x - xv[q].y * tv[q].y;
          const double im = xv[q].x * tv[q].y + xv[q].y * tv[q].x;
          acc[q].x = acc[q].x + re;
          acc[q].y = acc[q].y + im;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < kQ; q++) {
      const int m = mb + q * T;
      if (m < B) smem64[bitrev(m, logB)] = acc[q];
    }
  }
  __syncthreads();
  lds_fft(smem64, logB, a.tw64, a.tw_nmax / B);
  double2 *out = a.xs64 + (int64_t)s * a.xs +
                 loop_offset(l, a.Lloc, a.B_loc, a.B_est);
  for (int j = threadIdx.x; j < B; j += T) out[j] = smem64[j];
}

// ---------------------------------------------------------------- streamed fold
// The FP64 bucket sums, scheduled for the Infinity Cache (MALL): work items
// are (signal, loop, 64 buckets), one per wavefront, numbered signal-major,
// and a fixed set of resident wavefronts walks them in order.  Only a few
// signals are then in flight at once, so the second and later touches of a
// signal's 128-byte lines (20 loops gather ~20% of its samples) find them in
// the MALL instead of HBM.  Each lane sums its bucket in ascending tap order
// (identical arithmetic to the fused kernel); the B-point FFT follows in
// the bucket FFT batches.  Buckets are written pre-FFT into xs64.
constexpr int kFoldWaves = 4;  // wavefronts per block

struct FoldItems {
  int l0, nl;          // loop range
  int c_loc, c_est;    // 64-bucket chunks per location / estimation loop
  int n_loc;           // location loops in the range
  int64_t per_signal;  // items per signal
  int64_t total;
};

__host__ __device__ inline FoldItems fold_items(const V12Launch &a, int l0, int nl)
{
  FoldItems f;
  f.l0 = l0;
  f.nl = nl;
  f.c_loc = (a.B_loc + 63) / 64;
  f.c_est = (a.B_est + 63) / 64;
  const int l1 = l0 + nl;
  f.n_loc = max(0, min(l1, a.Lloc) - l0);
  f.per_signal = (int64_t)f.n_loc * f.c_loc + (int64_t)(nl - f.n_loc) * f.c_est;
  f.total = f.per_signal * a.S;
  return f;
}

__global__ void __launch_bounds__(64 * kFoldWaves) k_fold_stream(V12Launch a, FoldItems f)
{
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t)blockIdx.x * kFoldWaves + (threadIdx.x >> 6);
  const int64_t nwaves = (int64_t)gridDim.x * kFoldWaves;
  const unsigned mask = (unsigned)a.n - 1;
  const int stride = 2 * a.L + max(a.comb_loops, 1);
  for (int64_t item = wave; item < f.total; item += nwaves) {
    const int s = (int)(item / f.per_signal);
    const int64_t r0 = item % f.per_signal;
    int l, ch;
    if (r0 < (int64_t)f.n_loc * f.c_loc) {
      l = f.l0 + (int)(r0 / f.c_loc);
      ch = (int)(r0 % f.c_loc);
    } else {
      const int64_t r1 = r0 - (int64_t)f.n_loc * f.c_loc;
      l = f.l0 + f.n_loc + (int)(r1 / f.c_est);
      ch = (int)(r1 % f.c_est);
    }
    const bool isloc = l < a.Lloc;
    const int B = isloc ? a.B_loc : a.B_est;
    const int w = isloc ? a.w_loc : a.w_est;
    const int m = ch * 64 + lane;
    if (m >= B) continue;
    const unsigned ai = (unsigned)a.params[(int64_t)s * stride + a.L + l];
    const double2 *__restrict__ x = a.x[s];
    const int64_t out = (int64_t)s * a.xs + loop_offset(l, a.Lloc, a.B_loc, a.B_est) + m;
    const int R = (w - m + B - 1) / B;  // taps m, m+B, ... below w
    const double2 *__restrict__ taps = isloc ? a.taps_loc : a.taps_est;
    double2 acc = make_double2(0.0, 0.0);
    int r = 0;
    for (; r + 8 <= R; r += 8) {
      double2 xv[8], tv[8];
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const unsigned i = (unsigned)(m + (r + q) * B);
        xv[q] = x[(i * ai) & mask];
        tv[q] = taps[i];
      }
#pragma unroll
      for (int q = 0; q < 8; q++) {
        acc.x = acc.x + (xv[q].x * tv[q].x - xv[q].y * tv[q].y);
        acc.y = acc.y + (xv[q].x * tv[q].y + xv[q].y * tv[q].x);
      }
    }
    for (; r < R; r++) {
      const unsigned i = (unsigned)(m + r * B);
      const double2 xv = x[(i * ai) & mask], tv = taps[i];
      acc.x = acc.x + (xv.x * tv.x - xv.y * tv.y);
      acc.y = acc.y + (xv.x * tv.y + xv.y * tv.x);
    }
    a.xs64[out] = acc;
  }
}

// ---------------------------------------------------------------- chunked fold
// The FP32 form's bucket sums, grouped so that an XCD's L2 serves repeated
// touches of a 128-byte line.  Loop l adds x[p] * tap[i] into bucket
// b = i mod B for the taps i = b + kB < w at p = i*ai mod n; p mod B =
// b*ai mod B, so bucket b only ever reads residue class r = b*ai mod B, and
// b = r*a mod B (a*ai = 1 mod n).  A workgroup takes the buckets of kResPer
// consecutive residues in every loop of the launch, one (residue, loop) pair
// per thread: it is the only reader of those residues' lines (viewing the
// signal as n/B rows of B samples, p = r + t*B).  Rows are walked in chunks
// of 2^seg_log rows; each thread adds, chunk after chunk, its taps whose
// sample lies in the chunk, in ascending tap order, with compensated (2Sum)
// FP32 sums, and adds the chunks' sums in chunk order.  The workgroup's
// threads work on the same chunk at about the same time, so the touches of
// one line (taps of different loops in one row) fall within one memory round
// trip and the L2 serves all but the first.  The pairs' taps are staged in
// LDS by coalesced loads of their by-bucket runs.  This fixes the FP32 form's
// summation order; with one chunk it is plain ascending tap order.
constexpr int kResPer = 16;                        // residues per workgroup
constexpr int kFoldResBlock = kResPer * kMaxLoops;  // 512: one pair per thread
constexpr int kChunkBatch = 16;                    // gathers in flight per thread and round

struct FoldChunks {
  int l0, nl;        // loops [l0, l0 + nl), all with B buckets
  int B, logB;
  int res, groups;   // residues per workgroup, B / res
  int seg_log, nseg; // rows per chunk 2^seg_log, chunks
  int tR;            // LDS stride of a pair's taps (max ceil(w/B) of the loops)
};

__global__ void __launch_bounds__(kFoldResBlock) k_fold_chunks(V12Launch a, FoldChunks f)
{
  extern __shared__ float2 s_tap[];  // pair-major, f.tR taps each
  __shared__ int s_b[kFoldResBlock];
  const int g = (int)(blockIdx.x % (unsigned)f.groups);
  const int s = (int)(blockIdx.x / (unsigned)f.groups);
  const int tid = threadIdx.x;
  const int npairs = f.res * f.nl;
  const bool has = tid < npairs;
  const int e = tid % f.res, jl = has ? tid / f.res : 0;
  const int l = f.l0 + jl;
  const bool isloc = l < a.Lloc;
  const int *prm = a.params + (int64_t)s * (2 * a.L + max(a.comb_loops, 1));
  const unsigned av = (unsigned)prm[l], ai = (unsigned)prm[a.L + l];
  const int B = f.B, logB = f.logB;
  const unsigned mask = (unsigned)a.n - 1;
  const int b = (int)(((unsigned)(g * f.res + e) * av) & (unsigned)(B - 1));
  const int w = isloc ? a.w_loc : a.w_est;
  const int R = (has && b < w) ? (w - b + B - 1) / B : 0;
  const unsigned p0 = ((unsigned)b * ai) & mask;
  const unsigned step = ((unsigned)B * ai) & mask;
  s_b[tid] = b;
  __syncthreads();
  for (int idx = tid; idx < npairs * f.tR; idx += blockDim.x) {
    const int pr = idx / f.tR, k = idx - pr * f.tR;
    const bool lp = f.l0 + pr / f.res < a.Lloc;
    const int tRp = lp ? a.tR_loc : a.tR_est;
    const float2 *T = lp ? a.tapsT_loc32 : a.tapsT_est32;
    s_tap[idx] = k < tRp ? T[(int64_t)s_b[pr] * tRp + k] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  if (!has) return;  // no barrier below
  const float2 *mytap = s_tap + tid * f.tR;
  const double2 *__restrict__ x = a.x[s];
  const int cshift = logB + f.seg_log;
  float2 tot = make_float2(0.f, 0.f);
  for (int c = 0; c < f.nseg; c++) {
    unsigned m = 0, pp = p0;
    for (int k = 0; k < R; k++) {
      if ((int)(pp >> cshift) == c) m |= 1u << k;
      pp = (pp + step) & mask;
    }
    float2 sum = make_float2(0.f, 0.f), cmp = make_float2(0.f, 0.f);
    while (m) {
      double2 xv[kChunkBatch];
      int kk[kChunkBatch];
#pragma unroll
      for (int q = 0; q < kChunkBatch; q++) {
        kk[q] = m ? __ffs(m) - 1 : -1;
        if (kk[q] >= 0) {
          xv[q] = x[(p0 + (unsigned)kk[q] * step) & mask];
          m &= m - 1;
        }
      }
#pragma unroll
      for (int q = 0; q < kChunkBatch; q++) {
        if (kk[q] >= 0) {
          const float2 pr = cmul(make_float2((float)xv[q].x, (float)xv[q].y), mytap[kk[q]]);
          float t = sum.x + pr.x, bp = t - sum.x;
          cmp.x += (sum.x - (t - bp)) + (pr.x - bp);
          sum.x = t;
          t = sum.y + pr.y;
          bp = t - sum.y;
          cmp.y += (sum.y - (t - bp)) + (pr.y - bp);
          sum.y = t;
        }
      }
    }
    const float2 part = make_float2(sum.x + cmp.x, sum.y + cmp.y);
    tot = c == 0 ? part : make_float2(tot.x + part.x, tot.y + part.y);
  }
  a.xs32[(int64_t)s * a.xs + loop_offset(l, a.Lloc, a.B_loc, a.B_est) + b] = tot;
}

// ---------------------------------------------------------------- v1 votes
// One block per (signal, location loop): top-num buckets by |.|^2.
// LDS: B doubles.  v1 location loops are always FP64 (xs64): their top-2k
// choice is made among leakage-level bucket powers FP32 cannot resolve.
template <bool GLOBAL>
__global__ void __launch_bounds__(kBlock) k_select(V12Launch a)
{
  extern __shared__ double lds_mag[];
  __shared__ int sh[256 + 16];
  // above the LDS size the magnitudes go to mag_buf
  double *mag = GLOBAL ? a.mag_buf + (int64_t)blockIdx.x * a.B_loc : lds_mag;
  const int s = blockIdx.x / a.Lloc, l = blockIdx.x % a.Lloc;
  const int B = a.B_loc;
  const int64_t base = (int64_t)s * a.xs + (int64_t)l * a.B_loc;
  const double2 *xs = a.xs64 + base;
  for (int i = threadIdx.x; i < B; i += blockDim.x) {
    const double2 z = xs[i];
    mag[i] = z.x * z.x + z.y * z.y;
  }
  if (GLOBAL) __threadfence_block();
  __syncthreads();
  block_top_num(mag, B, a.num, a.J + ((int64_t)s * a.Lloc + l) * a.num, sh);
}

// A frequency f is a hit when, in at least thr of the location loops, its
// permuted position p = f*ai mod n lies in the range
// [ceil((J-1/2) n/B), ceil((J+1/2) n/B)) of a selected bucket J
// (src/computefourier-1.0-2.0.cc:91-116).  Candidates are enumerated from the
// ranges of loops 0..Lloc-thr (candidate c: loop c / (num*m), then selected
// bucket, then offset in the range); a candidate is counted only from the
// first such loop that selects it, so every hit appears once.  The candidate
// list of a signal is cut into chunks of kVoteChunk, one block each:
// k_vote_count counts each chunk's hits, k_vote_scan turns the counts into
// offsets, k_vote_write writes the hits in candidate order.  LDS: Lloc * B
// bits of selected-bucket flags.
constexpr int kVoteChunk = 8 * kBlock;
static_assert(kVoteChunk == 2048, "vote_chunks() assumes chunks of 2048 candidates");

struct VoteCtx {
  const int *prm, *J;
  const unsigned *selbits;
  int64_t total, per_loop;
  unsigned mask;
  int words, half, logm;
};

__device__ VoteCtx vote_setup(const V12Launch &a, int s, unsigned *selbits)
{
  VoteCtx v;
  const int B = a.B_loc, Lloc = a.Lloc, num = a.num, n = a.n;
  v.words = (B + 31) / 32;
  v.prm = a.params + (int64_t)s * (2 * a.L + max(a.comb_loops, 1));
  v.mask = (unsigned)n - 1;
  const int m = n / B;
  v.half = m / 2;
  v.logm = ilog2(m);
  v.per_loop = (int64_t)num * m;
  v.total = (int64_t)max(Lloc - a.thr + 1, 0) * v.per_loop;
  v.J = a.J + (int64_t)s * Lloc * num;
  v.selbits = selbits;
  for (int i = threadIdx.x; i < Lloc * v.words; i += blockDim.x) selbits[i] = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < Lloc * num; i += blockDim.x) {
    const int l = i / num, b = v.J[i];
    atomicOr(&selbits[l * v.words + (b >> 5)], 1u << (b & 31));
  }
  __syncthreads();
  return v;
}

__device__ __forceinline__ int vote_keep(const V12Launch &a, const VoteCtx &v, int64_t c,
                                         unsigned &loc)
{
  if (c >= v.total) return 0;
  const int l0 = (int)(c / v.per_loop);
  const int64_t r = c % v.per_loop;
  const int jb = v.J[(int64_t)l0 * a.num + (int)(r >> v.logm)];
  const int q = (int)(r & ((1 << v.logm) - 1));
  const unsigned m = 1u << v.logm;
  const unsigned p = ((unsigned)jb * m - (unsigned)v.half + (unsigned)q + (unsigned)a.n) & v.mask;
  loc = (p * (unsigned)v.prm[l0]) & v.mask;  // original frequency
  int score = 0, earlier = 0;
  for (int l = 0; l < a.Lloc; l++) {
    const unsigned pl = (loc * (unsigned)v.prm[a.L + l]) & v.mask;
    const unsigned bk = ((pl + (unsigned)v.half) & v.mask) >> v.logm;
    const int sel = (v.selbits[l * v.words + (bk >> 5)] >> (bk & 31)) & 1;
    score += sel;
    earlier |= (l < l0) & sel;
  }
  return (score >= a.thr) && !earlier;
}

// grid: S x chunks
__global__ void __launch_bounds__(kBlock) k_vote_count(V12Launch a, int chunks)
{
  extern __shared__ unsigned selbits[];
  __shared__ int sh[16];
  const int s = blockIdx.x / chunks, ch = blockIdx.x % chunks;
  const VoteCtx v = vote_setup(a, s, selbits);
  int mine = 0;
  for (int i = 0; i < kVoteChunk; i += blockDim.x) {
    unsigned loc;
    mine += vote_keep(a, v, (int64_t)ch * kVoteChunk + i + threadIdx.x, loc);
  }
  int tot;
  (void)block_exclusive_scan(mine, sh, tot);
  if (threadIdx.x == 0) a.vote_cnt[blockIdx.x] = tot;
}

// grid: S; exclusive scan of a signal's chunk counts in place, and its
// located count (at most hits_cap)
__global__ void __launch_bounds__(kBlock) k_vote_scan(V12Launch a, int chunks)
{
  __shared__ int sh[16];
  int *cnt = a.vote_cnt + (int64_t)blockIdx.x * chunks;
  int64_t base = 0;
  for (int c0 = 0; c0 < chunks; c0 += blockDim.x) {
    const int c = c0 + threadIdx.x;
    const int v = c < chunks ? cnt[c] : 0;
    int tot;
    const int pos = block_exclusive_scan(v, sh, tot);
    if (c < chunks) cnt[c] = (int)min(base + pos, (int64_t)INT_MAX);
    base += tot;
  }
  if (threadIdx.x == 0) a.count[blockIdx.x] = min(base, a.hits_cap);
}

// grid: S x chunks
__global__ void __launch_bounds__(kBlock) k_vote_write(V12Launch a, int chunks)
{
  extern __shared__ unsigned selbits[];
  __shared__ int sh[16];
  const int s = blockIdx.x / chunks, ch = blockIdx.x % chunks;
  const VoteCtx v = vote_setup(a, s, selbits);
  int *hits = a.hits + (int64_t)s * a.hits_cap;
  int64_t base = a.vote_cnt[blockIdx.x];
  for (int i = 0; i < kVoteChunk; i += blockDim.x) {
    unsigned loc = 0;
    const int keep = vote_keep(a, v, (int64_t)ch * kVoteChunk + i + threadIdx.x, loc);
    int tot;
    const int pos = block_exclusive_scan(keep, sh, tot);
    if (keep && base + pos < a.hits_cap) hits[base + pos] = (int)loc;
    base += tot;
  }
}

// ---------------------------------------------------------------- v2 votes (diagnostic)
// The v2 location votes (src/computefourier-1.0-2.0.cc:125-184), which never
// change the output (every frequency they append is already in the comb
// prefill), counted for parity tests.  The reference walks, for location
// loop l and selected bucket J, the positions of the range
// [ceil((J-1/2) n/B), ceil((J+1/2) n/B)) in the classes of the comb residues,
// adds one to score[h] of the frequency h = position * a mod n and appends h
// when its score reaches thr.  Those frequencies are the located set
// {j W + c}; a frequency's score after loop l is the number of loops <= l
// whose selected buckets hold h * ai mod n.  One thread per located
// frequency (prefill order) writes its append key -- (loop where the score
// reaches thr, rank of that bucket in J, offset of the position in the
// range), the order of the reference's appends -- or -1.
__global__ void __launch_bounds__(kBlock) k_comb_vote_keys(V12Launch a, int64_t chunks,
                                                          int64_t *keys, int64_t stride)
{
  extern __shared__ unsigned selbits[];
  const int s = (int)(blockIdx.x / chunks);
  const int64_t t = (blockIdx.x % chunks) * kBlock + threadIdx.x;
  const int B = a.B_loc, Lloc = a.Lloc, num = a.num;
  const int words = (B + 31) / 32;
  const int *J = a.J + (int64_t)s * Lloc * num;
  for (int i = threadIdx.x; i < Lloc * words; i += blockDim.x) selbits[i] = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < Lloc * num; i += blockDim.x) {
    const int l = i / num, b = J[i];
    atomicOr(&selbits[l * words + (b >> 5)], 1u << (b & 31));
  }
  __syncthreads();
  const int nc = a.num_comb[s];
  const int64_t located = (int64_t)nc * (a.n / a.W);
  if (t >= a.est_cap) return;
  if (t >= located) {  // capacity past the located set (several comb loops)
    keys[(int64_t)s * stride + 2 * t] = -1;
    keys[(int64_t)s * stride + 2 * t + 1] = -1;
    return;
  }
  const int *prm = a.params + (int64_t)s * (2 * a.L + max(a.comb_loops, 1));
  const unsigned mask = (unsigned)a.n - 1;
  const int m = a.n / B, half = m / 2, logm = ilog2(m);
  const unsigned h = (unsigned)(t / nc) * (unsigned)a.W +
                     (unsigned)a.approved[(int64_t)s * a.comb_loops * num + t % nc];
  int64_t key = -1;
  int score = 0;
  for (int l = 0; l < Lloc && key < 0; l++) {
    const unsigned p = (h * (unsigned)prm[a.L + l]) & mask;
    const unsigned bk = ((p + (unsigned)half) & mask) >> logm;
    const unsigned *sb = selbits + l * words;
    if (!((sb[bk >> 5] >> (bk & 31)) & 1)) continue;
    if (++score < a.thr) continue;
    int rank = __popc(sb[bk >> 5] & ((1u << (bk & 31)) - 1));
    for (unsigned wd = 0; wd < (bk >> 5); wd++) rank += __popc(sb[wd]);
    const unsigned low = (bk * (unsigned)m - (unsigned)half + (unsigned)a.n) & mask;
    key = (int64_t)l << 48 | (int64_t)rank << 28 | (int64_t)((p - low) & mask);
  }
  keys[(int64_t)s * stride + 2 * t] = key;
  keys[(int64_t)s * stride + 2 * t + 1] = (int64_t)h;
}

// ---------------------------------------------------------------- estimate
// Per located frequency h and loop l: the bucket h hashes to, divided by the
// filter response at h's offset from the bucket centre (a half-bucket tie
// rounds down, cf12 371-378); then the lower median (rank (L-1)/2) of the
// real and imaginary parts separately (cf12 406-413).
//
// Blocks are mapped so that every block of one signal lands on the same XCD
// (workgroups are dispatched round-robin over the 8 XCDs): the signal's
// buckets then stay in that XCD's L2 while its located set is estimated.
constexpr int kXcds = 8;

__device__ __forceinline__ bool xcd_block(int64_t b, int S, int64_t per_signal,
                                          int &s, int64_t &c)
{
  if (S < kXcds) {  // too few signals to give each XCD its own: spread them
    s = (int)(b / per_signal);
    c = b % per_signal;
    return s < S;
  }
  const int x = (int)(b % kXcds);
  const int64_t q = b / kXcds;
  s = x + kXcds * (int)(q / per_signal);
  c = q % per_signal;
  return s < S;
}

__host__ __device__ inline int64_t xcd_grid(int S, int64_t per_signal)
{
  return S < kXcds ? (int64_t)S * per_signal
                   : (int64_t)((S + kXcds - 1) / kXcds) * kXcds * per_signal;
}

__host__ __device__ constexpr int pow2_at_least(int v)
{
  return v <= 1 ? 1 : 2 * pow2_at_least((v + 1) / 2);
}

// Ascending sort of v[0..LMAX) with Batcher's odd-even merge network on the
// next power of two; comparators that touch a slot >= LMAX compare with an
// implicit +inf and are dropped at compile time.
template <int LMAX>
__device__ __forceinline__ void sort_net(double (&v)[LMAX])
{
  constexpr int N = pow2_at_least(LMAX);
#pragma unroll
  for (int p = 1; p < N; p <<= 1) {
#pragma unroll
    for (int k = p; k >= 1; k >>= 1) {
#pragma unroll
      for (int j = k % p; j + k < N; j += 2 * k) {
#pragma unroll
        for (int i = 0; i < k; i++) {
          const int lo = i + j, hi = i + j + k;
          if (hi < LMAX && i < N - j - k && lo / (2 * p) == hi / (2 * p)) {
            const double x = fmin(v[lo], v[hi]), y = fmax(v[lo], v[hi]);
            v[lo] = x;
            v[hi] = y;
          }
        }
      }
    }
  }
}

// Ascending sort of v[O..O+10) with a 29-comparator 10-input network
// (every 0-1 input checked).
template <int LMAX, int O>
__device__ __forceinline__ void sort10(double (&v)[LMAX])
{
  constexpr int net[29][2] = {{4, 9}, {3, 8}, {2, 7}, {1, 6}, {0, 5}, {1, 4}, {6, 9}, {0, 3},
                              {5, 8}, {0, 2}, {3, 6}, {7, 9}, {0, 1}, {2, 4}, {5, 7}, {8, 9},
                              {1, 2}, {4, 6}, {7, 8}, {3, 5}, {2, 5}, {6, 8}, {1, 3}, {4, 7},
                              {2, 3}, {6, 7}, {3, 4}, {5, 6}, {4, 5}};
#pragma unroll
  for (int c = 0; c < 29; c++) {
    const int lo = O + net[c][0], hi = O + net[c][1];
    const double x = fmin(v[lo], v[hi]), y = fmax(v[lo], v[hi]);
    v[lo] = x;
    v[hi] = y;
  }
}

// Lower median (rank 9) of 20 values: with both halves sorted, the 10th
// smallest of their union is max_i min(A[i], B[9-i]).
template <int LMAX>
__device__ __forceinline__ double median20(double (&v)[LMAX])
{
  sort10<LMAX, 0>(v);
  sort10<LMAX, 10>(v);
  double m = fmin(v[0], v[19]);
#pragma unroll
  for (int i = 1; i < 10; i++) m = fmax(m, fmin(v[i], v[19 - i]));
  return m;
}

template <int LMAX>
__device__ __forceinline__ double2 lower_median(double (&vr)[LMAX], double (&vi)[LMAX],
                                                int L)
{
  if constexpr (LMAX == 20) {
    if (L == 20) return make_double2(median20<LMAX>(vr), median20<LMAX>(vi));
  }
  sort_net<LMAX>(vr);
  sort_net<LMAX>(vi);
  const int rank = (L - 1) / 2;
  double mr = vr[0], mi = vi[0];
#pragma unroll
  for (int l = 1; l < LMAX; l++) {
    if (l == rank) {
      mr = vr[l];
      mi = vi[l];
    }
  }
  return make_double2(mr, mi);
}

__device__ __forceinline__ double2 load_bucket(const V12Launch &a, bool fp32, bool isloc,
                                               int64_t idx)
{
  if (fp32 && !(a.loc64 && isloc)) {
    const float2 u = a.xs32[idx];
    return make_double2((double)u.x, (double)u.y);
  }
  return a.xs64[idx];
}

// bucket / G with the reference's operation order (division by |G|^2)
__device__ __forceinline__ void divide_by_filter(double2 U, double2 G, double den,
                                                 bool conj, double &vr, double &vi)
{
  vr = (U.x * G.x + U.y * G.y) / den;
  vi = conj ? (U.x * G.y - U.y * G.x) / den : (U.y * G.x - U.x * G.y) / den;
}

// FP32 form: bucket u times g = G / |G|^2, one complex multiply instead of
// two divisions, each part one product and one fused multiply-add.  sign
// (V12Launch::conj_sign) is 1.0 for the reference's conjugated estimate and
// -1.0 for the other, whose imaginary part is the exact negation; it is a
// launch argument rather than a select on ref_conj, so the negation stays
// one multiplication.  Every FP32 estimation path calls this, so they
// agree bit for bit.
__device__ __forceinline__ void mul_by_inverse(double ux, double uy, double2 g, double sign,
                                               double &vr, double &vi)
{
  vr = fma(ux, g.x, uy * g.y);
  vi = sign * fma(ux, g.y, -(uy * g.x));
}

// Bucket and filter value of frequency h in loop l, computed directly.
__device__ __forceinline__ void hash_direct(const V12Launch &a, int l, unsigned ai,
                                            unsigned h, int &hashed, double2 &G)
{
  const bool isloc = l < a.Lloc;
  const int B = isloc ? a.B_loc : a.B_est;
  const int m = a.n / B;
  const unsigned p = (ai * h) & ((unsigned)a.n - 1);
  hashed = (int)(p / (unsigned)m);
  int dist = (int)(p % (unsigned)m);
  if (dist > m / 2) {
    hashed = (hashed + 1) & (B - 1);
    dist -= m;
  }
  G = (isloc ? a.win_loc : a.win_est)[dist + m / 2];
}

__device__ __forceinline__ void write_result(const V12Launch &a, int s, int64_t t,
                                             unsigned h, double2 val)
{
  if (a.o_freq && t < a.o_cap) {
    int32_t *of = a.o_freq[s];
    double2 *ov = a.o_val[s];
    if (of) of[t] = (int32_t)h;
    if (ov) ov[t] = val;
  }
  if (a.o_dense) {
    double2 *od = a.o_dense[s];
    if (od) od[h] = val;
  }
}

// v2: the located set is h = j*W + c for the comb residues c (in the
// reference's prefill order t = j*num_comb + i).  When W is a multiple of
// m = n/B, every h with the same residue has the same offset from its
// bucket centre and its bucket advances by (ai*W/m) mod B per alias row j,
// so the filter value and base bucket are tabulated per (residue, loop).
// Block tile: kTI residues x kTJ alias rows, kTI*kTJ/256 hits per thread.
constexpr int kTI = 64, kTJ = 16;

template <int LMAX, bool FP32, bool ALLFAST>
__global__ void __launch_bounds__(kBlock) k_estimate_comb(V12Launch a, int tiles_i,
                                                          int tiles_j)
{
  __shared__ int t_hash[kTI * LMAX];
  __shared__ double2 t_g[kTI * LMAX];
  __shared__ double t_den[FP32 ? 1 : kTI * LMAX];
  __shared__ unsigned s_ai[LMAX];
  __shared__ unsigned s_step[LMAX];
  __shared__ unsigned s_mask[LMAX];
  __shared__ int s_off[LMAX];
  int s;
  int64_t c;
  if (!xcd_block(blockIdx.x, a.S, (int64_t)tiles_i * tiles_j, s, c)) return;
  const int ti = (int)(c % tiles_i), tj = (int)(c / tiles_i);
  const int nc = a.num_comb[s];
  const int rows = a.n / a.W;
  if (ti * kTI >= nc || tj * kTJ >= rows) return;
  const int L = a.L;
  const int stride = 2 * L + max(a.comb_loops, 1);
  const int *approved = a.approved + (int64_t)s * a.comb_loops * a.num;
  if (threadIdx.x < L) {
    const int l = threadIdx.x;
    const unsigned ai = (unsigned)a.params[(int64_t)s * stride + L + l];
    const int B = l < a.Lloc ? a.B_loc : a.B_est;
    const int m = a.n / B;
    s_ai[l] = ai;
    s_mask[l] = (unsigned)(B - 1);
    s_off[l] = (int)loop_offset(l, a.Lloc, a.B_loc, a.B_est);
    // 0 marks a loop whose W is not a multiple of m (direct hashing)
    s_step[l] = (a.W % m == 0) ? ((ai * (unsigned)(a.W / m)) & (unsigned)(B - 1)) | 0x80000000u
                               : 0u;
  }
  __syncthreads();
  if (ALLFAST) {
    // every loop tabulated: a lean loop body with per-loop scalars in LDS
    for (int e = threadIdx.x; e < kTI * L; e += blockDim.x) {
      const int l = e / kTI, il = e % kTI;  // consecutive threads: consecutive il
      const int i = ti * kTI + il;
      if (i < nc) {
        int hashed;
        double2 G;
        hash_direct(a, l, s_ai[l], (unsigned)approved[i], hashed, G);
        const double den = G.x * G.x + G.y * G.y;
        t_hash[l * kTI + il] = hashed;
        if (FP32) {
          t_g[l * kTI + il] = make_double2(G.x / den, G.y / den);
        } else {
          t_g[l * kTI + il] = G;
          t_den[l * kTI + il] = den;
        }
      }
    }
    __syncthreads();
    const double2 *xs64 = a.xs64 + (int64_t)s * a.xs;
    const float2 *xs32 = FP32 ? a.xs32 + (int64_t)s * a.xs : nullptr;
    const bool conj = a.ref_conj;
    const double sign = a.conj_sign;
    for (int e = threadIdx.x; e < kTI * kTJ; e += kBlock) {
      // re-read the per-loop scalars from LDS every hit instead of holding
      // 3 x LMAX of them in registers across the loop (occupancy)
      asm volatile("" ::: "memory");
      const int jl = e / kTI;
      const int il = (e + jl) % kTI;  // skew: see below
      const int i = ti * kTI + il;
      const int j = tj * kTJ + jl;
      // hits past the tile's edge run too (clamped indices, nothing
      // written) so that every thread reaches the barriers below
      const bool valid = i < nc && j < rows;
      const int ic = min(i, nc - 1), jc = min(j, rows - 1);
      const unsigned h = (unsigned)jc * (unsigned)a.W + (unsigned)approved[ic];
      double vr[LMAX], vi[LMAX];
#pragma unroll
      for (int l = 0; l < LMAX; l++) {
        if (l < L) {
          const int hashed =
              (int)(((unsigned)t_hash[l * kTI + il] + (unsigned)jc * (s_step[l] & 0x7fffffffu)) &
                    s_mask[l]);
          const double2 g = t_g[l * kTI + il];
          double2 U;
          if (FP32) {
            const float2 u = xs32[s_off[l] + hashed];
            mul_by_inverse((double)u.x, (double)u.y, g, sign, vr[l], vi[l]);
          } else {
            U = xs64[s_off[l] + hashed];
            const double den = t_den[l * kTI + il];
            vr[l] = (U.x * g.x + U.y * g.y) / den;
            const double q = (U.x * g.y - U.y * g.x) / den;
            vi[l] = conj ? q : -q;
          }
        } else {
          vr[l] = __longlong_as_double(0x7ff0000000000000ll);  // +inf pads
          vi[l] = vr[l];
        }
        // every 4 loops the block's wavefronts meet, so that they read the
        // same loops' bucket arrays together (L1 reuse; measured -7%)
        if (l % 4 == 3) __syncthreads();
      }
      const double2 med = lower_median<LMAX>(vr, vi, L);
      if (valid) write_result(a, s, (int64_t)j * nc + i, h, med);
    }
    return;
  }
  for (int e = threadIdx.x; e < kTI * L; e += blockDim.x) {
    const int l = e / kTI, il = e % kTI;  // consecutive threads: consecutive il
    const int i = ti * kTI + il;
    if (i < nc && (s_step[l] & 0x80000000u)) {
      int hashed;
      double2 G;
      hash_direct(a, l, s_ai[l], (unsigned)approved[i], hashed, G);
      const double den = G.x * G.x + G.y * G.y;
      t_hash[l * kTI + il] = hashed;
      if (FP32) {
        t_g[l * kTI + il] = make_double2(G.x / den, G.y / den);
      } else {
        t_g[l * kTI + il] = G;
        t_den[l * kTI + il] = den;
      }
    }
  }
  __syncthreads();
  const int64_t sbase = (int64_t)s * a.xs;
  // Hit e of the tile: row jl = e / kTI, residue il skewed by the row so that
  // a thread's table entries change from hit to hit (keeps them in LDS rather
  // than hoisted into registers) while a wavefront still writes consecutive
  // output slots.
  for (int e = threadIdx.x; e < kTI * kTJ; e += kBlock) {
    const int jl = e / kTI;
    const int il = (e + jl) % kTI;
    const int i = ti * kTI + il;
    const int j = tj * kTJ + jl;
    if (i >= nc || j >= rows) continue;
    const unsigned h = (unsigned)j * (unsigned)a.W + (unsigned)approved[i];
    double vr[LMAX], vi[LMAX];
#pragma unroll
    for (int l = 0; l < LMAX; l++) {
      if (l < L) {
        const bool isloc = l < a.Lloc;
        const int B = isloc ? a.B_loc : a.B_est;
        const int64_t off = sbase + loop_offset(l, a.Lloc, a.B_loc, a.B_est);
        const unsigned step = s_step[l];
        if (step & 0x80000000u) {
          const int hashed = (t_hash[l * kTI + il] +
                              (int)(((unsigned)j * (step & 0x7fffffffu)) & (unsigned)(B - 1))) &
                             (B - 1);
          const double2 U = load_bucket(a, FP32, isloc, off + hashed);
          const double2 g = t_g[l * kTI + il];
          if (FP32) {
            mul_by_inverse(U.x, U.y, g, a.conj_sign, vr[l], vi[l]);
          } else {
            divide_by_filter(U, g, t_den[l * kTI + il], a.ref_conj, vr[l], vi[l]);
          }
        } else {
          int hashed;
          double2 G;
          hash_direct(a, l, s_ai[l], h, hashed, G);
          const double2 U = load_bucket(a, FP32, isloc, off + hashed);
          divide_by_filter(U, G, G.x * G.x + G.y * G.y, a.ref_conj, vr[l], vi[l]);
        }
      } else {
        vr[l] = __longlong_as_double(0x7ff0000000000000ll);  // +inf pads
        vi[l] = vr[l];
      }
    }
    write_result(a, s, (int64_t)j * nc + i, h, lower_median<LMAX>(vr, vi, L));
  }
}

// ------------------------------------------------------ v2: alias-row layout
// With W a multiple of m = n/B, frequency h = j*W + c of comb residue c
// lies in bucket hashed0(c) + j*ai*d (mod B) of a loop, d = W/m, for its
// alias rows j < rows = n/W = B/d.  Stored at position r*rows + (q * u^-1
// mod rows) for bucket b = r + d*q (u = ai mod rows, odd), the buckets of
// rows j, j+1, ... of one residue are adjacent: a wavefront of 64 rows reads
// 512 contiguous bytes instead of 64 scattered lines.  The products and the
// median are the same as in k_estimate_comb, so the results are identical.

// inverse of odd a modulo 2^32 (Newton's iteration from a*a = 1 mod 8)
__device__ __forceinline__ unsigned inv_odd(unsigned a)
{
  unsigned x = a;
#pragma unroll
  for (int i = 0; i < 4; i++) x *= 2u - a * x;
  return x;
}

// Bucket FFTs of loops [l0, l0 + nl), written in the alias-row layout.
template <bool FP32>
__global__ void __launch_bounds__(kBlock) k_bucket_fft_rows(V12Launch a, int l0, int nl)
{
  using T2 = typename std::conditional<FP32, float2, double2>::type;
  extern __shared__ char rraw[];
  T2 *rbuf = reinterpret_cast<T2 *>(rraw);
  const int s = blockIdx.x / nl, l = l0 + blockIdx.x % nl;
  const int B = l < a.Lloc ? a.B_loc : a.B_est;
  const int logB = ilog2(B);
  const int rows = a.n / a.W, d = B / rows;
  const int stride = 2 * a.L + max(a.comb_loops, 1);
  const unsigned ai = (unsigned)a.params[(int64_t)s * stride + a.L + l];
  const unsigned uinv = inv_odd(ai) & (unsigned)(rows - 1);
  T2 *io;
  if constexpr (FP32) io = a.xs32; else io = a.xs64;
  io += (int64_t)s * a.xs + loop_offset(l, a.Lloc, a.B_loc, a.B_est);
  for (int m = threadIdx.x; m < B; m += blockDim.x) rbuf[bitrev(m, logB)] = io[m];
  __syncthreads();
  if constexpr (FP32) lds_fft(rbuf, logB, a.tw32, a.tw_nmax / B);
  else lds_fft(rbuf, logB, a.tw64, a.tw_nmax / B);
  for (int b = threadIdx.x; b < B; b += blockDim.x) {
    const int r = b % d, q = b / d;
    io[r * rows + (int)(((unsigned)q * uinv) & (unsigned)(rows - 1))] = rbuf[b];
  }
}

// Residues per tile of k_estimate_rows (RI / 4 per wavefront): 16 for
// batches, 4 when a batch gives too few tiles to fill the chip.
constexpr int kRI = 16, kRIsmall = 4;

// Median value of one residue (table column il) at alias row jm: the L
// bucket reads of the lane, each at byte offset t_base + (t_q + jm mod
// rows) * sizeof(bucket) of the signal's bucket array xs (the alias-row
// layout needs the bucket FFTs in LDS, so a signal's buckets span a few MB
// and 32-bit offsets hold them).
template <int LMAX, bool FULL, bool FP32, int RI>
__device__ __forceinline__ double2 rows_median(const char *xs, const unsigned *t_base,
                                               const int *t_q, const double2 *t_g,
                                               const double *t_den, int il, int jm, int rows,
                                               int L, bool conj, double sign)
{
  constexpr unsigned kElt = FP32 ? sizeof(float2) : sizeof(double2);
  double vr[LMAX], vi[LMAX];
#pragma unroll
  for (int l = 0; l < LMAX; l++) {
    if (FULL || l < L) {
      const unsigned off =
          t_base[l * RI + il] + ((unsigned)(t_q[l * RI + il] + jm) & (unsigned)(rows - 1)) * kElt;
      const double2 g = t_g[l * RI + il];
      if constexpr (FP32) {
        const float2 u = *reinterpret_cast<const float2 *>(xs + off);
        mul_by_inverse((double)u.x, (double)u.y, g, sign, vr[l], vi[l]);
      } else {
        // the reference's operation order (cf12 383-398), as k_estimate_comb
        divide_by_filter(*reinterpret_cast<const double2 *>(xs + off), g, t_den[l * RI + il],
                         conj, vr[l], vi[l]);
      }
    } else {
      vr[l] = __longlong_as_double(0x7ff0000000000000ll);  // +inf pads
      vi[l] = vr[l];
    }
  }
  return lower_median<LMAX>(vr, vi, FULL ? LMAX : L);
}

// FULL: L == LMAX, so the loop over loops has no branch and the compiler
// issues all bucket loads of a frequency before the first wait.
template <int LMAX, bool FULL, bool FP32, int RI>
__global__ void __launch_bounds__(kBlock) k_estimate_rows(V12Launch a, int tiles_i, int tiles_j)
{
  using T2 = typename std::conditional<FP32, float2, double2>::type;
  __shared__ unsigned t_base[RI * LMAX];  // byte offset of loop l's row r: (off_l + r*rows) * elt
  __shared__ int t_q[RI * LMAX];        // q0 * u^-1 mod rows
  __shared__ double2 t_g[RI * LMAX];    // FP32 form: G / |G|^2; FP64: G
  __shared__ double t_den[FP32 ? 1 : RI * LMAX];  // FP64: |G|^2
  __shared__ double2 stage[64 * (RI + 1)];  // [row][residue], padded
  int s;
  int64_t c;
  if (!xcd_block(blockIdx.x, a.S, (int64_t)tiles_i * tiles_j, s, c)) return;
  const int ti = (int)(c % tiles_i), tj = (int)(c / tiles_i);
  const int nc = a.num_comb[s];
  const int rows = a.n / a.W;
  if (ti * RI >= nc) return;
  const int L = a.L;
  const int stride = 2 * L + max(a.comb_loops, 1);
  const int *approved = a.approved + (int64_t)s * a.comb_loops * a.num;
  for (int e = threadIdx.x; e < RI * L; e += blockDim.x) {
    const int l = e / RI, il = e % RI;
    const int i = ti * RI + il;
    if (i < nc) {
      const unsigned ai = (unsigned)a.params[(int64_t)s * stride + L + l];
      int hashed;
      double2 G;
      hash_direct(a, l, ai, (unsigned)approved[i], hashed, G);
      const int B = l < a.Lloc ? a.B_loc : a.B_est;
      const int d = B / rows;
      const unsigned uinv = inv_odd(ai) & (unsigned)(rows - 1);
      const double den = G.x * G.x + G.y * G.y;
      t_base[l * RI + il] =
          (unsigned)(loop_offset(l, a.Lloc, a.B_loc, a.B_est) + (hashed % d) * rows) *
          (unsigned)sizeof(T2);
      t_q[l * RI + il] = (int)(((unsigned)(hashed / d) * uinv) & (unsigned)(rows - 1));
      if (FP32) {
        t_g[l * RI + il] = make_double2(G.x / den, G.y / den);
      } else {
        t_g[l * RI + il] = G;
        t_den[l * RI + il] = den;
      }
    }
  }
  __syncthreads();
  const T2 *xsig;
  if constexpr (FP32) xsig = a.xs32 + (int64_t)s * a.xs;
  else xsig = a.xs64 + (int64_t)s * a.xs;
  const char *xs = reinterpret_cast<const char *>(xsig);
  const bool conj = a.ref_conj;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int j = tj * 64 + lane;             // one alias row per lane
  const int jm = j & (rows - 1);
  for (int q = 0; q < RI / 4; q++) {
    asm volatile("" ::: "memory");
    const int il = wid * (RI / 4) + q;     // one residue per wavefront
    if (ti * RI + il >= nc) break;         // wave-uniform; no barrier inside
    stage[lane * (RI + 1) + il] =
        rows_median<LMAX, FULL, FP32, RI>(xs, t_base, t_q, t_g, t_den, il, jm, rows, L, conj,
                                          a.conj_sign);
  }
  __syncthreads();
  // row by row: RI consecutive output slots per row
  for (int e = threadIdx.x; e < 64 * RI; e += blockDim.x) {
    const int jr = e / RI, il = e % RI;
    const int i = ti * RI + il, jj = tj * 64 + jr;
    if (i < nc && jj < rows)
      write_result(a, s, (int64_t)jj * nc + i, (unsigned)jj * (unsigned)a.W + (unsigned)approved[i],
                   stage[jr * (RI + 1) + il]);
  }
}

// v1: one thread per hit of the vote list.
template <int LMAX, bool FP32>
__global__ void __launch_bounds__(kBlock) k_estimate_hits(V12Launch a, int64_t chunks)
{
  __shared__ unsigned s_ai[LMAX];
  int s;
  int64_t c;
  if (!xcd_block(blockIdx.x, a.S, chunks, s, c)) return;
  const int L = a.L;
  const int stride = 2 * L + max(a.comb_loops, 1);
  if (threadIdx.x < L)
    s_ai[threadIdx.x] = (unsigned)a.params[(int64_t)s * stride + L + threadIdx.x];
  __syncthreads();
  const int64_t t = c * blockDim.x + threadIdx.x;
  if (t >= a.count[s]) return;
  const unsigned h = (unsigned)a.hits[(int64_t)s * a.hits_cap + t];
  const int64_t sbase = (int64_t)s * a.xs;
  double vr[LMAX], vi[LMAX];
#pragma unroll
  for (int l = 0; l < LMAX; l++) {
    if (l < L) {
      const bool isloc = l < a.Lloc;
      int hashed;
      double2 G;
      hash_direct(a, l, s_ai[l], h, hashed, G);
      const double2 U = load_bucket(a, FP32, isloc,
                                    sbase + loop_offset(l, a.Lloc, a.B_loc, a.B_est) + hashed);
      divide_by_filter(U, G, G.x * G.x + G.y * G.y, a.ref_conj, vr[l], vi[l]);
    } else {
      vr[l] = __longlong_as_double(0x7ff0000000000000ll);
      vi[l] = vr[l];
    }
  }
  write_result(a, s, t, h, lower_median<LMAX>(vr, vi, L));
}

__global__ void k_zero(double2 *out, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = i; j < n; j += stride) out[j] = make_double2(0.0, 0.0);
}

// Kernels whose dynamic LDS can exceed the default 64 KiB opt in to the
// device maximum (once per kernel; see raise_lds_limit).
template <typename K>
void allow_lds(K kernel, size_t bytes)
{
  if (bytes > 65536) raise_lds_limit(reinterpret_cast<const void *>(kernel));
}

int launched(const char *what)
{
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return -1;
  }
  return 0;
}

int launch_failed(const char *what)
{
  const hipError_t e = hipGetLastError();
  set_error(std::string(what) + ": " + (e != hipSuccess ? hipGetErrorString(e)
                                                         : "invalid transform shape"));
  return -1;
}

}  // namespace

// ---------------------------------------------------------------- scratch
// Device memory of the allocation-free exec (sfft_exec_sparse_noalloc): the
// workspace of one signal is carved from it instead of allocated.  Every GPU
// the library runs on holds its own copy for the life of the process.  In
// host code `scratch` is only a stand-in for the symbol: its device address
// comes from hipGetSymbolAddress, once per device, and only that address is
// passed to kernels and copies.
static __device__ __align__(256) char scratch[kNoAllocScratch];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};
}  // namespace

int noalloc_setup()
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("noalloc set-up: no current device");
    return -1;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_done[dev]) return 0;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    set_error(std::string("noalloc set-up: hipGetSymbolAddress: ") + hipGetErrorString(e));
    return -1;
  }
  g_scratch[dev] = ptr;
  g_scratch_done[dev] = true;
  return 0;
}

void *noalloc_scratch(int dev)
{
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch[dev];
}

// Below this many comb transforms per batch they run with kCombWide threads.
constexpr int kFewCombs = 64, kCombWide = 1024;

int comb(const V12Launch &a, hipStream_t st)
{
  if (a.W <= kMaxLdsComb) {
    const size_t lds = (size_t)a.W * sizeof(double2);
    if (a.S * a.comb_loops < kFewCombs) {
      allow_lds(k_comb<kCombWide>, lds);
      hipLaunchKernelGGL(k_comb<kCombWide>, dim3(a.S * a.comb_loops), dim3(kCombWide), lds, st,
                         a);
    } else {
      allow_lds(k_comb<kBlock>, lds);
      hipLaunchKernelGGL(k_comb<kBlock>, dim3(a.S * a.comb_loops), dim3(kBlock), lds, st, a);
    }
    if (launched("comb") != 0) return -1;
  } else {
    const int64_t total = (int64_t)a.S * a.comb_loops * a.W;
    hipLaunchKernelGGL(k_comb_gather, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                       a);
    if (launched("comb_gather") != 0) return -1;
    if (radix2_fft64(a.comb_buf, a.comb_tmp, a.W, (int64_t)a.S * a.comb_loops, 1, a.W, a.W,
                     a.tw64, a.tw_nmax, st) != 0) {
      set_error("comb FFT launch failed");
      return -1;
    }
    hipLaunchKernelGGL(k_comb_select_global, dim3(a.S * a.comb_loops), dim3(kBlock), 0, st, a);
    if (launched("comb_select") != 0) return -1;
  }
  if (a.comb_loops > 1) {
    const size_t lb = (size_t)((a.W + 31) / 32) * 4;
    hipLaunchKernelGGL(k_comb_merge, dim3(a.S), dim3(kBlock), lb, st, a);
    if (launched("comb_merge") != 0) return -1;
  }
  return 0;
}

// Loops of one bucket count folded by k_fold_chunks.
static int fold_chunks(const V12Launch &a, hipStream_t st, int l0, int nl, int B)
{
  FoldChunks f;
  f.l0 = l0;
  f.nl = nl;
  f.B = B;
  f.logB = 31 - __builtin_clz((unsigned)B);
  f.res = std::min(kResPer, B);
  f.groups = B / f.res;
  const int rows = a.n / B;
  const int seg = std::min(rows, a.seg_rows);
  f.seg_log = 31 - __builtin_clz((unsigned)seg);
  f.nseg = rows / seg;
  f.tR = 0;
  if (l0 < a.Lloc) f.tR = std::max(f.tR, a.tR_loc);
  if (l0 + nl > a.Lloc) f.tR = std::max(f.tR, a.tR_est);
  if (f.tR > 32) {  // tap masks are 32 bits; tolerances >= 5e-9 give at most 27 taps
    set_error("fold: more than 32 taps per bucket");
    return -1;
  }
  const int threads = (f.res * nl + 63) / 64 * 64;
  const size_t lds = (size_t)f.res * nl * f.tR * sizeof(float2);
  allow_lds(k_fold_chunks, lds);
  const dim3 grid((unsigned)((int64_t)a.S * f.groups));
  hipLaunchKernelGGL(k_fold_chunks, grid, dim3(threads), lds, st, a, f);
  return launched("fold_chunks");
}

int fold_fft(const V12Launch &a, hipStream_t st, int part)
{
  // loops [0, n64) in FP64, [n64, L) in FP32 (folded by k_fold_chunks, with
  // the FFTs in part 1 whatever the schedule)
  const int n64 = !a.fp32 ? a.L : (a.loc64 ? a.Lloc : 0);
  const int B64 = n64 == 0 ? 0 : (n64 > a.Lloc ? max(a.B_loc, a.B_est) : a.B_loc);
  int launches = 0;
  if (part == 0) {
    if (n64 > 0 && a.fold_mode >= 1) {
      const FoldItems f = fold_items(a, 0, n64);
      const int64_t g =
          std::min<int64_t>(a.stream_blocks, (f.total + kFoldWaves - 1) / kFoldWaves);
      hipLaunchKernelGGL(k_fold_stream, dim3((unsigned)g), dim3(64 * kFoldWaves), 0, st, a, f);
      if (launched("fold_stream64") != 0) return -1;
      launches++;
    } else if (n64 > 0) {
      allow_lds(k_fold_fft64, (size_t)B64 * sizeof(double2));
      hipLaunchKernelGGL(k_fold_fft64, dim3(a.S * n64), dim3(kBlock),
                         (size_t)B64 * sizeof(double2), st, a, 0, n64);
      if (launched("fold_fft64") != 0) return -1;
      launches++;
    }
    if (n64 < a.L) {
      // the FP32 loops, in runs of one bucket count
      auto run = [&](int l0, int nl, int B) { return fold_chunks(a, st, l0, nl, B); };
      if (n64 < a.Lloc && a.B_loc != a.B_est) {
        if (run(n64, a.Lloc - n64, a.B_loc) != 0 || run(a.Lloc, a.L - a.Lloc, a.B_est) != 0)
          return -1;
        launches += 2;
      } else {
        if (run(n64, a.L - n64, n64 < a.Lloc ? a.B_loc : a.B_est) != 0) return -1;
        launches++;
      }
    }
    return launches;
  }
  // part 1: the bucket FFTs of every loop not fused into its fold
  if (a.rows_layout) {
    // v2: every loop in one precision, written in the alias-row layout
    const int Bm = max(a.B_loc, a.B_est);
    if (a.fp32) {
      allow_lds(k_bucket_fft_rows<true>, (size_t)Bm * sizeof(float2));
      hipLaunchKernelGGL(k_bucket_fft_rows<true>, dim3(a.S * a.L), dim3(kBlock),
                         (size_t)Bm * sizeof(float2), st, a, 0, a.L);
    } else {
      allow_lds(k_bucket_fft_rows<false>, (size_t)Bm * sizeof(double2));
      hipLaunchKernelGGL(k_bucket_fft_rows<false>, dim3(a.S * a.L), dim3(kBlock),
                         (size_t)Bm * sizeof(double2), st, a, 0, a.L);
    }
    if (launched("bucket_fft_rows") != 0) return -1;
    return 1;
  }
  // the location and estimation loops as batches of one bucket count and
  // precision: in LDS up to the LDS size, else radix-2 runs plus global
  // passes; FP64 loops fused into their fold (schedule 0) are done
  const int f0 = a.fold_mode == 0 ? n64 : 0;  // first loop still to transform
  int groups[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  int ng = 0;
  auto add_run = [&](int l0, int l1) {
    // split [l0, l1) at Lloc (bucket counts) and n64 (precision)
    int cuts[4] = {l0, std::min(std::max(a.Lloc, l0), l1), std::min(std::max(n64, l0), l1), l1};
    std::sort(cuts, cuts + 4);
    for (int c = 0; c < 3; c++)
      if (cuts[c + 1] > cuts[c]) {
        groups[ng][0] = cuts[c];
        groups[ng][1] = cuts[c + 1];
        groups[ng][2] = cuts[c] < a.Lloc ? a.B_loc : a.B_est;
        ng++;
      }
  };
  add_run(f0, a.L);
  for (int gi = 0; gi < ng; gi++) {
    int l0 = groups[gi][0], l1 = groups[gi][1];
    const int B = groups[gi][2];
    // merge with the next run when bucket count and precision agree
    while (gi + 1 < ng && groups[gi + 1][2] == B && (groups[gi + 1][0] < n64) == (l0 < n64)) {
      l1 = groups[++gi][1];
    }
    const int64_t off = loop_offset_host(l0, a.Lloc, a.B_loc, a.B_est);
    const int rc = l0 < n64
                       ? radix2_fft64(a.xs64 + off, a.tmp64 ? a.tmp64 + off : nullptr, B,
                                      a.S, l1 - l0, a.xs, B, a.tw64, a.tw_nmax, st)
                       : radix2_fft32(a.xs32 + off, a.tmp32 ? a.tmp32 + off : nullptr, B,
                                      a.S, l1 - l0, a.xs, B, a.tw32, a.tw_nmax, st);
    if (rc != 0) return launch_failed("bucket FFT");
    launches++;
  }
  return launches;
}

int select_vote_v1(const V12Launch &a, hipStream_t st)
{
  if (a.B_loc <= kMaxLdsSelect) {
    allow_lds(k_select<false>, (size_t)a.B_loc * sizeof(double));
    hipLaunchKernelGGL(k_select<false>, dim3(a.S * a.Lloc), dim3(kBlock),
                       (size_t)a.B_loc * sizeof(double), st, a);
  } else {
    hipLaunchKernelGGL(k_select<true>, dim3(a.S * a.Lloc), dim3(kBlock), 0, st, a);
  }
  if (launched("select") != 0) return -1;
  const size_t lb = (size_t)a.Lloc * ((a.B_loc + 31) / 32) * 4;
  const int chunks = vote_chunks(a.n, a.B_loc, a.num, a.Lloc, a.thr);
  const dim3 grid((unsigned)((int64_t)a.S * chunks));
  hipLaunchKernelGGL(k_vote_count, grid, dim3(kBlock), lb, st, a, chunks);
  if (launched("vote_count") != 0) return -1;
  hipLaunchKernelGGL(k_vote_scan, dim3(a.S), dim3(kBlock), 0, st, a, chunks);
  if (launched("vote_scan") != 0) return -1;
  hipLaunchKernelGGL(k_vote_write, grid, dim3(kBlock), lb, st, a, chunks);
  return launched("vote_write");
}

int comb_vote_keys(const V12Launch &a, int64_t *keys, int64_t stride, hipStream_t st)
{
  if (a.B_loc <= kMaxLdsSelect) {
    allow_lds(k_select<false>, (size_t)a.B_loc * sizeof(double));
    hipLaunchKernelGGL(k_select<false>, dim3(a.S * a.Lloc), dim3(kBlock),
                       (size_t)a.B_loc * sizeof(double), st, a);
  } else {
    hipLaunchKernelGGL(k_select<true>, dim3(a.S * a.Lloc), dim3(kBlock), 0, st, a);
  }
  if (launched("select (v2 votes)") != 0) return -1;
  const int64_t chunks = (a.est_cap + kBlock - 1) / kBlock;
  const size_t lb = (size_t)a.Lloc * ((a.B_loc + 31) / 32) * 4;
  allow_lds(k_comb_vote_keys, lb);
  hipLaunchKernelGGL(k_comb_vote_keys, dim3((unsigned)((int64_t)a.S * chunks)), dim3(kBlock), lb,
                     st, a, chunks, keys, stride);
  return launched("comb_vote_keys");
}

// Blocks that keep every CU busy with several wavefronts.
constexpr int64_t kFillBlocks = 1024;

template <int LMAX, int RI>
void launch_rows(const V12Launch &a, hipStream_t st, int tiles_j)
{
  const int tiles_i = (a.comb_loops * a.num + RI - 1) / RI;
  const dim3 grid((unsigned)xcd_grid(a.S, (int64_t)tiles_i * tiles_j));
  if (a.fp32) {
    if (a.L == LMAX)
      hipLaunchKernelGGL((k_estimate_rows<LMAX, true, true, RI>), grid, dim3(kBlock), 0, st, a,
                         tiles_i, tiles_j);
    else
      hipLaunchKernelGGL((k_estimate_rows<LMAX, false, true, RI>), grid, dim3(kBlock), 0, st, a,
                         tiles_i, tiles_j);
  } else {
    if (a.L == LMAX)
      hipLaunchKernelGGL((k_estimate_rows<LMAX, true, false, RI>), grid, dim3(kBlock), 0, st, a,
                         tiles_i, tiles_j);
    else
      hipLaunchKernelGGL((k_estimate_rows<LMAX, false, false, RI>), grid, dim3(kBlock), 0, st,
                         a, tiles_i, tiles_j);
  }
}

template <int LMAX>
void launch_estimate(const V12Launch &a, hipStream_t st)
{
  if (a.comb_loops > 0 && a.rows_layout) {
    const int tiles_j = (a.n / a.W + 63) / 64;
    const int64_t big = (int64_t)a.S * ((a.comb_loops * a.num + kRI - 1) / kRI) * tiles_j;
    if (big >= kFillBlocks) launch_rows<LMAX, kRI>(a, st, tiles_j);
    else launch_rows<LMAX, kRIsmall>(a, st, tiles_j);
  } else if (a.comb_loops > 0) {
    const int tiles_i = (a.comb_loops * a.num + kTI - 1) / kTI;
    const int tiles_j = (a.n / a.W + kTJ - 1) / kTJ;
    const dim3 grid((unsigned)xcd_grid(a.S, (int64_t)tiles_i * tiles_j));
    const bool allfast = a.W % (a.n / a.B_loc) == 0 && a.W % (a.n / a.B_est) == 0;
    if (a.fp32) {
      if (allfast)
        hipLaunchKernelGGL((k_estimate_comb<LMAX, true, true>), grid, dim3(kBlock), 0, st, a,
                           tiles_i, tiles_j);
      else
        hipLaunchKernelGGL((k_estimate_comb<LMAX, true, false>), grid, dim3(kBlock), 0, st,
                           a, tiles_i, tiles_j);
    } else {
      if (allfast)
        hipLaunchKernelGGL((k_estimate_comb<LMAX, false, true>), grid, dim3(kBlock), 0, st,
                           a, tiles_i, tiles_j);
      else
        hipLaunchKernelGGL((k_estimate_comb<LMAX, false, false>), grid, dim3(kBlock), 0, st,
                           a, tiles_i, tiles_j);
    }
  } else {
    const int64_t chunks = (a.est_cap + kBlock - 1) / kBlock;
    const dim3 grid((unsigned)xcd_grid(a.S, chunks));
    if (a.fp32)
      hipLaunchKernelGGL((k_estimate_hits<LMAX, true>), grid, dim3(kBlock), 0, st, a, chunks);
    else
      hipLaunchKernelGGL((k_estimate_hits<LMAX, false>), grid, dim3(kBlock), 0, st, a, chunks);
  }
}

int estimate(const V12Launch &a, hipStream_t st)
{
  // the smallest median network that holds the loops
  if (a.L <= 8) launch_estimate<8>(a, st);
  else if (a.L <= 12) launch_estimate<12>(a, st);
  else if (a.L <= 16) launch_estimate<16>(a, st);
  else if (a.L <= 20) launch_estimate<20>(a, st);
  else if (a.L <= 24) launch_estimate<24>(a, st);
  else launch_estimate<32>(a, st);
  return launched("estimate");
}

int zero_dense(double2 *out, int64_t n, hipStream_t st)
{
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, st, out, n);
  return launched("zero_dense");
}

}  // namespace k
}  // namespace sfa
