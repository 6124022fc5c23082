/*
 * sfft_amd serving loop: continuous batching of single-signal sFFT requests
 * of any (n, k, version, precision) on one GPU.
 *
 * No counterpart in the reference: its harnesses (src/timing_many.cc:85-96,
 * 181-190) time sfft_exec_many over one plan.  A request here runs exactly
 * the plan sfft_make_plan(n, k, version) builds, with private RNG streams
 * seeded like srandom(random_seed) / srand48(drand48_seed) (as
 * sfft_plan_seed), so its result does not depend on which other requests
 * share its batch or the device.
 *
 * Scheduling: worker threads, each with its own HIP stream and workspace,
 * take the oldest queued request, batch it with the queued requests of the
 * same (n, k, version, precision) (up to max_batch) and run the batch;
 * batches of different workers overlap on the device.  Plans (filter taps,
 * windows, twiddles) are built on first use and stay device-resident.
 * Admission is by HBM footprint: a batch starts only when plan tables plus
 * every worker's workspace (including the batch's) fit in hbm_budget bytes;
 * the batch is cut down to fit, and unused plans are evicted (least recently
 * used first) to make room.
 */
#ifndef SFFT_AMD_SFFT_SERVER_H
#define SFFT_AMD_SFFT_SERVER_H

#include "sfft.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sfft_server sfft_server;

typedef struct sfft_request {
  int n, k;
  sfft_version version;
  sfft_precision precision;
  const complex_t *in;       /* device pointer, n values; untouched until done */
  int32_t *freq;             /* device, capacity entries (or NULL) */
  complex_t *val;            /* device, capacity entries (or NULL) */
  int64_t capacity;          /* >= sfft_output_capacity(n, k, version) */
  unsigned int random_seed;  /* this request's random() stream */
  long drand48_seed;         /* this request's drand48() stream */
} sfft_request;

typedef struct sfft_server_stats {
  int64_t submitted, completed, rejected, failed;
  int64_t batches, batched_requests;  /* batches run; requests they held */
  int64_t plans_created, plans_evicted, plans_resident;
  int64_t hbm_budget, hbm_in_use, hbm_peak;  /* tables + workspaces, bytes */
  int64_t max_concurrent_batches;            /* most batches on the GPU at once */
  double plan_build_ms;                      /* total time building plans */
} sfft_server_stats;

/* Located-set capacity a request of (n, k, version) needs, or -1 (and
 * sfft_last_error) when the plan cannot be built for it. */
int64_t sfft_output_capacity(int n, int k, sfft_version version);

/* workers >= 1 threads (each a HIP stream); max_batch >= 1 requests per
 * batch; hbm_budget in bytes (0: 80% of the device's free memory).  Uses
 * the calling thread's current device. */
sfft_server *sfft_server_create(int64_t hbm_budget, int workers, int max_batch);

/* Builds (or finds) the plan of a key ahead of its requests.  0 or -1. */
int sfft_server_prepare(sfft_server *s, int n, int k, sfft_version version,
                        sfft_precision precision);

/* Queues a request; returns its ticket (>= 0), or -1 when it is rejected
 * (infeasible (n, k, version), too small a capacity, null input) with the
 * reason in sfft_last_error. */
int64_t sfft_server_submit(sfft_server *s, const sfft_request *r);

/* 0 queued or running, 1 done, -1 failed or unknown ticket. */
int sfft_server_poll(sfft_server *s, int64_t ticket);

/* Blocks until the request finishes.  count: located frequencies written
 * (min(count, capacity) entries are valid); latency_ms: submit to the end
 * of its batch on the device.  Forgets the ticket.  0, or -1 with the
 * request's error in sfft_last_error. */
int sfft_server_wait(sfft_server *s, int64_t ticket, int64_t *count, double *latency_ms);

void sfft_server_get_stats(sfft_server *s, sfft_server_stats *out);

/* Finishes every queued request, then frees the plans and workspaces. */
void sfft_server_destroy(sfft_server *s);

#ifdef __cplusplus
}
#endif
#endif
