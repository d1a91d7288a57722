/* dnaln_serve.h — request serving on the resident index (libdnaln_host.so).
 *
 * No counterpart in the reference: its CLI maps one file per process (src/main.cpp:12-75). Here a
 * server keeps the reference and index resident on the GPU (one dnaln_hip_ctx) and maps
 * requests of any size as they arrive, by continuous batching:
 *   - requests wait in arrival order;
 *   - whenever a batch slot (a dnaln_hip_batch with its own HIP stream) is free, the waiting
 *     requests at the head of the queue are packed into one batch, up to max_batch_reads and
 *     max_batch_bases, provided that the batch's estimated device bytes fit into what the HBM
 *     budget has left after the batches in flight;
 *   - each slot has a worker thread that uploads, runs and downloads its batch, so batches on
 *     different slots overlap on the device and on the host; finished batches are split back
 *     into their requests. Slots are sized for their largest batch when the server starts, so
 *     serving allocates no device memory.
 * Every read's record depends only on the read (src/mapper.cpp:75-223), so a request's records
 * are the same alone and in any mix; the GPU tests check exactly that.
 *
 * One host thread drives a server (submit / progress / take), or dnaln_serve_run_trace replays
 * a timed arrival trace and measures per-request latency.
 */
#ifndef DNALN_SERVE_H
#define DNALN_SERVE_H

#include <stddef.h>
#include <stdint.h>

#include "dnaln_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t slots;              /* batches in flight at most (default 6) */
  int32_t mode;               /* DNALN_MODE_BATCHED or DNALN_MODE_PLAIN for every slot */
  uint64_t hbm_budget;        /* device bytes the batches in flight may use (0: free HBM / 2) */
  uint64_t max_batch_reads;   /* reads packed into one batch at most */
  uint64_t max_batch_bases;
  uint64_t dir_budget;        /* traceback direction bytes per SW launch in a slot (0: budget/slots/4) */
  int32_t max_read_len;       /* slots are sized for reads up to this long (longer ones still map) */
  int32_t reserved;
} dnaln_serve_config;

typedef struct {
  uint64_t submitted, completed, batches, reads;
  uint64_t queued_requests;     /* waiting now */
  uint64_t inflight_batches;    /* running now */
  uint64_t max_inflight_batches;/* most batches that were running at once */
  uint64_t inflight_bytes;      /* estimated device bytes of the batches running now */
  uint64_t max_inflight_bytes;
  uint64_t budget_bytes;        /* the HBM budget in force */
  uint64_t deferred_by_budget;  /* admissions postponed because the budget was short */
} dnaln_serve_stats;

void dnaln_serve_default_config(dnaln_serve_config* cfg);
int dnaln_serve_create(struct dnaln_hip_ctx* ctx, const dnaln_serve_config* cfg, struct dnaln_server** out);
void dnaln_serve_destroy(struct dnaln_server* s);

/* Estimated device bytes of one batch of n reads with `bases` bases, the longest maxlen. */
uint64_t dnaln_serve_batch_bytes(const struct dnaln_server* s, uint64_t n, uint64_t bases, int32_t maxlen);

/* Queues a request (the data is copied). A request larger than max_batch_reads/bases or the
 * whole budget is rejected with DNALN_E_CAPACITY. */
int dnaln_serve_submit(struct dnaln_server* s, uint64_t id, const uint8_t* codes, const uint64_t* offs,
                       const int32_t* lens, const uint32_t* name_fnv1a, uint64_t n);

/* Collects finished batches and admits waiting requests into free slots; never blocks.
 * Returns the number of requests that finished during the call, or a negative status. */
int dnaln_serve_progress(struct dnaln_server* s);

/* Pops one finished request: its id, its n records (in submission order) and their CIGARs.
 * n_out receives the read count; records may be NULL to query sizes. DNALN_E_PENDING when no
 * request has finished; DNALN_E_CAPACITY when out or cigar_pool is too small (nothing popped). */
int dnaln_serve_take(struct dnaln_server* s, uint64_t* id, dnaln_record* out, uint64_t out_cap, uint32_t* cigar_pool,
                     uint64_t cigar_cap, uint64_t* n_out, uint64_t* cigar_len_out);

int dnaln_serve_stats_get(const struct dnaln_server* s, dnaln_serve_stats* out);
/* Message of the last failure seen by dnaln_serve_progress (a slot's batch error). */
const char* dnaln_serve_last_error(const struct dnaln_server* s);

/* Replays an arrival trace in real time: request r holds reads [req_first[r], req_first[r] +
 * req_count[r]) of the read set (ranges may repeat) and arrives arrival_s[r] seconds after the
 * start (non-decreasing). Requests are submitted when their time comes, the server progresses
 * in between, and latency_s[r] is the time from arrival to the moment its batch was seen
 * finished. records (may be NULL) gets each read's record at its read-set position (a later
 * request over the same reads writes the same records again), with CIGARs appended to
 * cigar_pool. wall_s receives the time from the first arrival to the last completion. */
int dnaln_serve_run_trace(struct dnaln_server* s, const uint8_t* codes, const uint64_t* offs, const int32_t* lens,
                          const uint32_t* name_fnv1a, uint64_t n_reads, const uint64_t* req_first,
                          const uint64_t* req_count, const double* arrival_s, uint64_t n_req, double* latency_s,
                          dnaln_record* records, uint32_t* cigar_pool, uint64_t cigar_cap, double* wall_s);

#ifdef __cplusplus
}
#endif

#endif
