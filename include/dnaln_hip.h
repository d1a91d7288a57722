/* dnaln_hip.h — C ABI of the MI355X (gfx950) seed-chain-extend path of dnaln.
 *
 * Plain C, no STL or HIP types in the signatures: the reference's C++ host (src/mapper.cpp)
 * reaches the GPU only through these calls. Every call returns 0 (DNALN_OK) or a negative
 * status; dnaln_hip_last_error() gives the message of the calling thread's last failure.
 * A context is used by one host thread at a time (the reference's Mapper and Aligner are not
 * thread-safe either, src/mapper.hpp:51-94).
 *
 * Reference interfaces replaced (paths relative to the dnaln checkout):
 *   dnaln_hip_create / _destroy   Mapper::Mapper (src/mapper.cpp:8-12) with the resident
 *                                 Reference (src/reference.hpp:16-30) and MinimizerIndex
 *                                 (src/index.hpp:21-39)
 *   dnaln_hip_align               Aligner::align (src/align.hpp:62-68)
 *   dnaln_hip_map_batch           Mapper::map (src/mapper.hpp:54, src/mapper.cpp:75-223), minus
 *                                 SAM text, which stays on the host (append_sam)
 *   dnaln_hip_batch_*             the same pipeline on device-resident batches (plain
 *                                 per-chunk launches, or the batched form: a fixed kernel
 *                                 sequence that reads its task counts from device memory),
 *                                 used by the serving loop and the benchmark
 */
#ifndef DNALN_HIP_H
#define DNALN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DNALN_HIP_ABI_VERSION 1

enum {
  DNALN_OK = 0,
  DNALN_E_INVALID = -1,   /* bad argument or parameter outside the validated range */
  DNALN_E_CAPACITY = -2,  /* a caller-provided output buffer is too small */
  DNALN_E_DEVICE = -3,    /* HIP runtime error */
  DNALN_E_NOMEM = -4,     /* device or host allocation failed */
  DNALN_E_INTERNAL = -5,  /* an internal consistency check failed */
  DNALN_E_PENDING = -6,   /* dnaln_hip_batch_poll: still running */
  DNALN_E_IO = -7         /* index file: unreadable, truncated, corrupt, or built from other codes */
};

/* ScoreParams (src/align.hpp:34-40). A gap of length L costs gap_open + L * gap_ext. */
typedef struct {
  int32_t match, mismatch, gap_open, gap_ext, ambig, bw;
} dnaln_score_params;

/* ChainParams (src/chain.hpp:30-36). */
typedef struct {
  int32_t max_gap, band, lookback, min_score, min_anchors;
} dnaln_chain_params;

/* MapParams (src/mapper.hpp:25-35); `simd` and `prefetch` have no meaning on the device. */
typedef struct {
  int32_t k, w;
  uint32_t max_occ;
  dnaln_score_params sc;
  dnaln_chain_params ch;
  int32_t max_cand;
  double pri_ratio;
  int32_t min_aln_score;
} dnaln_map_params;

/* Contig (src/reference.hpp:10-14) without the name (names stay on the host). */
typedef struct {
  uint64_t offset;
  uint32_t len;
  uint32_t reserved;
} dnaln_contig;

/* MinimizerIndex layout (src/index.hpp:49-56): 16-byte buckets, linear probing, empty key ~0;
 * pos[] holds (global_pos << 1 | strand) grouped by key. n_buckets is a power of two. */
typedef struct {
  uint64_t key;
  uint32_t offset, count;
} dnaln_bucket;

typedef struct {
  const dnaln_bucket* table;
  uint64_t n_buckets;
  const uint32_t* pos;
  uint64_t n_pos;
  int32_t k, w;
} dnaln_index_view;

/* AlignTask (src/align.hpp:42-47) as offsets into one host byte pool of codes 0..4. The target
 * should hold qlen + 2*bw codes; a shorter one is N-padded, as in the reference. */
typedef struct {
  uint64_t q_off;
  uint64_t t_off;
  int32_t qlen;
  int32_t tlen;
} dnaln_task;

/* AlignResult (src/align.hpp:49-60); the CIGAR (M/I/D, BAM encoding len << 4 | op) is stored
 * at cigar_pool[cigar_off .. cigar_off + cigar_len). */
typedef struct {
  int32_t score, qb, qe, tb, te, nm;
  uint64_t cigar_off;
  uint32_t cigar_len;
  uint32_t reserved;
} dnaln_result;

/* MapRecord (src/mapper.hpp:37-44): contig -1 = unmapped; pos is 0-based within the contig;
 * the CIGAR includes soft clips. */
typedef struct {
  int32_t contig;
  uint8_t rev, mapq;
  uint16_t reserved;
  int64_t pos;
  int32_t score, nm, n_cand;
  uint32_t cigar_len;
  uint64_t cigar_off;
} dnaln_record;

/* MapStats (src/mapper.hpp:46-49) plus device timings of the last batch (ms). cells counts
 * band cells as the reference does (rank pass + traceback pass). anchors counts only the
 * anchors of reads on the global-memory seed path (the LDS path keeps its anchors on chip);
 * global_path_reads counts those reads in the batched form, and fused_retry_reads the reads
 * that the LDS path's 32-anchor shape declined and its 64-anchor shape took again. */
typedef struct {
  uint64_t reads, mapped, anchors, candidates, cells, rank_tasks, tb_tasks;
  double ms_seed, ms_chain, ms_rank, ms_traceback, ms_total;
  uint64_t global_path_reads, fused_retry_reads;
} dnaln_map_stats;

/* Fills the reference's defaults (k 15, w 10, max_occ 500, 2/4/4/2/1 bw 16, chain
 * 5000/16/50/20/2, max_cand 10, pri_ratio 0.5, min score 40). */
void dnaln_hip_default_params(dnaln_map_params* p);

/* Uploads the reference codes (0..4) and the index once; they stay resident in HBM.
 * Validates the parameters as src/cli.hpp:93-95 and src/index.cpp:11-12 do. With index NULL
 * the index is built on the GPU from the uploaded codes (as dnaln_hip_index_build; k <= 16) in
 * place of the host's MinimizerIndex::build before the Mapper is made (src/main.cpp:34-36). */
int dnaln_hip_create(const dnaln_map_params* params, const uint8_t* ref_codes, uint64_t ref_len,
                     const dnaln_contig* contigs, int32_t n_contigs, const dnaln_index_view* index, int32_t device,
                     struct dnaln_hip_ctx** out);
void dnaln_hip_destroy(struct dnaln_hip_ctx* ctx);
const char* dnaln_hip_last_error(void);
/* Free and total device memory of the context's GPU, bytes (no reference counterpart; the
 * serving loop sizes its HBM budget from it). */
int dnaln_hip_device_memory(struct dnaln_hip_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);
/* The context's parameters after validation (the chain band clamped to the SW band). */
int dnaln_hip_context_params(struct dnaln_hip_ctx* ctx, dnaln_map_params* out);
/* Device allocations the library has made in this process (count and bytes), for callers that
 * check that a warmed-up path allocates nothing (no reference counterpart). */
void dnaln_hip_alloc_counters(uint64_t* events, uint64_t* bytes);
/* Measured packed-int16 VALU peak of a device (no reference counterpart; the roofline the SW
 * fill is judged against): sixteen independent chains of v_pk_add_u16 + v_pk_max_i16 per lane,
 * eight waves per SIMD on every CU, `iters` iterations; *tops = int16 operations (two per packed
 * instruction and lane) per second / 1e12, *ms = the timed launch. */
int dnaln_hip_peak_int16(int32_t device, int32_t iters, double* tops, double* ms);

/* MinimizerIndex::build (src/index.cpp:10-45) on the GPU: every contig sketched in segments of
 * 1,024 k-mer starts, one per thread; the occurrences radix-sorted by (hash, pos); each key placed
 * by one thread with atomic linear probing. The bucket count and pos[] equal the host build's; a
 * key may sit in a different slot of its probe chain, so every lookup (src/index.hpp:24-38)
 * returns the same slice. Needs k <= 16. The result stays on the device until downloaded into
 * host arrays of n_buckets buckets and n_pos positions (dnaln_hip_index_info). No reference
 * counterpart for the handle: the reference builds its index on the host at start-up
 * (src/main.cpp:34-36). */
typedef struct dnaln_hip_index dnaln_hip_index;
int dnaln_hip_index_build(const uint8_t* ref_codes, uint64_t ref_len, const dnaln_contig* contigs, int32_t n_contigs,
                          int32_t k, int32_t w, int32_t device, dnaln_hip_index** out);
int dnaln_hip_index_info(const dnaln_hip_index* ix, uint64_t* n_buckets, uint64_t* n_pos, uint64_t* n_keys,
                         uint32_t* max_count, double* build_ms);
int dnaln_hip_index_download(const dnaln_hip_index* ix, dnaln_bucket* table, uint32_t* pos);
void dnaln_hip_index_free(dnaln_hip_index* ix);
/* Page-locked host memory (hipHostMalloc) for staging reads and records, so that batch uploads
 * and downloads run at full PCIe rate and overlap with the device (no reference counterpart). */
int dnaln_hip_host_alloc(uint64_t bytes, void** out);
void dnaln_hip_host_free(void* p);
/* Sizes of the context's resident index, and the wall time of its GPU build in ms (0 when the
 * index came from the host). */
int dnaln_hip_context_index(struct dnaln_hip_ctx* ctx, uint64_t* n_buckets, uint64_t* n_pos, double* build_ms);

/* Aligner::align: n independent banded alignments; traceback = 0 computes the score only.
 * sp may be NULL (use the context's scoring). Results are bit-identical to the reference's
 * scalar and SIMD kernels. */
int dnaln_hip_align(struct dnaln_hip_ctx* ctx, const dnaln_score_params* sp, const dnaln_task* tasks, uint64_t n,
                    const uint8_t* seq_pool, uint64_t pool_len, int traceback, dnaln_result* out,
                    uint32_t* cigar_pool, uint64_t cigar_cap);

/* Device scratch of dnaln_hip_align_static: one array of this size per GPU the library runs on,
 * held until the process ends. */
#define DNALN_STATIC_SCRATCH_BYTES (64ull << 20)

/* Aligner::align exactly as dnaln_hip_align (same arguments, results and errors), for callers
 * that cannot allocate device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc. No counterpart in the reference. It makes no device
 * allocation: every buffer of the call (the copied tasks and sequence pool, results, traceback
 * directions, CIGARs) lies in one static device array of DNALN_STATIC_SCRATCH_BYTES per GPU,
 * and a call whose buffers do not fit returns DNALN_E_CAPACITY. Every call on a GPU shares that
 * GPU's array, so calls that overlap on one GPU (from several threads or contexts) race; the
 * caller serialises them. Nothing of a previous call's contents is read. */
int dnaln_hip_align_static(struct dnaln_hip_ctx* ctx, const dnaln_score_params* sp, const dnaln_task* tasks,
                           uint64_t n, const uint8_t* seq_pool, uint64_t pool_len, int traceback, dnaln_result* out,
                           uint32_t* cigar_pool, uint64_t cigar_cap);

/* Mapper::map on n reads (forward-strand codes 0..4 at read_codes[offs[i] .. offs[i]+lens[i]),
 * name_fnv1a[i] = FNV-1a of the read name for the equal-best tie-break). Records are identical
 * to the reference's. stats may be NULL. */
int dnaln_hip_map_batch(struct dnaln_hip_ctx* ctx, const uint8_t* read_codes, const uint64_t* offs,
                        const int32_t* lens, const uint32_t* name_fnv1a, uint64_t n, dnaln_record* out,
                        uint32_t* cigar_pool, uint64_t cigar_cap, dnaln_map_stats* stats);

/* ------------------------------------------------------------------ device-resident batches */
enum { DNALN_MODE_PLAIN = 0, DNALN_MODE_BATCHED = 1 };

typedef struct {
  int32_t mode;             /* DNALN_MODE_PLAIN: one launch sequence per chunk of chunk_reads, with
                               host reads of the task counts between launches;
                               DNALN_MODE_BATCHED: the whole batch as one chunk: reads bucketed
                               by length into the fused seed + chain kernel's LDS shapes (the
                               rest on the global-memory path), SW tasks sorted by length; the
                               launches are separate, with a few host reads of counts (`graph`
                               selects the forms without them) */
  uint32_t chunk_reads;     /* plain form: reads per chunk (65536 by default) */
  uint64_t max_reads;       /* capacity of the batch */
  uint64_t max_bases;
  uint32_t anchors_per_read;/* batched form: first anchor capacity of the global path, per read
                               (0: estimated from the read lengths); a run that needs more grows
                               it and runs again */
  int32_t tune;             /* batched form: 1 = run the autotuner on the first batch */
  int32_t len_split[2];     /* batched form: read-length limits of the two LDS shapes of the fused
                               seed+chain kernel; longer reads take the global-memory path
                               (0: the largest lengths the shapes hold on average; -1: bucket off) */
  uint64_t dir_budget;      /* bytes of traceback direction words per SW launch
                               (0: min(16 GiB, free HBM / 8)) */
  int32_t walk_kind;        /* traceback walk: 0 one thread per task, 1 one block per fill wave
                               through LDS windows of flag words */
  int32_t graph;            /* batched form: 0 = the form above (default); 2 = one fixed kernel
                               sequence whose grids cover the batch's worst case and whose blocks
                               read the live counts from device memory (no host read until the
                               batch is polled); 1 = that sequence captured as one HIP graph per
                               plan. 1 and 2 are refused (DNALN_E_INVALID) by the device build:
                               their first run faulted the GPU and the cause is not found (the
                               CPU emulation runs them) */
  int32_t sw_lanes;         /* lanes per pair of alignments in the packed SW fill: 1 = the band in
                               one lane's registers; 2 (bw 32) or 4 (bw 64) = the band split over
                               that many lanes, walked one thread per task (walk_kind is then
                               ignored); 0 = the default (the split for the batched form where the
                               band width has one, else 1); a split the band width lacks is
                               refused (DNALN_E_INVALID) */
  int32_t reserved;         /* 0 */
} dnaln_batch_config;

void dnaln_hip_default_batch_config(dnaln_batch_config* c);

/* A batch owns device buffers for inputs, workspace and outputs, and a HIP stream. */
int dnaln_hip_batch_create(struct dnaln_hip_ctx* ctx, const dnaln_batch_config* cfg, struct dnaln_hip_batch** out);
void dnaln_hip_batch_destroy(struct dnaln_hip_batch* b);
/* Asynchronous H2D copy of n reads into the batch (host buffers must stay valid until sync). */
int dnaln_hip_batch_upload(struct dnaln_hip_batch* b, const uint8_t* read_codes, const uint64_t* offs,
                           const int32_t* lens, const uint32_t* name_fnv1a, uint64_t n);
/* Sizes the batch's device buffers for batches of up to n reads, `bases` bases and reads of up
 * to maxlen bases (anchors_per_read anchors per read on the global path), so that running
 * such batches allocates nothing more in the common case. No reference counterpart: for
 * serving slots, which must not stall the device with allocations. */
int dnaln_hip_batch_reserve(struct dnaln_hip_batch* b, uint64_t n, uint64_t bases, int32_t maxlen,
                            uint32_t anchors_per_read);
/* Runs the pipeline on the resident batch, asynchronously where the mode allows. */
int dnaln_hip_batch_run(struct dnaln_hip_batch* b);
/* Waits for the batch (wait = 1) or reports DNALN_E_PENDING (wait = 0). */
int dnaln_hip_batch_poll(struct dnaln_hip_batch* b, int wait);
/* Copies records and CIGARs to the host (after poll returned DNALN_OK). */
int dnaln_hip_batch_download(struct dnaln_hip_batch* b, dnaln_record* out, uint32_t* cigar_pool, uint64_t cigar_cap,
                             dnaln_map_stats* stats);
/* A candidate locus of one read after seeding and chaining (Mapper::Cand, src/mapper.hpp:61-67):
 * strand, contig, diagonal (the window starts at diag - bw, src/mapper.cpp:60) and the score of
 * the chain it came from (Chain::score, src/chain.hpp:22-28). */
typedef struct {
  uint32_t rev;
  int32_t contig;
  int64_t diag;
  int32_t chain_score;
  int32_t reserved;
} dnaln_candidate;

/* The candidates the last run's seed and chain stages left for the rank pass (the state of
 * Mapper::cands_ after seed_and_chain, src/mapper.cpp:14-63): n_cand[i], and candidate c of read
 * i at cands[i * per_read + c] in the reference's order; per_read must be at least max_cand.
 * The plain form keeps one chunk's candidates, so there the batch must fit one chunk. For
 * checks of the seed and chain stages; the mapping path never needs it. */
int dnaln_hip_batch_candidates(struct dnaln_hip_batch* b, int32_t* n_cand, dnaln_candidate* cands, uint64_t per_read);
/* Device bytes a batch of this configuration holds (for admission by HBM budget). */
uint64_t dnaln_hip_batch_device_bytes(struct dnaln_hip_ctx* ctx, const dnaln_batch_config* cfg);
/* Time of the last run measured with HIP events on the batch stream, per family (ms):
 * [0] seed, [1] chain, [2] rank fill, [3] traceback fill, [4] traceback walk, [5] whole run.
 * The fill and walk families time exactly their kernel launches. launches (may be NULL)
 * receives the number of kernel launches of families 2-4. Arrays hold DNALN_TIMING_FAMILIES. */
#define DNALN_TIMING_FAMILIES 6
int dnaln_hip_batch_timings(struct dnaln_hip_batch* b, double* ms, uint64_t* launches);

/* Autotuner of the batched form (no counterpart in the reference, whose CPU path has no such
 * choices). Runs the resident batch under each candidate setting and keeps the fastest:
 * read-length bucket limits (which reads chain one per lane in the fused LDS kernel, in which
 * LDS shape, and which one per wave through global memory) and the traceback direction budget
 * per SW launch. Every candidate gives the same records; only the time differs. */
typedef struct {
  int32_t len_split[2];     /* chosen bucket limits (0 = bucket unused) */
  uint64_t dir_budget;      /* chosen direction budget, bytes */
  int32_t walk_kind;        /* chosen walk */
  int32_t n_tried;
  int32_t chosen;           /* index into ms[] */
  int32_t sw_lanes;         /* chosen lanes per pair of the packed SW fill */
  double ms[16];            /* whole-run time of each candidate, in the order tried */
  int32_t cand_split[16][2];
  uint64_t cand_budget[16];
  int32_t cand_walk[16];
  int32_t cand_lanes[16];
} dnaln_tune_report;

int dnaln_hip_batch_tune(struct dnaln_hip_batch* b, dnaln_tune_report* report);

#ifdef __cplusplus
}
#endif

#endif
