/* dnaln_host.h — host-side services around the HIP path (libdnaln_host.so).
 *
 *   dnaln_index_build   MinimizerIndex::build (src/index.cpp:10-45): sketch every contig, sort
 *                       occurrences by (hash, pos), place keys with linear probing at load <= 0.5.
 *                       Produces the view dnaln_hip_create uploads.
 *   dnaln_index_build_gpu, dnaln_index_save / dnaln_index_load
 *                       the same index built on the GPU, and written to / read from a file.
 *   dnaln_map_file      reads file -> SAM records with input, device and output overlapped.
 *   dnaln_evaluate, dnaln_truth_check
 *                       accuracy of records and alignments against simulated truth.
 *   dnaln_sam_format    Mapper::append_sam (src/mapper.cpp:231-270) for one record.
 *   dnaln_fnv1a         fnv1a (src/common.hpp:136-140), the equal-best tie-break hash.
 *   dnaln_encode        encode (src/common.hpp:36-38).
 */
#ifndef DNALN_HOST_H
#define DNALN_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "dnaln_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  dnaln_index_view view;  /* table / pos owned by this struct */
  uint64_t n_keys;
  uint32_t max_count;
} dnaln_index;

int dnaln_index_build(const uint8_t* ref_codes, const dnaln_contig* contigs, int32_t n_contigs, int32_t k, int32_t w,
                      dnaln_index* out);
void dnaln_index_free(dnaln_index* idx);

/* The same index built on GPU `device` (dnaln_hip_index_build, k <= 16) and downloaded into
 * host arrays: bucket count, pos[], key count and largest count equal dnaln_index_build's.
 * *build_ms (optional) receives the device build's wall time. */
int dnaln_index_build_gpu(const uint8_t* ref_codes, const dnaln_contig* contigs, int32_t n_contigs, int32_t k,
                          int32_t w, int32_t device, dnaln_index* out, double* build_ms);

/* Serialized index (no reference counterpart: src/main.cpp:34-36 rebuilds it on every run).
 * Little-endian file: a 64-byte header (magic "DNALNIX1", version 1, k, w, largest count,
 * bucket count, position count, key count, a 64-bit digest of the contigs' extents and codes),
 * then the buckets (16 bytes each) and the positions (4 bytes each). Load checks the header,
 * the exact file size, the digest against the codes given, and that every bucket's slice lies
 * inside pos[]; any failure returns DNALN_E_IO. */
int dnaln_index_save(const dnaln_index* idx, const uint8_t* ref_codes, const dnaln_contig* contigs,
                     int32_t n_contigs, const char* path);
int dnaln_index_load(const char* path, const uint8_t* ref_codes, const dnaln_contig* contigs, int32_t n_contigs,
                     dnaln_index* out);
/* The digest stored in an index file: contig offsets and lengths, then their codes. */
uint64_t dnaln_reference_digest(const uint8_t* ref_codes, const dnaln_contig* contigs, int32_t n_contigs);

/* The algorithmic HBM bytes of the seed stage of a batch, for the roofline of the probe kernels (no
 * reference counterpart; a measurement aid): each read sketched as src/minimizer.cpp:5-68 does, each
 * minimizer looked up by linear probing as src/index.hpp:24-38 does, and the occurrences of keys
 * with 1 <= count <= max_occ read as src/mapper.cpp:24-40 reads them. bytes = read_bases +
 * 16 * buckets + 4 * occurrences. Reads are split over `threads` host threads. */
typedef struct {
  uint64_t read_bases, minimizers, buckets, occurrences, bytes;
} dnaln_seed_traffic_counts;
int dnaln_seed_traffic(const dnaln_index_view* idx, const uint8_t* codes, const uint64_t* offs, const int32_t* lens,
                       uint64_t n, uint32_t max_occ, int32_t threads, dnaln_seed_traffic_counts* out);

uint32_t dnaln_fnv1a(const char* s, size_t n);
void dnaln_fnv1a_many(const char* names, const uint64_t* offs, const int32_t* lens, uint64_t n, uint32_t* out);
void dnaln_encode(const char* s, size_t n, uint8_t* out);

/* Appends one SAM line to buf (capacity cap); returns its length, or -1 when cap is short.
 * qual may be NULL (FASTA input). */
long dnaln_sam_format(const char* name, size_t name_len, const char* seq, const char* qual, size_t len,
                      const dnaln_record* rec, const uint32_t* cigar_pool, const char* contig_name, char* buf,
                      size_t cap);

/* Mapper::map + append_sam over a whole FASTA/FASTQ file, plain or gzip (the driver loop of
 * src/main.cpp:47-61), as three stages that overlap on consecutive batches: input (parse,
 * encode, name hashes into page-locked staging; plain 4-line FASTQ parsed by `threads` threads,
 * anything else by the reference's sequential rules), device (two batches of configuration cfg,
 * one thread each: one's copies overlap the other's kernels) and output (SAM text formatted by `threads`
 * threads, written to out_fd in input order; no header). A batch holds about batch_bytes of
 * input (0: 64 MiB). The records are those of dnaln_hip_map_batch + dnaln_sam_format. */
typedef struct {
  uint64_t reads, mapped, bases, sam_bytes, batches;
  double s_wall;    /* the whole call */
  double s_input;   /* busy time of each stage (the stages overlap) */
  double s_device;
  double s_format;
  double s_write;
  int32_t parallel_fastq; /* 1: every batch went through the parallel 4-line FASTQ parser */
  int32_t reserved;
} dnaln_pipeline_stats;

int dnaln_map_file(struct dnaln_hip_ctx* ctx, const dnaln_batch_config* cfg, const char* reads_path, int out_fd,
                   const char* const* contig_names, int32_t n_contigs, uint64_t batch_bytes, int32_t threads,
                   dnaln_pipeline_stats* stats);
/* The same on two batches the caller created on one context (dnaln_hip_batch_create) and keeps
 * for later calls, which then allocate nothing on the device once warm. */
int dnaln_map_file_on(struct dnaln_hip_batch* b0, struct dnaln_hip_batch* b1, const char* reads_path, int out_fd,
                      const char* const* contig_names, int32_t n_contigs, uint64_t batch_bytes, int32_t threads,
                      dnaln_pipeline_stats* stats);
/* Mapper::map (src/mapper.cpp:75-223, minus SAM text) over n reads already encoded in host memory
 * (the arguments of dnaln_hip_map_batch), through n_batches device batches of one context at
 * once: each batch's thread takes the next chunk of chunk_reads reads (0: 262,144), uploads it
 * from the caller's arrays (page-locked memory, dnaln_hip_host_alloc, makes these copies direct),
 * runs it and downloads its records into out, so one batch's copies overlap another's kernels.
 * CIGAR words are appended to cigar_pool in read order (*cigar_used words); records are those of
 * dnaln_hip_map_batch. stats: reads, mapped, batches (chunks) and s_device = s_wall, the call's
 * time; may be NULL. The batches stay with the caller (warm calls allocate no device memory). */
int dnaln_map_encoded(struct dnaln_hip_batch* const* batches, int32_t n_batches, const uint8_t* codes,
                      const uint64_t* offs, const int32_t* lens, const uint32_t* fnv, uint64_t n, uint64_t chunk_reads,
                      dnaln_record* out, uint32_t* cigar_pool, uint64_t cigar_cap, uint64_t* cigar_used,
                      dnaln_pipeline_stats* stats);
/* Accuracy against simulated truth (no part of the mapping path; validation of its results).
 * dnaln_evaluate is Evaluator::add (src/evaluate.cpp:40-84) over n records, record i against
 * truth i: truth_start[i] is the true 0-based start within contig truth_contig[i] (-1: not in
 * the truth set) and truth_cigar[truth_cigar_off[i] .. [i + 1]) its true CIGAR in reference
 * orientation (truth_cigar NULL: no base concordance). A mapped record is correct on the true
 * contig and strand with its unclipped start within tol bp of the true start. */
typedef struct {
  uint64_t n_truth, records, mapped, correct, bases, bases_ok;
  uint64_t mq_mapped[61], mq_wrong[61]; /* by MAPQ */
} dnaln_eval_stats;
int dnaln_evaluate(const dnaln_record* recs, const uint32_t* cigar_pool, uint64_t n, const int32_t* truth_contig,
                   const int64_t* truth_start, const uint8_t* truth_rev, const uint32_t* truth_cigar,
                   const uint64_t* truth_cigar_off, int32_t tol, dnaln_eval_stats* out);
/* The true-locus checks of src/bench.cpp:272-314 on aligner results (dnaln_hip_align with
 * traceback): task i is read i in reference orientation against the window from its true start
 * minus bw (src/groundtruth.cpp:11-44), truth_cigar its true CIGAR. Counts results whose CIGAR
 * re-scores to their score, tasks whose true path stays in the band and, of those, results
 * scoring at least the true path, and query bases placed on their true target column. */
typedef struct {
  uint64_t n, rescored_ok, in_band, optimal_ok, majority_ok, bases, bases_ok;
} dnaln_truth_stats;
int dnaln_truth_check(const dnaln_score_params* sp, const uint8_t* pool, const dnaln_task* tasks, uint64_t n,
                      const dnaln_result* res, const uint32_t* cigar_pool, const uint32_t* truth_cigar,
                      const uint64_t* truth_cigar_off, dnaln_truth_stats* out);

/* Message of the calling thread's last failed dnaln_map_file, dnaln_map_file_on or dnaln_map_encoded. */
const char* dnaln_host_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
