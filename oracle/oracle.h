/* oracle.h — CPU restatement of dnaln's seed-chain-extend path, used ONLY as a checker.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only to check or time against the HIP path; the product never calls it.
 *
 * Every function follows one piece of the reference (paths relative to the dnaln checkout):
 *   or_encode / or_revcomp       src/common.hpp:23-42
 *   or_hash64                    src/minimizer.hpp:21-30
 *   or_sketch                    src/minimizer.cpp:5-68 (leftmost window minimum, dedupe)
 *   or_index_*                   src/index.hpp:24-38, src/index.cpp:10-45
 *   or_chain                     src/chain.cpp:12-71
 *   or_align                     src/align.cpp:31-84 (scalar kernel), 87-143 (traceback)
 *   or_score_cigar               src/align.cpp:145-162
 *   or_map                       src/mapper.cpp:8-223 (seed, chain, windows, rank, select,
 *                                traceback pass, records, MAPQ)
 *   or_candidates                src/mapper.cpp:14-63 (seed_and_chain alone)
 *   or_format_sam                src/mapper.cpp:231-270, src/common.cpp:6-22
 *   or_fnv1a                     src/common.hpp:136-140
 *
 * Pinning (tests/test_oracle.py): hash64 closed-form bijection; sketch against brute force on
 * tiny inputs; index against a dictionary; SW score against a full-matrix DP; traceback by
 * re-scoring the CIGAR and the perfect-match KAT; chain against the reference KAT and a brute
 * force over all predecessor sets; the whole mapper against SAM records written by the
 * reference binary (tests/golden).
 */
#ifndef DNALN_ORACLE_H
#define DNALN_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int match, mismatch, gap_open, gap_ext, ambig, bw;
} or_score_t;

typedef struct {
  int max_gap, band, lookback, min_score, min_anchors;
} or_chain_params_t;

typedef struct {
  int k, w;
  uint32_t max_occ;
  or_score_t sc;
  or_chain_params_t ch;
  int max_cand;
  double pri_ratio;
  int min_aln_score;
} or_map_params_t;

typedef struct {
  uint64_t hash;
  uint32_t pos, strand;
} or_mini_t;

typedef struct {
  uint64_t key;
  uint32_t offset, count;
} or_bucket_t;

typedef struct {
  or_bucket_t* table;
  uint64_t mask;
  uint32_t* pos;
  uint64_t n_pos, n_keys, cap;
  uint32_t max_count;
  int k, w;
} or_index_t;

typedef struct {
  uint32_t rpos;
  int32_t qpos;
  uint32_t strand;
} or_anchor_t;

typedef struct {
  int32_t score;
  uint32_t strand;
  int32_t n_anchors;
  int64_t dmin, dmax;
  uint32_t rpos;
} or_chain_t;

typedef struct {
  int32_t score, qb, qe, tb, te, nm;
  int32_t n_cigar; /* ops written to the caller's buffer, -1 if it was too small */
} or_result_t;

typedef struct {
  int32_t contig; /* -1: unmapped */
  int32_t pad_;
  int64_t pos;
  uint8_t rev, mapq, pad2_[2];
  int32_t score, nm, n_cand;
  int32_t n_cigar; /* including soft clips */
  uint64_t cigar_off;
} or_record_t;

void or_default_params(or_map_params_t* p);

void or_encode(const char* s, size_t n, uint8_t* out);
void or_revcomp(const uint8_t* in, size_t n, uint8_t* out);
uint64_t or_hash64(uint64_t key, uint64_t mask);
uint32_t or_fnv1a(const char* s, size_t n);

/* Writes at most len minimizers; returns the count. */
size_t or_sketch(const uint8_t* seq, uint32_t len, int k, int w, uint32_t offset, or_mini_t* out);

int or_index_build(or_index_t* idx, const uint8_t* seq, const uint64_t* contig_off, const uint32_t* contig_len,
                   int n_contigs, int k, int w);
const uint32_t* or_index_lookup(const or_index_t* idx, uint64_t h, uint32_t* n);
void or_index_free(or_index_t* idx);

/* Sorts a[] in place; writes at most n chains to out; returns the count. */
size_t or_chain(or_anchor_t* a, size_t n, int k, const or_chain_params_t* p, or_chain_t* out);

/* One banded alignment. cigar receives at most cigar_cap ops (M/I/D only). */
int or_align(const or_score_t* sp, const uint8_t* q, int qlen, const uint8_t* t, int tlen, int traceback,
             or_result_t* r, uint32_t* cigar, int cigar_cap);
int or_score_cigar(const or_score_t* sp, const uint8_t* q, const uint8_t* t, int qb, int tb, const uint32_t* cigar,
                   int n_cigar);

/* Maps n reads (ASCII). names: concatenated, name_off[i]/name_len[i]. Records go to out[n];
 * CIGARs to cigar_pool (capacity cap ops). Returns 0, or -1 if cigar_pool is too small. */
int or_map(const or_map_params_t* p, const uint8_t* ref, const uint64_t* contig_off, const uint32_t* contig_len,
           int n_contigs, const or_index_t* idx, const char* seq, const uint64_t* seq_off, const int32_t* seq_len,
           const char* names, const uint64_t* name_off, const int32_t* name_len, size_t n, or_record_t* out,
           uint32_t* cigar_pool, size_t cap, uint64_t* cells);

/* Candidates of n reads (ASCII) after seeding and chaining: n_cand[r], and 4 values per candidate
 * (rev, contig, diag, chain score) at cand[4 * (r * max_cand + c)]. */
int or_candidates(const or_map_params_t* p, const uint64_t* contig_off, int n_contigs, const or_index_t* idx,
                  const char* seq, const uint64_t* seq_off, const int32_t* seq_len, size_t n, int32_t* n_cand,
                  int64_t* cand);

/* Appends one SAM line (with '\n') to buf; returns the length written, or -1 if cap is short. */
long or_format_sam(const char* name, int name_len, const char* seq, const char* qual, int len,
                   const or_record_t* m, const uint32_t* cigar_pool, const char* contig_name, char* buf, size_t cap);

#ifdef __cplusplus
}
#endif

#endif
