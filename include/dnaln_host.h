/* dnaln_host.h — host-side services around the HIP path (libdnaln_host.so).
 *
 *   dnaln_index_build   MinimizerIndex::build (src/index.cpp:10-45): sketch every contig, sort
 *                       occurrences by (hash, pos), place keys with linear probing at load <= 0.5.
 *                       Produces the view dnaln_hip_create uploads.
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

uint32_t dnaln_fnv1a(const char* s, size_t n);
void dnaln_fnv1a_many(const char* names, const uint64_t* offs, const int32_t* lens, uint64_t n, uint32_t* out);
void dnaln_encode(const char* s, size_t n, uint8_t* out);

/* Appends one SAM line to buf (capacity cap); returns its length, or -1 when cap is short.
 * qual may be NULL (FASTA input). */
long dnaln_sam_format(const char* name, size_t name_len, const char* seq, const char* qual, size_t len,
                      const dnaln_record* rec, const uint32_t* cigar_pool, const char* contig_name, char* buf,
                      size_t cap);

#ifdef __cplusplus
}
#endif

#endif
