// sam_text.hpp — one SAM record as text (src/mapper.cpp:231-270), written with table lookups and
// hand-rolled integer formatting so that several threads can format a batch at memory speed.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <cstring>

#include "dnaln_hip.h"

namespace dnaln_amd {

// kCompChar of src/common.cpp:6-14: ACGT and acgt complemented, n kept, anything else N.
struct CompTable {
  char c[256];
  CompTable() {
    for (int i = 0; i < 256; ++i) c[i] = 'N';
    c[(unsigned char)'A'] = 'T';
    c[(unsigned char)'C'] = 'G';
    c[(unsigned char)'G'] = 'C';
    c[(unsigned char)'T'] = 'A';
    c[(unsigned char)'a'] = 't';
    c[(unsigned char)'c'] = 'g';
    c[(unsigned char)'g'] = 'c';
    c[(unsigned char)'t'] = 'a';
    c[(unsigned char)'n'] = 'n';
  }
};
inline const CompTable& comp_table() {
  static const CompTable t;
  return t;
}

inline char* put_u64(char* o, uint64_t v) {
  char tmp[20];
  int n = 0;
  do {
    tmp[n++] = (char)('0' + v % 10);
    v /= 10;
  } while (v);
  while (n) *o++ = tmp[--n];
  return o;
}
inline char* put_i64(char* o, int64_t v) {
  if (v < 0) {
    *o++ = '-';
    return put_u64(o, (uint64_t)(-(v + 1)) + 1);
  }
  return put_u64(o, (uint64_t)v);
}

// Bytes one record can take at most (the bound dnaln_sam_format checks against).
inline size_t sam_bound(size_t name_len, size_t len, uint32_t cigar_len, size_t contig_len) {
  return name_len + 2 * len + 96 + (size_t)cigar_len * 12 + contig_len;
}

// Writes the record at o (room for sam_bound bytes) and returns the end. qual may be null (FASTA).
inline char* sam_record(char* o, const char* name, size_t name_len, const char* seq, const char* qual, size_t len,
                        const dnaln_record& m, const uint32_t* cigar_pool, const char* contig, size_t contig_len) {
  static const char kOps[] = "MIDNSHP=X";
  std::memcpy(o, name, name_len);
  o += name_len;
  if (m.contig < 0) {
    static const char kUnmapped[] = "\t4\t*\t0\t0\t*\t*\t0\t0\t";
    std::memcpy(o, kUnmapped, sizeof kUnmapped - 1);
    o += sizeof kUnmapped - 1;
    if (len == 0) *o++ = '*';
    else {
      std::memcpy(o, seq, len);
      o += len;
    }
    *o++ = '\t';
    if (len == 0 || !qual) *o++ = '*';
    else {
      std::memcpy(o, qual, len);
      o += len;
    }
    *o++ = '\n';
    return o;
  }
  *o++ = '\t';
  if (m.rev) *o++ = '1', *o++ = '6';
  else *o++ = '0';
  *o++ = '\t';
  std::memcpy(o, contig, contig_len);
  o += contig_len;
  *o++ = '\t';
  o = put_i64(o, m.pos + 1);
  *o++ = '\t';
  o = put_u64(o, m.mapq);
  *o++ = '\t';
  for (uint32_t c = 0; c < m.cigar_len; ++c) {
    const uint32_t x = cigar_pool[m.cigar_off + c];
    o = put_u64(o, x >> 4);
    *o++ = kOps[x & 0xf];
  }
  std::memcpy(o, "\t*\t0\t0\t", 7);
  o += 7;
  if (m.rev) {
    const char* t = comp_table().c;
    for (size_t i = 0; i < len; ++i) o[i] = t[(unsigned char)seq[len - 1 - i]];
    o += len;
    *o++ = '\t';
    if (len == 0 || !qual) *o++ = '*';
    else {
      for (size_t i = 0; i < len; ++i) o[i] = qual[len - 1 - i];
      o += len;
    }
  } else {
    std::memcpy(o, seq, len);
    o += len;
    *o++ = '\t';
    if (len == 0 || !qual) *o++ = '*';
    else {
      std::memcpy(o, qual, len);
      o += len;
    }
  }
  std::memcpy(o, "\tNM:i:", 6);
  o = put_i64(o + 6, m.nm);
  std::memcpy(o, "\tAS:i:", 6);
  o = put_i64(o + 6, m.score);
  *o++ = '\n';
  return o;
}

}  // namespace dnaln_amd
