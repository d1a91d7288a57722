// host_api.cpp — index build, encoding and SAM formatting for the HIP path (include/dnaln_host.h).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "device/seed_chain.hpp"
#include "dnaln_host.h"

namespace {

uint8_t nt4(unsigned char c) {  // src/common.hpp:23-31
  switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': case 'U': case 'u': return 3;
    default: return 4;
  }
}

char comp_char(char c) {  // src/common.cpp:6-14
  switch (c) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
    case 'a': return 't';
    case 'c': return 'g';
    case 'g': return 'c';
    case 't': return 'a';
    case 'n': return 'n';
    default: return 'N';
  }
}

struct Occ {
  uint64_t hash;
  uint32_t pos;  // pos << 1 | strand
};

}  // namespace

extern "C" {

int dnaln_index_build(const uint8_t* ref, const dnaln_contig* contigs, int32_t n_contigs, int32_t k, int32_t w,
                      dnaln_index* out) {
  if (!out || !contigs || n_contigs < 1 || k < 4 || k > 28 || w < 1 || w > 255) return DNALN_E_INVALID;
  std::memset(out, 0, sizeof *out);
  uint64_t total = 0;
  for (int c = 0; c < n_contigs; ++c) total = std::max<uint64_t>(total, contigs[c].offset + contigs[c].len);
  if (total >= (1ull << 31)) return DNALN_E_INVALID;
  std::vector<Occ> occ;
  occ.reserve(total * 2 / (w + 1) + 1024);
  std::vector<uint64_t> rh(w);
  std::vector<uint32_t> rps(w);
  dn::Ring ring{rh.data(), rps.data(), 1};
  for (int c = 0; c < n_contigs; ++c) {
    const uint32_t base = (uint32_t)contigs[c].offset;
    dn::sketch_stream(dn::PtrReader{ref + contigs[c].offset}, (int)contigs[c].len, k, w, ring,
                      [&](uint64_t h, uint32_t p, uint32_t s) { occ.push_back({h, (p + base) << 1 | s}); });
  }
  std::sort(occ.begin(), occ.end(),
            [](const Occ& a, const Occ& b) { return a.hash != b.hash ? a.hash < b.hash : a.pos < b.pos; });
  uint64_t n_keys = 0;
  for (size_t i = 0; i < occ.size(); ++i) n_keys += i == 0 || occ[i].hash != occ[i - 1].hash;
  uint64_t cap = 16;
  while (cap < n_keys * 2) cap <<= 1;
  dnaln_bucket* table = new dnaln_bucket[cap];
  for (uint64_t s = 0; s < cap; ++s) table[s] = {~0ULL, 0, 0};
  uint32_t* pos = new uint32_t[occ.size() + 1];
  const uint64_t mask = cap - 1;
  uint32_t max_count = 0;
  for (size_t i = 0; i < occ.size();) {
    size_t j = i;
    for (; j < occ.size() && occ[j].hash == occ[i].hash; ++j) pos[j] = occ[j].pos;
    uint64_t s = occ[i].hash & mask;
    while (table[s].key != ~0ULL) s = (s + 1) & mask;
    table[s] = {occ[i].hash, (uint32_t)i, (uint32_t)(j - i)};
    max_count = std::max(max_count, (uint32_t)(j - i));
    i = j;
  }
  out->view.table = table;
  out->view.n_buckets = cap;
  out->view.pos = pos;
  out->view.n_pos = occ.size();
  out->view.k = k;
  out->view.w = w;
  out->n_keys = n_keys;
  out->max_count = max_count;
  return DNALN_OK;
}

void dnaln_index_free(dnaln_index* idx) {
  if (!idx) return;
  delete[] idx->view.table;
  delete[] idx->view.pos;
  std::memset(idx, 0, sizeof *idx);
}

uint32_t dnaln_fnv1a(const char* s, size_t n) {
  uint32_t h = 2166136261u;
  for (size_t i = 0; i < n; ++i) h = (h ^ (unsigned char)s[i]) * 16777619u;
  return h;
}

void dnaln_encode(const char* s, size_t n, uint8_t* out) {
  for (size_t i = 0; i < n; ++i) out[i] = nt4((unsigned char)s[i]);
}

long dnaln_sam_format(const char* name, size_t name_len, const char* seq, const char* qual, size_t len,
                      const dnaln_record* m, const uint32_t* cigar_pool, const char* contig_name, char* buf,
                      size_t cap) {
  static const char kOps[] = "MIDNSHP=X";
  const size_t need = name_len + 2 * len + 96 + (size_t)m->cigar_len * 12 + (contig_name ? std::strlen(contig_name) : 0);
  if (need > cap) return -1;
  char* o = buf;
  std::memcpy(o, name, name_len);
  o += name_len;
  if (m->contig < 0) {
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
    return (long)(o - buf);
  }
  o += std::sprintf(o, "\t%d\t%s\t%lld\t%d\t", m->rev ? 16 : 0, contig_name, (long long)(m->pos + 1), (int)m->mapq);
  for (uint32_t c = 0; c < m->cigar_len; ++c) {
    const uint32_t x = cigar_pool[m->cigar_off + c];
    o += std::sprintf(o, "%u%c", x >> 4, kOps[x & 0xf]);
  }
  std::memcpy(o, "\t*\t0\t0\t", 7);
  o += 7;
  if (m->rev) {
    for (size_t i = 0; i < len; ++i) *o++ = comp_char(seq[len - 1 - i]);
    *o++ = '\t';
    if (len == 0 || !qual) *o++ = '*';
    else
      for (size_t i = 0; i < len; ++i) *o++ = qual[len - 1 - i];
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
  o += std::sprintf(o, "\tNM:i:%d\tAS:i:%d\n", m->nm, m->score);
  return (long)(o - buf);
}

}  // extern "C"

extern "C" {

// FNV-1a of n names packed back to back (names + offs[i], lens[i]).
void dnaln_fnv1a_many(const char* names, const uint64_t* offs, const int32_t* lens, uint64_t n, uint32_t* out) {
  for (uint64_t i = 0; i < n; ++i) out[i] = dnaln_fnv1a(names + offs[i], (size_t)lens[i]);
}

}  // extern "C"
