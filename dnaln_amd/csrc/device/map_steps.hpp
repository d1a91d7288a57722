// map_steps.hpp — per-read steps of Mapper::map after chaining (src/mapper.cpp:103-222).
#pragma once

#include "common.hpp"
#include "seed_chain.hpp"
#include "dnaln_hip.h"

namespace dn {

struct MapGeom {
  uint64_t ref;            // device address of the reference codes
  const uint64_t* coff;    // contig offsets
  const uint32_t* clen;
  int bw;
};

// The window [wstart, wstart + qlen + 2bw) of candidate c reads N outside its contig
// (src/mapper.cpp:103-121, src/reference.cpp:40-47).
DN_HD SwTask make_task(const MapGeom& g, uint64_t q, int qlen, const Cand& c) {
  SwTask t;
  t.q = q;
  t.qlen = qlen;
  t.flags = c.rev;
  const int64_t ws = cand_wstart(c, g.bw);
  t.t = g.ref + (uint64_t)ws;
  const int64_t wl = (int64_t)qlen + 2 * (int64_t)g.bw;
  const int64_t cb = (int64_t)g.coff[c.contig], ce = cb + (int64_t)g.clen[c.contig];
  int64_t lo = cb - ws, hi = ce - ws;
  lo = lo < 0 ? 0 : (lo > wl ? wl : lo);
  hi = hi < lo ? lo : (hi > wl ? wl : hi);
  t.tlo = (int32_t)lo;
  t.thi = (int32_t)hi;
  return t;
}

// Rank and select (src/mapper.cpp:142-174): s1 = best score, ties = its count, s2 = s1 when
// tied else the best other score (0 if none); the pick among ties is fnv1a(name) % ties.
template <class Score>
DN_HD void select_read(int nc, Score sw, uint32_t fnv, int32_t& chosen, int32_t& second) {
  chosen = -1;
  second = 0;
  if (nc == 1) chosen = 0;
  if (nc < 2) return;
  int s1 = -1, ties = 0;
  for (int c = 0; c < nc; ++c) {
    if (sw(c) > s1) {
      s1 = sw(c);
      ties = 1;
    } else if (sw(c) == s1) {
      ++ties;
    }
  }
  int s2 = 0;
  if (ties > 1) s2 = s1;
  else
    for (int c = 0; c < nc; ++c)
      if (sw(c) != s1 && sw(c) > s2) s2 = sw(c);
  uint32_t pick = fnv % (uint32_t)ties;
  for (int c = 0; c < nc; ++c) {
    if (sw(c) == s1 && pick-- == 0) {
      chosen = c;
      break;
    }
  }
  second = s2;
}

// Record and MAPQ (src/mapper.cpp:189-222). mapq_lut[d] = min(60, 6.02 d / match + 0.499) as
// the reference computes it in double, for d < lut_len; larger d give 60.
DN_HD void record_read(const MapGeom& g, int qlen, int nc, const Cand* c, int chosen, int second,
                       const SwResult* res, int min_aln_score, const uint8_t* mapq_lut, int lut_len,
                       dnaln_record& m) {
  m.contig = -1;
  m.rev = 0;
  m.mapq = 0;
  m.reserved = 0;
  m.pos = 0;
  m.score = 0;
  m.nm = 0;
  m.n_cand = nc;
  m.cigar_len = 0;
  m.cigar_off = 0;
  if (chosen < 0 || res->score < min_aln_score) return;
  const Cand& cd = c[chosen];
  m.contig = cd.contig;
  m.rev = (uint8_t)cd.rev;
  m.pos = cand_wstart(cd, g.bw) + res->tb - (int64_t)g.coff[cd.contig];
  m.score = res->score;
  m.nm = res->nm;
  m.cigar_len = (uint32_t)res->n_cigar + (res->qb > 0) + (qlen > res->qe);
  if (nc <= 1) {
    m.mapq = 60;
  } else {
    const int delta = res->score - second;
    m.mapq = delta <= 0 ? 0 : (delta < lut_len ? mapq_lut[delta] : 60);
  }
}

// CIGAR with soft clips; the walk stored its ops in reverse.
DN_HD void gather_cigar(int qlen, const SwResult& r, const uint32_t* rev_ops, uint32_t* out) {
  int o = 0;
  if (r.qb > 0) out[o++] = (uint32_t)r.qb << 4 | 4;
  for (int c = r.n_cigar - 1; c >= 0; --c) out[o++] = rev_ops[c];
  if (qlen > r.qe) out[o++] = (uint32_t)(qlen - r.qe) << 4 | 4;
}

}  // namespace dn
