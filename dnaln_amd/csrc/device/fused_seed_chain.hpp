// fused_seed_chain.hpp — one read per thread from bases to candidates, with the read's anchors in
// LDS (the batched form's seed + chain stage).
//
// Same results as seed_count/seed_emit/chain_read (src/minimizer.cpp:5-68, src/mapper.cpp:14-63,
// src/chain.cpp:12-71), reorganised so that nothing but the candidates leaves the CU:
//   * anchors go straight from the probes into LDS (forward-strand ones from the front,
//     reverse-strand ones from the back, so both runs arrive nearly sorted);
//   * f (int16) and prev (int8) live beside them; anchor j of the thread sits at element
//     j * stride, with the 64 threads of a wave interleaved, so LDS accesses are conflict-free;
//   * chain extraction replays the reference order (f desc, index asc) by repeated argmax over a
//     64-bit used mask, first to find the best chain score, then to collect the chains scoring
//     >= pri_ratio * best into a register list kept in stable score order;
//   * candidates are selected from that list exactly as src/mapper.cpp:44-62 does.
// A read whose anchors exceed the LDS capacity, or whose candidate selection would need more
// chains than the register list holds, returns -1; the caller routes it to the global-memory
// path, which has no such limits.
#pragma once

#include "seed_chain.hpp"

namespace dn {

struct LdsView {
  uint64_t* keys;
  int16_t* f;
  int8_t* prev;
  int stride;
};

struct ChainLite {
  int32_t score;
  int32_t dmin, dmax;        // diagonals rpos - qpos fit 32 bits (rpos < 2^31, qpos >= 0)
  uint32_t rpos_strand;      // rpos | strand << 31
};

// Anchor index with the largest f among those not in `used` (ties: smallest index), or -1.
DN_HD int next_end(const LdsView& v, int n, uint64_t used) {
  int best = -1, bf = -1;
  for (int i = 0; i < n; ++i) {
    if ((used >> i) & 1) continue;
    const int fi = v.f[i * v.stride];
    if (fi > bf) {
      bf = fi;
      best = i;
    }
  }
  return best;
}

// Walks the chain ending at `end` (src/chain.cpp:50-61); marks its anchors used.
DN_HD ChainLite walk_chain(const LdsView& v, int end, uint64_t& used, int& cnt) {
  int i = end;
  cnt = 0;
  int64_t dmin = INT64_MAX, dmax = INT64_MIN;
  while (i >= 0 && !((used >> i) & 1)) {
    used |= 1ull << i;
    ++cnt;
    const uint64_t k = v.keys[i * v.stride];
    const int64_t d = (int64_t)key_rpos(k) - key_qpos(k);
    dmin = d < dmin ? d : dmin;
    dmax = d > dmax ? d : dmax;
    i = v.prev[i * v.stride];
  }
  ChainLite c;
  c.score = v.f[end * v.stride] - (i >= 0 ? v.f[i * v.stride] : 0);
  c.dmin = (int32_t)dmin;
  c.dmax = (int32_t)dmax;
  const uint64_t ke = v.keys[end * v.stride];
  c.rpos_strand = key_rpos(ke) | key_strand(ke) << 31;
  return c;
}

template <int MAXCH>
DN_HD int seed_chain_lds(const SeedArgs& s, const ChainArgs& a, uint32_t r, Ring ring, LdsView v, int cap) {
  // 1. anchors
  const int len = s.lens[r], k = s.k;
  int lo = 0, hi = cap - 1;
  bool overflow = false;
  auto on = [&](uint32_t mpos, uint32_t mstrand, uint32_t off, uint32_t cnt, uint32_t first) {
    if (overflow) return;
    if ((int)cnt > hi - lo + 1) {
      overflow = true;
      return;
    }
    for (uint32_t j = 0; j < cnt; ++j) {
      const uint32_t val = j == 0 ? first : s.pos[off + j];
      const uint32_t st = (val & 1) ^ mstrand;
      const int32_t qpos = st ? len - (int32_t)mpos - k : (int32_t)mpos;
      const uint64_t key = anchor_key(st, val >> 1, qpos);
      if (st) v.keys[(hi--) * v.stride] = key;
      else v.keys[(lo++) * v.stride] = key;
    }
  };
  ProbePipe<true, decltype(on)> pipe(s, on);
  sketch_stream(ChunkReader(s.codes + s.offs[r]), len, k, s.w, ring,
                [&](uint64_t h, uint32_t p, uint32_t st) { pipe.push(h, p, st); });
  pipe.drain();
  if (overflow) return -1;
  const int nrev = cap - 1 - hi, n = lo + nrev;
  for (int x = 0; x < nrev; ++x) v.keys[(lo + x) * v.stride] = v.keys[(hi + 1 + x) * v.stride];
  if (n == 0) return 0;

  // 2. sort (nearly sorted: insertion sort) and DP (src/chain.cpp:21-45)
  for (int i = 1; i < n; ++i) {
    const uint64_t x = v.keys[i * v.stride];
    int j = i - 1;
    while (j >= 0 && x < v.keys[j * v.stride]) {
      v.keys[(j + 1) * v.stride] = v.keys[j * v.stride];
      --j;
    }
    v.keys[(j + 1) * v.stride] = x;
  }
  for (int i = 0; i < n; ++i) {
    const uint64_t ki = v.keys[i * v.stride];
    const uint32_t si = key_strand(ki);
    const int64_t ri = key_rpos(ki), qi = key_qpos(ki);
    int best = k, best_j = -1;
    const int jlo = i - a.lookback > 0 ? i - a.lookback : 0;
    for (int j = i - 1; j >= jlo; --j) {
      const uint64_t kj = v.keys[j * v.stride];
      if (key_strand(kj) != si) break;
      const int64_t dr = ri - key_rpos(kj);
      if (dr > a.max_gap) break;
      const int64_t dq = qi - key_qpos(kj);
      if (dr <= 0 || dq <= 0) continue;
      const int64_t dd = dr > dq ? dr - dq : dq - dr;
      if (dd > a.band) continue;
      const int64_t m = dq < dr ? dq : dr;
      const int sc = v.f[j * v.stride] + (int)(m < k ? m : k) - a.cost[dd];
      if (sc > best) {
        best = sc;
        best_j = j;
      }
    }
    v.f[i * v.stride] = (int16_t)best;
    v.prev[i * v.stride] = (int8_t)best_j;
  }

  // 3. extraction, pass 1: best qualifying chain score
  uint64_t used = 0;
  int best_score = -1;
  for (int end = next_end(v, n, used); end >= 0; end = next_end(v, n, used)) {
    int cnt;
    const ChainLite c = walk_chain(v, end, used, cnt);
    if (c.score >= a.min_score && cnt >= a.min_anchors && c.score > best_score) best_score = c.score;
  }
  if (best_score < 0) return 0;
  // pass 2: chains scoring >= pri_ratio * best, in stable score order, at most MAXCH kept
  const double threshold = best_score * a.pri_ratio;
  ChainLite top[MAXCH];
  int count = 0;
  bool truncated = false;
  used = 0;
  for (int end = next_end(v, n, used); end >= 0; end = next_end(v, n, used)) {
    int cnt;
    ChainLite cur = walk_chain(v, end, used, cnt);
    if (cur.score < a.min_score || cnt < a.min_anchors || cur.score < threshold) continue;
    bool inserted = false;
#pragma unroll
    for (int x = 0; x < MAXCH; ++x) {
      const bool take = inserted || x >= count || top[x].score < cur.score;
      if (take) {
        const ChainLite t = top[x];
        top[x] = cur;
        cur = t;
        inserted = true;
      }
    }
    if (count == MAXCH || !inserted) truncated = true;
    count = count < MAXCH ? count + 1 : MAXCH;
  }

  // 4. candidates (src/mapper.cpp:44-62)
  Cand* cands = a.cands + (uint64_t)r * (uint64_t)a.max_cand;
  const int64_t bw = a.bw;
  int nc = 0;
  uint32_t chosen = 0;  // bit x: top[x] became a candidate
  bool done = false;
#pragma unroll
  for (int x = 0; x < MAXCH; ++x) {
    if (done || x >= count) continue;
    if (nc >= a.max_cand) {
      done = true;
      continue;
    }
    const uint32_t strand = top[x].rpos_strand >> 31;
    const int64_t diag = (int64_t)top[x].dmin + ((int64_t)top[x].dmax - top[x].dmin) / 2;
    bool dup = false;
#pragma unroll
    for (int y = 0; y < x; ++y) {
      if (!((chosen >> y) & 1)) continue;
      const int64_t dy = (int64_t)top[y].dmin + ((int64_t)top[y].dmax - top[y].dmin) / 2;
      const int64_t dd = dy - diag;
      dup = dup || ((top[y].rpos_strand >> 31) == strand && (dd < 0 ? -dd : dd) <= bw);
    }
    if (dup) continue;
    Cand cd;
    cd.rev = strand;
    cd.contig = contig_of(a.contig_off, a.n_contigs, top[x].rpos_strand & 0x7fffffffu);
    cd.diag = diag;
    cd.wstart = diag - bw;
    cands[nc++] = cd;
    chosen |= 1u << x;
  }
  // The list ran out before max_cand candidates while chains were dropped from it.
  if (truncated && !done && nc < a.max_cand) return -1;
  return nc;
}

}  // namespace dn
