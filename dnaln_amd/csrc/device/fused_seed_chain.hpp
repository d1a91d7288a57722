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
//     64-bit used mask, in one pass that keeps the qualifying chains in a register list in
//     stable score order (at most MAXCH) and remembers the best score that fell off it;
//   * candidates are selected from that list exactly as src/mapper.cpp:44-62 does.
// A read whose anchors exceed the LDS capacity, or whose candidate selection would need a chain
// that fell off the register list, returns -1; the caller routes it to the global-memory
// path, which has no such limits.
#pragma once

#include "seed_chain.hpp"

namespace dn {

// The read's anchors in LDS, one u64 word each: anchor j of the thread at keys[j * stride], the
// 64 threads of a wave interleaved. A word is the anchor key (strand << 63 | rpos << 32 | qpos)
// with the chain DP's results packed into the low dword above a 12-bit qpos (reads < 4096 bp):
// qpos bits 0..11; bits 12..13 the DP's headroom dm = (max f over the strand's anchors up to
// this one) - f, 3 meaning 3 or more; f bits 14..24 (1 <= f <= cap * k <= 64 * 28 < 2^11);
// prev + 1 bits 25..31 (prev in -1..63). The DP fields are 0 until the DP writes them, so
// sorting the words sorts the keys.
struct LdsView {
  uint64_t* keys;
  int stride;
  DN_HD uint64_t word(int j) const { return keys[j * stride]; }
  // Low dword of anchor j (qpos, f, prev).
  DN_HD void set_low(int j, uint32_t lo) const {
#if defined(__HIP_DEVICE_COMPILE__)
    reinterpret_cast<uint32_t*>(keys + j * stride)[0] = lo;
#else
    keys[j * stride] = (keys[j * stride] & ~0xffffffffull) | lo;
#endif
  }
};
constexpr int kFusedMaxLen = 4096;  // reads the fused path takes (12-bit qpos)
DN_HD int w_qpos(uint64_t x) { return (int)(x & 0xfffu); }
DN_HD int w_dm(uint64_t x) { return (int)((x >> 12) & 3u); }
DN_HD int w_f(uint64_t x) { return (int)((x >> 14) & 0x7ffu); }
DN_HD int w_prev(uint64_t x) { return (int)((x >> 25) & 0x7fu) - 1; }
DN_HD uint32_t w_low(int qpos, int dm, int f, int prev) {
  return (uint32_t)qpos | (uint32_t)dm << 12 | (uint32_t)f << 14 | (uint32_t)(prev + 1) << 25;
}

struct ChainLite {
  int32_t score;
  int32_t dmin, dmax;        // diagonals rpos - qpos fit 32 bits (rpos < 2^31, qpos >= 0)
  uint32_t rpos_strand;      // rpos | strand << 31
};

// Anchor index with the largest f among those not in `used` (ties: smallest index), or -1;
// *f_out = its f.
DN_HD int next_end(const LdsView& v, int n, uint64_t used, int* f_out) {
  int best = -1, bf = -1;
  for (int i = 0; i < n; ++i) {
    if ((used >> i) & 1) continue;
    const int fi = w_f(v.word(i));
    if (fi > bf) {
      bf = fi;
      best = i;
    }
  }
  *f_out = bf;
  return best;
}

// Walks the chain ending at `end` (src/chain.cpp:50-61); marks its anchors used.
DN_HD ChainLite walk_chain(const LdsView& v, int end, uint64_t& used, int& cnt) {
  int i = end;
  cnt = 0;
  int32_t dmin = INT32_MAX, dmax = INT32_MIN;
  int f_stop = 0;
  while (i >= 0 && !((used >> i) & 1)) {
    used |= 1ull << i;
    ++cnt;
    const uint64_t x = v.word(i);
    const int32_t d = (int32_t)key_rpos(x) - w_qpos(x);
    dmin = d < dmin ? d : dmin;
    dmax = d > dmax ? d : dmax;
    i = w_prev(x);
  }
  if (i >= 0) f_stop = w_f(v.word(i));
  const uint64_t ke = v.word(end);
  ChainLite c;
  c.score = w_f(ke) - f_stop;
  c.dmin = dmin;
  c.dmax = dmax;
  c.rpos_strand = key_rpos(ke) | key_strand(ke) << 31;
  return c;
}

// The bases of a read staged in LDS: the aligned 16-byte chunks around the read, stored as u64
// words in the thread's (interleaved) anchor slots; byte j of the read sits at byte off + j.
struct LdsSeq {
  const uint64_t* words;
  int stride;
  uint32_t off;
  DN_HD uint8_t operator()(int j) const {
    const uint32_t x = off + (uint32_t)j;
    return (uint8_t)(words[(x >> 3) * stride] >> (8 * (x & 7)));
  }
};

// Loads the read's chunks into the u64 slots of `dst` (stride apart), sixteen chunks at a time
// with all loads of a batch issued before its first wait, and checks the read for N (codes
// above 3) on the way. True when the read is staged and holds no N; false when it holds an N or
// the slots cannot hold it plus `tail` bytes that a reader may look at past its end. Reads up
// to 15 bytes before and after the read: the read buffer carries guard bytes.
DN_HD bool stage_read(const uint8_t* seq, int len, int tail, uint64_t* dst, int stride, int slots, LdsSeq& out) {
  constexpr int kBatch = 16;
  const uintptr_t a = (uintptr_t)seq;
  const uint32_t off = (uint32_t)(a & 15);
  const int nch = (int)((off + (uint32_t)len + 15) >> 4);
  if ((int)((off + (uint32_t)(len + tail) + 7) >> 3) > slots || 2 * nch > slots) return false;
  const uint64_t src = (uint64_t)(a - off);
  uint32_t bad = 0;
  for (int b0 = 0; b0 < nch; b0 += kBatch) {
    Quad c[kBatch];
#pragma unroll
    for (int x = 0; x < kBatch; ++x)
      if (b0 + x < nch) c[x] = load16(src + 16 * (uint64_t)(b0 + x));
#pragma unroll
    for (int x = 0; x < kBatch; ++x)
      if (b0 + x < nch) {
        dst[(2 * (b0 + x)) * stride] = (uint64_t)c[x].y << 32 | c[x].x;
        dst[(2 * (b0 + x) + 1) * stride] = (uint64_t)c[x].w << 32 | c[x].z;
        // Bytes of the read in each dword (the first and last chunks hold bytes of neighbours).
        const int q = 16 * (b0 + x) - (int)off;  // read position of the chunk's byte 0
        const uint32_t dw[4] = {c[x].x, c[x].y, c[x].z, c[x].w};
#pragma unroll
        for (int d = 0; d < 4; ++d) bad |= dw[d] & bytes_from(-(q + 4 * d)) & bytes_below(len - (q + 4 * d));
      }
  }
  out.words = dst;
  out.stride = stride;
  out.off = off;
  return (bad & 0xFCFCFCFCu) == 0;
}

// Chain DP over sorted anchors in LDS (src/chain.cpp:21-45). The anchors are sorted by strand
// first, so the predecessors that the reference reaches before its strand break are the same-
// strand ones: j >= the strand's first anchor (nf, the forward-strand count, for the reverse
// strand). Within a strand rpos is sorted, so dr = ri - rj is the difference of the words' high
// halves (the strand bits cancel) and never negative, and the reference's stop at dr > max_gap
// is monotone in j. Predecessors are visited nearest first in groups of four whose loads are
// issued together; each is tested on its own, without branches, and the scan ends after the
// first group whose farthest predecessor is stopped. It also ends once no predecessor left can
// beat the best: a predecessor's score is at most f[j] + k (gain <= k, cost >= 0), and the
// farthest anchor of a group carries dm, so max f over the anchors before it is known; along a
// colinear chain f grows, so the scan usually ends after one group. The result is the
// reference's either way. NIB: gap costs from the four-bit table (drift < 16) and cost_at_band
// (drift == band), valid when cost_nib_ok; a drift above the band is never taken, so its cost
// does not matter. Arithmetic is 32-bit: rpos < 2^31, qpos < 2^12.
// Sorts the n anchor words ascending (n <= N): the words go to registers (+inf past n), a bitonic
// network of N keys with compile-time indices sorts them, and they go back. Every comparator puts
// the smaller word at the lower index (each merge starts with a flip stage, partner i ^ (k - 1)),
// so the +inf padding stays at the end. Anchors from several loci arrive interleaved, so an
// insertion sort's dependent LDS loads grew with the square of their number; the network costs
// N loads, N log2(N) (log2(N) + 1) / 4 compare-exchanges in VGPRs and n stores, whatever the
// order.
template <int N>
DN_HD void sort_lds(const LdsView& v, int n) {
  uint64_t x[N];
#pragma unroll
  for (int i = 0; i < N; ++i) x[i] = i < n ? v.word(i) : ~0ull;
#pragma unroll
  for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int l = j == (k >> 1) ? (i ^ (k - 1)) : (i ^ j);
        if (l > i) {
          const uint64_t a = x[i], b = x[l];
          const bool sw = b < a;
          x[i] = sw ? b : a;
          x[l] = sw ? a : b;
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i)
    if (i < n) v.keys[i * v.stride] = x[i];
}

template <bool NIB>
DN_HD void chain_dp_lds(const ChainArgs& a, const LdsView& v, int n, int nf, int k) {
  const uint32_t max_gap = (uint32_t)a.max_gap;
  const int band = a.band;
  int fmax = 0;  // max f over this strand's anchors so far
  // Anchors i-1 .. i-8 with their DP fields, in registers: the first two groups of predecessors
  // need no LDS load (a wave steps until its last lane's scan ends, and the lane that needs a
  // second group is common), and anchor i's word is loaded one step ahead.
  uint64_t win[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint64_t k_next = v.word(0);
  for (int i = 0; i < n; ++i) {
    const uint64_t ki = k_next;
    if (i + 1 < n) k_next = v.word(i + 1);
    const uint32_t hi = (uint32_t)(ki >> 32);
    const int qi = w_qpos(ki);
    int best = k, best_j = -1;
    const int s0 = i < nf ? 0 : nf;
    const int jlo = i - a.lookback > s0 ? i - a.lookback : s0;
    for (int j0 = i - 1; j0 >= jlo; j0 -= 4) {
      uint64_t kj[4];
      if (j0 == i - 1) {
#pragma unroll
        for (int u = 0; u < 4; ++u) kj[u] = win[u];  // past jlo (u > i-1-jlo): unused, the group ends the scan
      } else if (j0 == i - 5) {
#pragma unroll
        for (int u = 0; u < 4; ++u) kj[u] = win[4 + u];
      } else {
#pragma unroll
        for (int u = 0; u < 4; ++u) kj[u] = v.word(j0 - u >= jlo ? j0 - u : jlo);
      }
      bool alive = true;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = j0 - u;
        const uint32_t dr = hi - (uint32_t)(kj[u] >> 32);
        alive = j >= jlo && dr <= max_gap;
        const int dq = qi - w_qpos(kj[u]);
        const uint32_t t = dr - (uint32_t)dq;                      // dr - dq, modulo 2^32
        const uint32_t dd = (int32_t)t < 0 ? 0u - t : t;           // |dr - dq| where it matters
        int cost;
        if (NIB) {
          const int c = (int)((a.cost_nib >> (4 * (dd & 15))) & 15u);
          cost = dd == (uint32_t)band ? a.cost_at_band : c;
        } else {
          cost = gap_cost(a, dd <= (uint32_t)band ? (int)dd : band);
        }
        const int drk = dr < (uint32_t)k ? (int)dr : k;
        const int gain = dq < drk ? dq : drk;
        const int sc = w_f(kj[u]) + gain - cost;
        const bool take = alive && dr > 0 && dq > 0 && dd <= (uint32_t)band && sc > best;
        best = take ? sc : best;
        best_j = take ? j : best_j;
      }
      if (!alive) break;
      const int dm = w_dm(kj[3]);  // the group's farthest anchor (or jlo's, where the scan ends)
      if (dm < 3 && w_f(kj[3]) + dm + k <= best) break;
    }
    fmax = (i == 0 || i == nf || best > fmax) ? best : fmax;
    const uint32_t low = w_low(qi, fmax - best < 3 ? fmax - best : 3, best, best_j);
    v.set_low(i, low);
#pragma unroll
    for (int u = 7; u > 0; --u) win[u] = win[u - 1];
    win[0] = (ki & ~0xffffffffull) | low;
  }
}

template <int MAXCH>
DN_HD int chain_extract_lds(const ChainArgs& a, const LdsView& v, int n, uint32_t r) {
  // One pass over the chains in extraction order: the qualifying ones go to a register list
  // kept in stable score order (at most MAXCH); dropped_max is the best score that fell off.
  uint64_t used = 0;
  ChainLite top[MAXCH];
  int count = 0, dropped_max = -1;
  // A chain scores f[end] - f[stop] <= f[end] (f >= k > 0 everywhere), and ends come in
  // descending f: once f[end] < min_score no chain left can qualify, so the pass ends there.
  int fe = 0;
  for (int end = next_end(v, n, used, &fe); end >= 0 && fe >= a.min_score; end = next_end(v, n, used, &fe)) {
    int cnt;
    ChainLite cur = walk_chain(v, end, used, cnt);
    if (cur.score < a.min_score || cnt < a.min_anchors) continue;
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
    if (count == MAXCH) dropped_max = cur.score > dropped_max ? cur.score : dropped_max;
    count = count < MAXCH ? count + 1 : MAXCH;
  }
  if (count == 0) return 0;
  const double threshold = top[0].score * a.pri_ratio;

  // candidates (src/mapper.cpp:44-62)
  Cand* cands = a.cands + (uint64_t)r * (uint64_t)a.max_cand;
  const int64_t bw = a.bw;
  int nc = 0;
  uint32_t chosen = 0;  // bit x: top[x] became a candidate
  bool done = false;
#pragma unroll
  for (int x = 0; x < MAXCH; ++x) {
    if (done || x >= count) continue;
    if (top[x].score < threshold || nc >= a.max_cand) {  // the reference stops here
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
    cd.score = top[x].score;
    cd.reserved = 0;
    cands[nc++] = cd;
    chosen |= 1u << x;
  }
  // The list ran out while the reference would go on to a chain that fell off it.
  if (!done && dropped_max >= threshold && nc < a.max_cand) return -1;
  return nc;
}

// Phases per read (one thread): 1. minimizers into a list (`mins`, maxmin entries mstride
// apart: the thread's column of a global scratch array); 2. probes, eight minimizers at a time,
// anchors into LDS; 3. sort, DP, extraction, candidates. `ring` (the fallback sketch for reads
// with N or a window length other than WT) may alias the anchor words, which are free until
// phase 2. P8: probes through the 8-byte table s.ctab (probe_table.hpp), else the 16-byte
// buckets. CAP: the anchor slots per read (the LDS shape, 32 or 64), which also sizes the sort
// network. Returns the candidate count or -1 (the read does not fit; global path).
template <int MAXCH, int WT, bool P8, int CAP>
DN_HD int seed_chain_lds(const SeedArgs& s, const ChainArgs& a, uint32_t r, LdsView v, uint64_t* mins, int mstride,
                         int maxmin, Ring ring) {
  constexpr int cap = CAP;
  const int len = s.lens[r], k = s.k;
  if (len >= kFusedMaxLen || 2 * k > 46) return -1;
  const uint8_t* seq = s.codes + s.offs[r];
  int nmin = 0;
  int got = 0;
  LdsSeq ls;
  // The read's bases go to LDS first (into the anchor slots, free until phase 2), so the
  // sketch waits for memory once instead of at every chunk boundary.
  if (s.w == WT && stage_read(seq, len, WT, v.keys, v.stride, cap, ls))
  {
    int m = 0;
    auto put = [&](uint64_t h, uint32_t p, uint32_t st) {
      if (m < maxmin) mins[m * mstride] = pack_min(h, p, st);
      ++m;
    };
    if (k <= 15) sketch_vh<WT, uint32_t>(ls, len, k, put);
    else sketch_vh<WT, uint64_t>(ls, len, k, put);
    nmin = m;
    got = m > maxmin ? -1 : 1;
  }
  if (got < 0) return -1;
  if (got == 0) {
    nmin = 0;
    bool full = false;
    sketch_stream(ChunkReader(seq), len, k, s.w, ring, [&](uint64_t h, uint32_t p, uint32_t st) {
      if (nmin >= maxmin) full = true;
      else mins[(nmin++) * mstride] = pack_min(h, p, st);
    });
    if (full) return -1;
  }

  // 2. probes, PG minimizers in flight
  constexpr int PG = 8;
  int lo = 0, hi = cap - 1;
  bool overflow = false;
  for (int m0 = 0; m0 < nmin && !overflow; m0 += PG) {
    uint64_t e[PG];
    uint32_t off[PG], cnt[PG], first[PG];
    if (P8) {
      // The 8-byte table (s.ctab set): the four slots from each home slot in one round of loads, resolved in
      // registers; then the count and first occurrence of every key with several, in a second
      // round. Chains longer than four slots continue one slot at a time (rare).
      uint64_t q[PG][4];
#pragma unroll
      for (int u = 0; u < PG; ++u) {
        e[u] = m0 + u < nmin ? mins[(m0 + u) * mstride] : 0;
        if (m0 + u < nmin) c_load4(s.ctab, (e[u] >> 17) & s.mask, q[u]);
      }
      bool many[PG];
#pragma unroll
      for (int u = 0; u < PG; ++u) {
        off[u] = cnt[u] = first[u] = 0;
        many[u] = false;
        if (m0 + u >= nmin) continue;
        const uint32_t h = (uint32_t)(e[u] >> 17);
        uint64_t ent = 0;
        int f = c_find4(q[u], h, ent);
        if (f < 0) f = c_probe_from(s.ctab, s.mask, h, ((e[u] >> 17) + 4) & s.mask, ent);
        const uint32_t tag = c_key(ent) & 3u;
        if (f > 0 && tag == kCOne) {
          cnt[u] = 1;
          first[u] = c_val(ent);
        } else if (f > 0 && tag == kCMany) {
          off[u] = c_val(ent) + 1;
          many[u] = true;
        }
      }
#pragma unroll
      for (int u = 0; u < PG; ++u)
        if (many[u]) {
          cnt[u] = s.pos[off[u] - 1];
          first[u] = s.pos[off[u]];
        }
    } else {
      Bucket b[PG];
#pragma unroll
      for (int u = 0; u < PG; ++u) {
        e[u] = m0 + u < nmin ? mins[(m0 + u) * mstride] : 0;
        if (m0 + u < nmin) b[u] = s.table[(e[u] >> 17) & s.mask];
      }
#pragma unroll
      for (int u = 0; u < PG; ++u) {
        off[u] = cnt[u] = first[u] = 0;
        if (m0 + u >= nmin) continue;
        const uint64_t h = e[u] >> 17;
        if (b[u].key == h) {
          off[u] = b[u].offset;
          cnt[u] = b[u].count;
        } else if (b[u].key != ~0ULL) {
          probe_from(s.table, s.mask, h, ((h & s.mask) + 1) & s.mask, off[u], cnt[u]);
        }
        if (cnt[u] > s.max_occ) cnt[u] = 0;
        if (cnt[u]) first[u] = first_pos(s, off[u], cnt[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < PG; ++u) {
      if (!cnt[u] || overflow) continue;
      if ((int)cnt[u] > hi - lo + 1) {
        overflow = true;
        continue;
      }
      const uint32_t mpos = (uint32_t)(e[u] >> 1) & 0xffffu, mstrand = (uint32_t)(e[u] & 1);
      const auto put = [&](uint32_t val) {
        const uint32_t st = (val & 1) ^ mstrand;
        const int32_t qpos = st ? len - (int32_t)mpos - k : (int32_t)mpos;
        const uint64_t key = anchor_key(st, val >> 1, qpos);
        if (st) v.keys[(hi--) * v.stride] = key;
        else v.keys[(lo++) * v.stride] = key;
      };
      put(first[u]);
      // Further occurrences four loads at a time (indices clamped into the slice).
      for (uint32_t j0 = 1; j0 < cnt[u]; j0 += 4) {
        uint32_t w[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) w[x] = s.pos[off[u] + (j0 + x < cnt[u] ? j0 + x : cnt[u] - 1)];
#pragma unroll
        for (int x = 0; x < 4; ++x)
          if (j0 + x < cnt[u]) put(w[x]);
      }
    }
  }
  if (overflow) return -1;
  const int nrev = cap - 1 - hi, n = lo + nrev;
  for (int x = 0; x < nrev; ++x) v.keys[(lo + x) * v.stride] = v.word(hi + 1 + x);
  if (n == 0) return 0;

  // 3. sort, DP, extraction, candidates
  sort_lds<CAP>(v, n);
  if (a.cost_nib_ok) chain_dp_lds<true>(a, v, n, lo, k);
  else chain_dp_lds<false>(a, v, n, lo, k);
  return chain_extract_lds<MAXCH>(a, v, n, r);
}

}  // namespace dn
