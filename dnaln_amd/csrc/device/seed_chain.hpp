// seed_chain.hpp — per-read programs for sketching, index probes, anchors, chaining and
// candidate selection (one read per thread).
//
//   sketch        src/minimizer.cpp:5-68 — leftmost minimum of every window of w k-mers inside
//                 an N-free run, deduplicated against the previous valid window's choice.
//                 Streamed with a w-entry ring and a running minimum that is rescanned only
//                 when it leaves the window.
//   probe         src/index.hpp:24-38 — linear probing from hash & mask over 16-byte buckets.
//   anchors       src/mapper.cpp:24-40 — n = 0 or n > max_occ skipped; strand = (v & 1) ^ m.strand;
//                 qpos = strand ? len - pos - k : pos. Stored as one sortable 64-bit key
//                 (strand << 63 | rpos << 32 | qpos), which orders exactly as chain.cpp:14-18.
//   chain         src/chain.cpp:12-71 — DP over the previous `lookback` anchors (strict >, so the
//                 nearest predecessor wins ties), extraction by f desc / index asc, stable sort
//                 of chains by score desc.
//   candidates    src/mapper.cpp:44-62.
#pragma once

#include "common.hpp"
#include "probe_table.hpp"
#include "sketch_vh.hpp"

namespace dn {

struct SeedArgs {
  const uint8_t* codes;     // reads, forward codes 0..4
  const uint64_t* offs;
  const int32_t* lens;
  uint32_t n_reads;         // entries to process (of `list` when set)
  const uint32_t* list;     // read ids to process (null: 0 .. n_reads-1)
  const Bucket* table;
  uint64_t mask;
  const uint32_t* pos;      // occurrences; pos2 of probe_table.hpp when ctab is set
  const uint64_t* ctab;     // the 8-byte probe table (probe_table.hpp), or null
  int pos_inline;           // a bucket with count 1 holds its occurrence itself in `offset`
  int k, w;
  uint32_t max_occ;
  uint32_t* anchor_count;   // seed_count output, by list position
  const uint64_t* anchor_off;   // exclusive scan of anchor_count
  uint64_t* keys;           // seed_emit output
  // The probe results of seed_count_read, which seed_emit_read replays instead of sketching and
  // probing again (null: it does both again): per minimizer with an anchor, three words
  // (mpos << 1 | strand, off, cnt), read r's at hits[3 * offs[r] ..] (a read has at most len
  // minimizers, so the regions of distinct reads never overlap), and their number by list
  // position in n_hits.
  uint32_t* hits;
  uint32_t* n_hits;
};

DN_HD uint32_t read_id(const uint32_t* list, uint32_t i) { return list ? list[i] : i; }

DN_HD uint64_t hash64(uint64_t key, uint64_t mask) {  // src/minimizer.hpp:21-30
  key = (~key + (key << 21)) & mask;
  key = key ^ key >> 24;
  key = ((key + (key << 3)) + (key << 8)) & mask;
  key = key ^ key >> 14;
  key = ((key + (key << 2)) + (key << 4)) & mask;
  key = key ^ key >> 28;
  key = (key + (key << 31)) & mask;
  return key;
}

// First occurrence of a minimizer whose bucket gave (off, cnt >= 1).
DN_HD uint32_t first_pos(const SeedArgs& s, uint32_t off, uint32_t cnt) {
  return s.pos_inline && cnt == 1 ? off : s.pos[off];
}

// Ring of the last w k-mers: hash and (pos << 1 | strand), at ring[slot * stride].
struct Ring {
  uint64_t* h;
  uint32_t* ps;
  int stride;
};

// Reads a read's codes front to back in 16-byte aligned chunks, one chunk of look-ahead, so a
// 150-bp read costs ~11 independent vector loads instead of 150 dependent byte loads. It touches
// up to 15 bytes before and 31 bytes after the read: the buffer carries guard bytes.
struct ChunkReader {
  const uint32_t* base;  // 16-byte aligned chunk holding byte 0
  uint32_t off;          // offset of byte 0 in that chunk
  int cur;
  uint32_t c0, c1, c2, c3, n0, n1, n2, n3;
  DN_HD explicit ChunkReader(const uint8_t* p) {
    const uintptr_t a = (uintptr_t)p;
    base = (const uint32_t*)(a & ~(uintptr_t)15);
    off = (uint32_t)(a & 15);
    cur = 0;
    c0 = base[0], c1 = base[1], c2 = base[2], c3 = base[3];
    n0 = base[4], n1 = base[5], n2 = base[6], n3 = base[7];
  }
  // Sequential access only: i grows by one per call.
  DN_HD uint8_t operator()(int i) {
    const uint32_t x = off + (uint32_t)i;
    const int c = (int)(x >> 4);
    if (c != cur) {
      c0 = n0, c1 = n1, c2 = n2, c3 = n3;
      cur = c;
      const uint32_t* q = base + 4 * (c + 1);
      n0 = q[0], n1 = q[1], n2 = q[2], n3 = q[3];
    }
    const uint32_t sel = (x >> 2) & 3;
    const uint32_t wv = sel == 0 ? c0 : (sel == 1 ? c1 : (sel == 2 ? c2 : c3));
    return (uint8_t)(wv >> (8 * (x & 3)));
  }
};

// Plain indexed reader (host-side genome sketching).
struct PtrReader {
  const uint8_t* p;
  DN_HD uint8_t operator()(int i) const { return p[i]; }
};

// Calls emit(hash, pos, strand) for every minimizer of seq[0, len), in position order.
template <class Reader, class Emit>
DN_HD void sketch_stream(Reader&& seq, int len, int k, int w, Ring ring, Emit&& emit) {
  const uint64_t mask = (1ULL << (2 * k)) - 1;
  const int shift = 2 * (k - 1);
  const uint64_t kInvalid = ~0ULL;
  uint64_t fwd = 0, rev = 0;
  int run = 0, kidx = 0, slot = 0;
  uint64_t mh = kInvalid;
  uint32_t mps = 0;
  int midx = -1;
  int64_t last = -1;
  for (int i = 0; i < len; ++i) {
    const uint8_t c = seq(i);
    if (c > 3) {
      run = 0;
      kidx = 0;
      slot = 0;
      mh = kInvalid;
      midx = -1;
      continue;
    }
    fwd = ((fwd << 2) | c) & mask;
    rev = (rev >> 2) | ((uint64_t)(3 - c) << shift);
    if (++run < k) continue;
    const uint32_t strand = rev < fwd;
    const uint64_t h = fwd != rev ? hash64(strand ? rev : fwd, mask) : kInvalid;
    const uint32_t ps = (uint32_t)(i - k + 1) << 1 | strand;
    ring.h[slot * ring.stride] = h;
    ring.ps[slot * ring.stride] = ps;
    // The window is [kidx - w + 1, kidx]; rescan when the running minimum left it.
    if (midx >= 0 && midx < kidx - w + 1) {
      const int lo = kidx - w + 1;
      int s = lo % w;
      mh = ring.h[s * ring.stride];
      mps = ring.ps[s * ring.stride];
      midx = lo;
      for (int x = lo + 1; x <= kidx; ++x) {
        s = s + 1 == w ? 0 : s + 1;
        const uint64_t hx = ring.h[s * ring.stride];
        if (hx < mh) {
          mh = hx;
          mps = ring.ps[s * ring.stride];
          midx = x;
        }
      }
    } else if (midx < 0 || h < mh) {
      mh = h;
      mps = ps;
      midx = kidx;
    }
    if (kidx + 1 >= w && mh != kInvalid) {
      const uint32_t p = mps >> 1;
      if ((int64_t)p != last) emit(mh, p, mps & 1);
      last = p;
    }
    ++kidx;
    slot = slot + 1 == w ? 0 : slot + 1;
  }
}

DN_HD uint64_t anchor_key(uint32_t strand, uint32_t rpos, int32_t qpos) {
  return (uint64_t)strand << 63 | (uint64_t)rpos << 32 | (uint32_t)qpos;
}
DN_HD uint32_t key_strand(uint64_t k) { return (uint32_t)(k >> 63); }
DN_HD uint32_t key_rpos(uint64_t k) { return (uint32_t)(k >> 32) & 0x7fffffffu; }
DN_HD int32_t key_qpos(uint64_t k) { return (int32_t)(uint32_t)k; }

// Continues linear probing for h from slot i (index.hpp:24-38).
DN_HD void probe_from(const Bucket* table, uint64_t mask, uint64_t h, uint64_t i, uint32_t& off, uint32_t& cnt) {
  for (;;) {
    const Bucket b = table[i];
    if (b.key == h) {
      off = b.offset;
      cnt = b.count;
      return;
    }
    if (b.key == ~0ULL) {
      off = cnt = 0;
      return;
    }
    i = (i + 1) & mask;
  }
}

// Probes minimizers as the sketch emits them, two deep: the home bucket of minimizer m is
// loaded while minimizer m-1's bucket is resolved (and its first occurrence loaded) and m-2's
// occurrences are consumed, so one read keeps several independent loads in flight. Calls
// on(mpos, mstrand, off, cnt, first) in minimizer order for every minimizer with
// 1 <= cnt <= max_occ (src/mapper.cpp:24-30); first = pos[off] when NeedPos. With the 8-byte
// table (s.ctab) a probe loads the four slots from the home slot at once, and for a key with
// several occurrences `off` indexes pos2 (s.pos), whose word before it holds the count.
template <bool NeedPos, class OnHit>
struct ProbePipe {
  const SeedArgs& s;
  OnHit& on;
  bool v1 = false, v2 = false;
  uint64_t h1 = 0;
  uint32_t mp1 = 0, ms1 = 0;
  Bucket b1{};
  uint64_t q1[4] = {0, 0, 0, 0};
  uint32_t mp2 = 0, ms2 = 0, off2 = 0, cnt2 = 0, first2 = 0;
  DN_HD ProbePipe(const SeedArgs& a, OnHit& f) : s(a), on(f) {}
  DN_HD void advance() {
    if (v2) {
      on(mp2, ms2, off2, cnt2, first2);
      v2 = false;
    }
    if (v1) {
      uint32_t off = 0, cnt = 0, first = 0;
      bool hit;
      if (s.ctab) {
        uint64_t ent = 0;
        int f = c_find4(q1, (uint32_t)h1, ent);
        if (f < 0) f = c_probe_from(s.ctab, s.mask, (uint32_t)h1, ((h1 & s.mask) + 4) & s.mask, ent);
        const uint32_t tag = c_key(ent) & 3u;
        hit = f > 0 && tag != kCSkip;
        if (hit && tag == kCOne) {
          cnt = 1;
          off = first = c_val(ent);  // as with inlined buckets: a single occurrence is its own offset
        } else if (hit) {
          off = c_val(ent) + 1;
          cnt = s.pos[off - 1];
          if (NeedPos) first = s.pos[off];
        }
      } else {
        if (b1.key == h1) {
          off = b1.offset;
          cnt = b1.count;
        } else if (b1.key != ~0ULL) {
          probe_from(s.table, s.mask, h1, ((h1 & s.mask) + 1) & s.mask, off, cnt);
        }
        hit = cnt != 0 && cnt <= s.max_occ;
        if (hit && NeedPos) first = first_pos(s, off, cnt);
      }
      if (hit) {
        mp2 = mp1;
        ms2 = ms1;
        off2 = off;
        cnt2 = cnt;
        first2 = first;
        v2 = true;
      }
      v1 = false;
    }
  }
  DN_HD void push(uint64_t h, uint32_t mpos, uint32_t mstrand) {
    Bucket bn{};
    uint64_t qn[4] = {0, 0, 0, 0};
    if (s.ctab) c_load4(s.ctab, h & s.mask, qn);
    else bn = s.table[h & s.mask];
    advance();
    h1 = h;
    mp1 = mpos;
    ms1 = mstrand;
    b1 = bn;
#pragma unroll
    for (int x = 0; x < 4; ++x) q1[x] = qn[x];
    v1 = true;
  }
  DN_HD void drain() {
    advance();
    advance();
  }
};

// The minimizers of read r, in position order: the register sketch for window length 10 and an
// N-free read (one branch-free step per k-mer), else the streamed sketch with its ring.
template <class Emit>
DN_HD void sketch_read(const SeedArgs& a, uint32_t r, Ring ring, Emit&& emit) {
  const uint8_t* seq = a.codes + a.offs[r];
  const int len = a.lens[r];
  if (a.w == 10 && !read_has_n(seq, len)) {
    if (a.k <= 15) sketch_vh<10, uint32_t>(ChunkReader(seq), len, a.k, emit);
    else sketch_vh<10, uint64_t>(ChunkReader(seq), len, a.k, emit);
  } else {
    sketch_stream(ChunkReader(seq), len, a.k, a.w, ring, emit);
  }
}

DN_HD uint32_t seed_count_read(const SeedArgs& a, uint32_t i, Ring ring) {
  const uint32_t r = read_id(a.list, i);
  uint32_t total = 0, nh = 0;
  uint32_t* hit = a.hits ? a.hits + 3 * a.offs[r] : nullptr;
  auto on = [&](uint32_t mpos, uint32_t mstrand, uint32_t off, uint32_t cnt, uint32_t) {
    total += cnt;
    if (hit) {
      hit[3 * nh] = mpos << 1 | mstrand;
      hit[3 * nh + 1] = off;
      hit[3 * nh + 2] = cnt;
      ++nh;
    }
  };
  ProbePipe<false, decltype(on)> pipe(a, on);
  sketch_read(a, r, ring, [&](uint64_t h, uint32_t p, uint32_t st) { pipe.push(h, p, st); });
  pipe.drain();
  if (hit) a.n_hits[i] = nh;
  return total;
}

// Forward-strand anchors fill the slice from the front in emission order, reverse-strand
// anchors from the back; both runs then come out nearly ascending, which the insertion sort of
// the chainer exploits. The final order is the unique sorted order either way.
DN_HD void seed_emit_read(const SeedArgs& a, uint32_t i, Ring ring) {
  const uint32_t r = read_id(a.list, i);
  const int len = a.lens[r];
  uint64_t* out = a.keys + a.anchor_off[i];
  int64_t lo = 0, hi = (int64_t)a.anchor_count[i] - 1;
  const int k = a.k;
  auto on = [&](uint32_t mpos, uint32_t mstrand, uint32_t off, uint32_t cnt, uint32_t first) {
    const auto put = [&](uint32_t v) {
      const uint32_t st = (v & 1) ^ mstrand;
      const int32_t qpos = st ? len - (int32_t)mpos - k : (int32_t)mpos;
      const uint64_t key = anchor_key(st, v >> 1, qpos);
      if (st) out[hi--] = key;
      else out[lo++] = key;
    };
    put(first);
    // Further occurrences eight loads at a time (indices clamped into the slice), so a
    // repeated minimizer's hundreds of occurrences do not cost a round trip each.
    for (uint32_t j0 = 1; j0 < cnt; j0 += 8) {
      uint32_t w[8];
#pragma unroll
      for (int x = 0; x < 8; ++x) w[x] = a.pos[off + (j0 + x < cnt ? j0 + x : cnt - 1)];
#pragma unroll
      for (int x = 0; x < 8; ++x)
        if (j0 + x < cnt) put(w[x]);
    }
  };
  if (a.hits) {
    // The count pass's probe results, four at a time with their first occurrences loaded together.
    const uint32_t* hit = a.hits + 3 * a.offs[r];
    const uint32_t nh = a.n_hits[i];
    for (uint32_t x0 = 0; x0 < nh; x0 += 4) {
      uint32_t mps[4], off[4], cnt[4], first[4];
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        const uint32_t k = x0 + x < nh ? x0 + x : x0;
        mps[x] = hit[3 * k];
        off[x] = hit[3 * k + 1];
        cnt[x] = hit[3 * k + 2];
      }
#pragma unroll
      for (int x = 0; x < 4; ++x) first[x] = first_pos(a, off[x], cnt[x]);
#pragma unroll
      for (int x = 0; x < 4; ++x)
        if (x0 + x < nh) on(mps[x] >> 1, mps[x] & 1, off[x], cnt[x], first[x]);
    }
    return;
  }
  ProbePipe<true, decltype(on)> pipe(a, on);
  sketch_read(a, r, ring, [&](uint64_t h, uint32_t p, uint32_t st) { pipe.push(h, p, st); });
  pipe.drain();
}

// ------------------------------------------------------------------------------------ chaining
struct ChainRec {
  int32_t score;
  uint32_t strand;
  int32_t n_anchors;
  uint32_t rpos;
  int64_t dmin, dmax;
};

// A candidate locus (src/mapper.hpp:61-67). The reference's window start is always diag - bw
// (src/mapper.cpp:60), so it is not stored; `score` is the score of the chain it came from.
struct Cand {
  uint32_t rev;
  int32_t contig;
  int64_t diag;
  int32_t score;
  int32_t reserved;
};

DN_HD int64_t cand_wstart(const Cand& c, int bw) { return c.diag - bw; }

struct ChainArgs {
  uint32_t n_reads;        // entries to process (of `list` when set)
  const uint32_t* list;    // read ids (null: identity); anchor arrays are by list position
  const uint32_t* anchor_count;
  const uint64_t* anchor_off;
  uint64_t* keys;          // sorted in place
  int32_t* f;              // per anchor
  int32_t* prev;
  uint64_t* order;         // per anchor scratch: extraction order
  ChainRec* chains;        // per anchor scratch: chains of the read
  const int32_t* cost;     // gap cost by |dr - dq|, 0..band
  uint64_t cost_pk[4];     // the same, one byte per drift 0..31, when cost_packed
  int cost_packed;         // band <= 31 and every cost in [0, 255]
  uint64_t cost_nib;       // the same, four bits per drift 0..15, when cost_nib_ok
  int cost_at_band;        // cost[band]
  int cost_nib_ok;         // band <= 16 and every cost in [0, 15]
  int k, max_gap, band, lookback, min_score, min_anchors;
  int max_cand, bw;
  const uint64_t* contig_off;
  int n_contigs;
  double pri_ratio;        // keep chains scoring >= pri_ratio * best
  Cand* cands;             // n_reads * max_cand
  int32_t* n_cand;
};

template <class T, class Less>
DN_HD void insertion_sort(T* a, int n, Less less) {
  for (int i = 1; i < n; ++i) {
    const T x = a[i];
    int j = i - 1;
    while (j >= 0 && less(x, a[j])) {
      a[j + 1] = a[j];
      --j;
    }
    a[j + 1] = x;
  }
}

template <class T, class Less>
DN_HD void heap_sort(T* a, int n, Less less) {
  auto sift = [&](int root, int end) {
    while (2 * root + 1 < end) {
      int c = 2 * root + 1;
      if (c + 1 < end && less(a[c], a[c + 1])) ++c;
      if (!less(a[root], a[c])) return;
      const T t = a[root];
      a[root] = a[c];
      a[c] = t;
      root = c;
    }
  };
  for (int s = n / 2 - 1; s >= 0; --s) sift(s, n);
  for (int end = n - 1; end > 0; --end) {
    const T t = a[0];
    a[0] = a[end];
    a[end] = t;
    sift(0, end);
  }
}

// Gap cost of drift dd (0 <= dd <= band): from the kernel-argument copy when it fits one, so
// the chain DP's inner loop does not wait on memory.
DN_HD int gap_cost(const ChainArgs& a, int64_t dd) {
  if (a.cost_packed) {
    const int x = (int)dd;
    const uint64_t w = x < 16 ? (x < 8 ? a.cost_pk[0] : a.cost_pk[1]) : (x < 24 ? a.cost_pk[2] : a.cost_pk[3]);
    return (int)((w >> (8 * (x & 7))) & 0xff);
  }
  return a.cost[dd];
}

// Packs cost[0..band] into a.cost_pk when band <= 31 and all costs fit a byte, and into
// a.cost_nib (drifts 0..15) and a.cost_at_band when band <= 16 and all costs fit four bits.
inline void pack_gap_cost(ChainArgs& a, const int32_t* host_cost, int band) {
  a.cost_nib = 0;
  a.cost_at_band = band >= 0 ? host_cost[band] : 0;
  a.cost_nib_ok = band >= 0 && band <= 16;
  for (int d = 0; d <= band && d <= 16; ++d) {
    if (host_cost[d] < 0 || host_cost[d] > 15) a.cost_nib_ok = 0;
    else if (d < 16) a.cost_nib |= (uint64_t)host_cost[d] << (4 * d);
  }
  a.cost_packed = 0;
  a.cost_pk[0] = a.cost_pk[1] = a.cost_pk[2] = a.cost_pk[3] = 0;
  if (band > 31) return;
  for (int d = 0; d <= band; ++d) {
    if (host_cost[d] < 0 || host_cost[d] > 255) return;
    a.cost_pk[d >> 3] |= (uint64_t)host_cost[d] << (8 * (d & 7));
  }
  a.cost_packed = 1;
}

// Chain and candidate parameters of one run (src/chain.hpp:30-36 after the band clamp of
// src/mapper.cpp:9, and src/mapper.cpp:44-62).
struct ChainSetup {
  int k, max_gap, band, lookback, min_score, min_anchors, max_cand, bw;
  double pri_ratio;
};

// Fills the parameter part of a ChainArgs. The gap-cost table cost[0..band] (host copy) is
// packed for the kernels only after the band is set, from the band the DP uses.
inline void init_chain_args(ChainArgs& a, const ChainSetup& s, const int32_t* cost_dev, const int32_t* cost_host,
                            const uint64_t* contig_off, int n_contigs) {
  a.k = s.k;
  a.max_gap = s.max_gap;
  a.band = s.band;
  a.lookback = s.lookback;
  a.min_score = s.min_score;
  a.min_anchors = s.min_anchors;
  a.max_cand = s.max_cand;
  a.bw = s.bw;
  a.pri_ratio = s.pri_ratio;
  a.cost = cost_dev;
  pack_gap_cost(a, cost_host, a.band);
  a.contig_off = contig_off;
  a.n_contigs = n_contigs;
}

DN_HD int contig_of(const uint64_t* off, int nc, uint64_t gpos) {  // src/reference.cpp:25-32
  int lo = 0, hi = nc;
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (off[mid] <= gpos) lo = mid;
    else hi = mid;
  }
  return lo;
}

DN_HD void chain_read(const ChainArgs& a, uint32_t li) {
  const uint32_t r = read_id(a.list, li);
  const int n = (int)a.anchor_count[li];
  const uint64_t base = a.anchor_off[li];
  Cand* cands = a.cands + (uint64_t)r * (uint64_t)a.max_cand;
  int nc = 0;
  if (n > 0) {
    uint64_t* key = a.keys + base;
    int32_t* f = a.f + base;
    int32_t* prev = a.prev + base;
    auto lt = [](uint64_t x, uint64_t y) { return x < y; };
    // Nearly sorted on arrival (see seed_emit_read): insertion sort unless the read is large.
    if (n <= 64) insertion_sort(key, n, lt);
    else heap_sort(key, n, lt);

    const int k = a.k;
    for (int i = 0; i < n; ++i) {
      const uint64_t ki = key[i];
      const uint32_t si = key_strand(ki);
      const int64_t ri = key_rpos(ki), qi = key_qpos(ki);
      int best = k, best_j = -1;
      const int lo = i - a.lookback > 0 ? i - a.lookback : 0;
      for (int j = i - 1; j >= lo; --j) {
        const uint64_t kj = key[j];
        if (key_strand(kj) != si) break;
        const int64_t dr = ri - key_rpos(kj);
        if (dr > a.max_gap) break;
        const int64_t dq = qi - key_qpos(kj);
        if (dr <= 0 || dq <= 0) continue;
        const int64_t dd = dr > dq ? dr - dq : dq - dr;
        if (dd > a.band) continue;
        const int64_t m = dq < dr ? dq : dr;
        const int sc = f[j] + (int)(m < k ? m : k) - gap_cost(a, dd);
        if (sc > best) {
          best = sc;
          best_j = j;
        }
      }
      f[i] = best;
      prev[i] = best_j;
    }

    // Extraction order: f desc, index asc, as one ascending 64-bit key.
    uint64_t* order = a.order + base;
    for (int i = 0; i < n; ++i) order[i] = (uint64_t)(uint32_t)(0x7fffffff - f[i]) << 32 | (uint32_t)i;
    if (n <= 64) insertion_sort(order, n, lt);
    else heap_sort(order, n, lt);

    ChainRec* ch = a.chains + base;
    int m = 0;
    for (int o = 0; o < n; ++o) {
      const int end = (int)(uint32_t)order[o];
      if (prev[end] == -2) continue;  // used
      int i = end, cnt = 0;
      int64_t dmin = INT64_MAX, dmax = INT64_MIN;
      const int fend = f[end];
      while (i >= 0 && prev[i] != -2) {
        const int nx = prev[i];
        prev[i] = -2;
        ++cnt;
        const int64_t d = (int64_t)key_rpos(key[i]) - key_qpos(key[i]);
        dmin = d < dmin ? d : dmin;
        dmax = d > dmax ? d : dmax;
        i = nx;
      }
      const int score = fend - (i >= 0 ? f[i] : 0);
      if (score < a.min_score || cnt < a.min_anchors) continue;
      ChainRec c;
      c.score = score;
      c.strand = key_strand(key[end]);
      c.n_anchors = cnt;
      c.rpos = key_rpos(key[end]);
      c.dmin = dmin;
      c.dmax = dmax;
      ch[m++] = c;
    }
    // Stable sort by score desc (insertion sort is stable).
    insertion_sort(ch, m, [](const ChainRec& x, const ChainRec& y) { return x.score > y.score; });

    // Candidates (mapper.cpp:44-62): score >= pri_ratio * best, at most max_cand, deduplicated.
    const int64_t bw = a.bw;
    const double min_score = ch[0].score * a.pri_ratio;
    for (int c = 0; c < m; ++c) {
      if (ch[c].score < min_score || nc >= a.max_cand) break;
      const int64_t diag = ch[c].dmin + (ch[c].dmax - ch[c].dmin) / 2;
      bool dup = false;
      for (int x = 0; x < nc && !dup; ++x) {
        const int64_t dd = cands[x].diag - diag;
        dup = cands[x].rev == ch[c].strand && (dd < 0 ? -dd : dd) <= bw;
      }
      if (dup) continue;
      Cand cd;
      cd.rev = ch[c].strand;
      cd.contig = contig_of(a.contig_off, a.n_contigs, ch[c].rpos);
      cd.diag = diag;
      cd.score = ch[c].score;
      cd.reserved = 0;
      cands[nc++] = cd;
    }
  }
  a.n_cand[r] = nc;
}

}  // namespace dn
