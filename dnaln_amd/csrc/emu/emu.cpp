// emu.cpp — CPU build of the device lane programs (test infrastructure).
//
// Runs sw_pair / sw_generic_one exactly as a wave64 would, lane by lane, with the same
// task pairing (lane l holds tasks l and l + 64 of each 128-task group), the same
// direction-slot layout and the same dispatch rules, so the kernels' integer tricks can be
// checked against the oracle and the reference vectors without a GPU. The split fill
// (sw_split.hpp) runs its P lanes of a pair phase by phase, in the order the wave's lock step
// gives them.
#include <stdint.h>
#include <string.h>

#include <cstring>

#include <algorithm>
#include <vector>

#include "device/bw_list.hpp"
#include "device/fused_seed_chain.hpp"
#include "device/index_build.hpp"
#include "device/sw_generic.hpp"
#include "device/sw_packed.hpp"
#include "device/sw_split.hpp"

using namespace dn;

namespace {

// The walk of k_walk_win<W> for one fill wave, thread by thread: rows outside the loaded window
// read as a poison value, so a step outside [lo, top] gives a wrong result.
template <int W>
void walk_wave_window(const SwArgs& a, uint32_t wave, int rows) {
  constexpr int G = (W + 7) / 8;
  struct Th {
    bool active = false;
    uint32_t id = 0;
    int score = 0, i_end = 0, d_end = 0;
    SwTask tk{};
    Walk w{};
    uint32_t buf[kWalkSeqDwords];
    WalkSeq seq;
    uint32_t* out = nullptr;
  };
  std::vector<Th> th(128);
  for (Th& x : th) x.seq = WalkSeq(x.buf, 1);
  for (int t = 0; t < 128; ++t) {
    const uint32_t k = wave * 128 + t;
    if (k >= a.n) continue;
    Th& x = th[t];
    x.id = a.order ? a.order[k] : k;
    SwResult& r = a.res[x.id];
    x.score = r.score;
    x.i_end = r.qe;
    x.d_end = r.te;
    if (x.score <= 0) {
      r.qb = r.qe = r.tb = r.te = r.nm = r.n_cigar = r.status = 0;
      continue;
    }
    x.tk = a.tasks[x.id];
    x.seq.begin(x.tk);
    x.out = a.cigar + (uint64_t)x.id * (uint64_t)a.cigar_cap;
    x.w = walk_init(x.i_end, x.d_end, x.score);
    x.active = true;
  }
  const uint64_t* slot = dir_slot(a, wave);
  std::vector<uint64_t> win((size_t)rows * G * 64);
  for (;;) {
    int top = -1;
    for (const Th& x : th)
      if (x.active && !x.w.done) top = std::max(top, x.w.i);
    if (top < 0) break;
    const bool gran = flag_granules<W, 1>();  // whole granules: [lo, top | 3], lo a multiple of 4
    const int lo = gran ? std::max(0, (top - rows + 4) & ~3) : std::max(0, top - rows + 1);
    const int hi = gran ? (top | 3) : top;
    const uint64_t n_words = (uint64_t)(hi - lo + 1) * G * 64;
    std::fill(win.begin(), win.end(), 0x5555aaaa5555aaaaULL);
    std::copy(slot + (uint64_t)lo * G * 64, slot + (uint64_t)lo * G * 64 + n_words, win.begin());
    for (int t = 0; t < 128; ++t) {
      Th& x = th[t];
      while (x.active && !x.w.done && (x.w.i >= lo || x.w.i < 0))
        walk_step<W>(x.w, x.tk,
                     reinterpret_cast<const uint32_t*>(win.data()) + (gran ? ((t >> 6) * 64 + (t & 63)) * 4 : (t >> 6) * 64 + (t & 63)),
                     a.sp, x.out, a.cigar_cap, x.seq, lo);
    }
  }
  for (Th& x : th)
    if (x.active) walk_finish(x.w, x.i_end, x.d_end, x.score, a.cigar_cap, a.res[x.id]);
}

template <int W, bool TB>
void run_packed(const SwArgs& a0) {
  constexpr int G = (W + 7) / 8;
  int qall = 1;
  for (uint32_t k = 0; k < a0.n; ++k) qall = std::max(qall, a0.tasks[k].qlen);
  const uint64_t slot_words = fill_rows(qall) * G * 64;
  const uint32_t waves = (a0.n + 127) / 128;
  std::vector<uint64_t> dir(TB ? (size_t)waves * slot_words : 1, 0xdeadbeefdeadbeefULL);
  SwArgs a = a0;
  a.dir = dir.data();
  a.dir_slot_words = slot_words;
  for (uint32_t wv = 0; wv < waves; ++wv) {
    const uint32_t base = wv * 128;
    int qmax = 0;
    for (uint32_t k = base; k < a.n && k < base + 128; ++k) qmax = std::max(qmax, a.tasks[k].qlen);
    for (int lane = 0; lane < 64; ++lane) {
      const uint32_t ia = base + lane, ib = base + 64 + lane;
      uint16_t bh[2 * W];
      uint32_t stage[8 * G];  // one step's flag words (flag granules)
      sw_pair<W, TB>(a, ia < a.n ? ia : 0, ib < a.n ? ib : 0, ia < a.n, ib < a.n, qmax,
                     TB ? a.dir + (uint64_t)wv * slot_words : nullptr, lane, bh, 1, stage, 1);
    }
  }
  if (TB) {
    // Both walks: windows of 3 rows (4, one granule, with flag granules; fewer than the device's,
    // so many window moves) on even fill waves, one thread per task on odd ones.
    for (uint32_t wv = 0; wv < waves; ++wv) {
      if (wv % 2 == 0) walk_wave_window<W>(a, wv, flag_granules<W, 1>() ? 4 : 3);
      else
        for (uint32_t k = wv * 128; k < a.n && k < wv * 128 + 128; ++k) {
          uint32_t buf[kWalkSeqDwords];
          WalkSeq seq(buf, 1);
          walk_task<W>(a, k, seq);
        }
    }
  }
}

// The split fill over P lanes per pair and its walk, one fill group (128 tasks, P waves) at a
// time. Each wave runs to its own longest query, as on the device.
template <int W, int P, bool TB>
void run_split(const SwArgs& a0) {
  using F = SegFill<W, P, TB>;
  constexpr int U = F::U, S = F::S;
  int qall = 1;
  for (uint32_t k = 0; k < a0.n; ++k) qall = std::max(qall, a0.tasks[k].qlen);
  const uint64_t slot_words = slot_rows(qall, P) * slot_row_words(W, P);
  const uint32_t groups = (a0.n + 127) / 128;
  std::vector<uint64_t> dir(TB ? (size_t)groups * slot_words : 1, 0xdeadbeefdeadbeefULL);
  SwArgs a = a0;
  a.dir = dir.data();
  a.dir_slot_words = slot_words;
  std::vector<F> L(P);
  std::vector<uint32_t> bh((size_t)P * S);  // each lane's best row: S words of two half-words
  std::vector<uint32_t> stage((size_t)P * seg_groups(W, P) * 8);  // each lane's flag staging
  const pk16 zero = pk_splat(0);
  for (uint32_t g = 0; g < groups; ++g) {
    for (int j = 0; j < P; ++j) {
      int qmax = 0;
      for (int pl = 0; pl < 64 / P; ++pl) {
        const uint32_t ka = g * 128 + j * (64 / P) + pl, kb = ka + 64;
        if (ka < a.n) qmax = std::max(qmax, a.tasks[ka].qlen);
        if (kb < a.n) qmax = std::max(qmax, a.tasks[kb].qlen);
      }
      for (int pl = 0; pl < 64 / P; ++pl) {
        const uint32_t ka = g * 128 + j * (64 / P) + pl, kb = ka + 64;
        const bool va = ka < a.n, vb = kb < a.n;
        for (int seg = 0; seg < P; ++seg)
          L[seg].init(a, a.tasks[va ? ka : 0], a.tasks[vb ? kb : 0], va, vb, seg,
                      TB ? a.dir + (uint64_t)g * slot_words : nullptr, j, pl * P + seg,
                      reinterpret_cast<uint16_t*>(bh.data() + (size_t)seg * S), 1,
                      stage.data() + (size_t)seg * seg_groups(W, P) * 8, 1);
        const int steps = qmax + P - 1;
        for (int s0 = 0; s0 < steps; s0 += U) {
          for (F& x : L) x.step_begin();
          for (int u = 0; u < U; ++u) {
            // Cell 0 reads the left neighbour's last cell of its previous row (rest has not run
            // yet this row); the rest reads the right neighbour's cell 0 of this row.
            for (int seg = 0; seg < P; ++seg)
              L[seg].cell0(u, (s0 & 3) + u, s0 + u, seg ? L[seg - 1].x_hl : zero, seg ? L[seg - 1].x_fl : zero);
            for (int seg = 0; seg < P; ++seg)
              L[seg].rest(u, s0 + u, seg + 1 < P ? L[seg + 1].x_h0 : zero, seg + 1 < P ? L[seg + 1].x_e0 : zero);
          }
          for (F& x : L) x.step_end(s0);
        }
        uint64_t kmax[2] = {0, 0};
        for (int seg = 0; seg < P; ++seg) {
          uint64_t k2[2];
          L[seg].keys(k2[0], k2[1]);
          kmax[0] = std::max(kmax[0], k2[0]);
          kmax[1] = std::max(kmax[1], k2[1]);
        }
        if (va) split_result<TB>(a.res[ka], kmax[0]);
        if (vb) split_result<TB>(a.res[kb], kmax[1]);
      }
    }
  }
  if (TB)
    for (uint32_t k = 0; k < a.n; ++k) {
      uint32_t buf[kWalkSeqDwords];
      WalkSeq seq(buf, 1);
      walk_task<W, P>(a, k, seq);
    }
}

template <bool TB>
bool dispatch_split(int bw, int lanes, const SwArgs& a) {
#define DN_CASE(B, P)                     \
  if (bw == B && lanes == P) {            \
    run_split<2 * B + 1, P, TB>(a);       \
    return true;                          \
  }
  DN_SPLIT_LIST(DN_CASE)
#undef DN_CASE
  return false;
}

template <bool TB>
bool dispatch(int bw, const SwArgs& a) {
  switch (bw) {
#define DN_CASE(B)                   \
  case B:                            \
    run_packed<2 * B + 1, TB>(a);    \
    return true;
    DN_PACKED_BW_LIST(DN_CASE)
#undef DN_CASE
    default:
      return false;
  }
}

// The 8-byte probe table and pos2 of probe_table.hpp from an inlined 16-byte table, built as
// k_ct_need / k_ct_fill build them on the device (slot by slot, blocks in slot order).
void build_probe_table(const std::vector<Bucket>& tab, const uint32_t* pos, uint32_t max_occ,
                       std::vector<uint64_t>& ctab, std::vector<uint32_t>& pos2) {
  const uint64_t n = tab.size();
  std::vector<uint64_t> at(n);
  uint64_t total = 0;
  for (uint64_t i = 0; i < n; ++i) {
    at[i] = total;
    total += c_need(tab[i], max_occ);
  }
  pos2.assign(total + 1, 0);
  ctab.assign(n + kCTail, 0);
  for (uint64_t i = 0; i < n + kCTail; ++i) {
    const uint64_t slot = i < n ? i : (i - n) & (n - 1);
    const Bucket& b = tab[slot];
    ctab[i] = c_entry(b, max_occ, b.offset, (uint32_t)at[slot]);
    if (i >= n || !c_need(b, max_occ)) continue;
    pos2[at[i]] = b.count;
    for (uint32_t j = 0; j < b.count; ++j) pos2[at[i] + 1 + j] = pos[b.offset + j];
  }
}

}  // namespace

extern "C" {

// Same contract as the reference shim's ref_align (res6 = score qb qe tb te nm), CIGAR in
// forward order. force: 0 auto, 1 generic int32 kernel, 2 packed (fails if unavailable).
int emu_align_ex(const int* sp6, const uint8_t* pool, const int64_t* q_off, const int32_t* q_len,
                 const int64_t* t_off, const int32_t* t_len, const uint8_t* rev, const int32_t* tlo, size_t n,
                 int traceback, int32_t* res6, uint32_t* cig, int64_t* cig_off, int32_t* cig_n, size_t cap,
                 int force);
int emu_align_lanes(const int* sp6, const uint8_t* pool, const int64_t* q_off, const int32_t* q_len,
                    const int64_t* t_off, const int32_t* t_len, const uint8_t* rev, const int32_t* tlo, size_t n,
                    int traceback, int32_t* res6, uint32_t* cig, int64_t* cig_off, int32_t* cig_n, size_t cap,
                    int force, int lanes);

int emu_align(const int* sp6, const uint8_t* pool, const int64_t* q_off, const int32_t* q_len, const int64_t* t_off,
              const int32_t* t_len, size_t n, int traceback, int32_t* res6, uint32_t* cig, int64_t* cig_off,
              int32_t* cig_n, size_t cap, int force) {
  return emu_align_ex(sp6, pool, q_off, q_len, t_off, t_len, nullptr, nullptr, n, traceback, res6, cig, cig_off,
                      cig_n, cap, force);
}

// rev[i] != 0 aligns the reverse complement of query i (read backwards, as for reverse-strand
// candidates); target positions below tlo[i] read as N (a window starting before its contig).
int emu_align_ex(const int* sp6, const uint8_t* pool, const int64_t* q_off, const int32_t* q_len,
                 const int64_t* t_off, const int32_t* t_len, const uint8_t* rev, const int32_t* tlo, size_t n,
                 int traceback, int32_t* res6, uint32_t* cig, int64_t* cig_off, int32_t* cig_n, size_t cap,
                 int force) {
  return emu_align_lanes(sp6, pool, q_off, q_len, t_off, t_len, rev, tlo, n, traceback, res6, cig, cig_off, cig_n,
                         cap, force, 1);
}

// lanes > 1 with force = 2: the packed fill split over that many lanes per pair (fails with -2
// where the band width has no such form).
int emu_align_lanes(const int* sp6, const uint8_t* pool, const int64_t* q_off, const int32_t* q_len,
                    const int64_t* t_off, const int32_t* t_len, const uint8_t* rev, const int32_t* tlo, size_t n,
                    int traceback, int32_t* res6, uint32_t* cig, int64_t* cig_off, int32_t* cig_n, size_t cap,
                    int force, int lanes) {
  ScoreP sp{sp6[0], sp6[1], sp6[2], sp6[3], sp6[4], sp6[5]};
  // The packed kernel streams aligned dwords around each sequence: give the pool guard bytes.
  uint64_t pool_len = 0;
  for (size_t i = 0; i < n; ++i)
    pool_len = std::max<uint64_t>(pool_len, std::max<uint64_t>(q_off[i] + q_len[i], t_off[i] + t_len[i]));
  if (pool_len > (1ull << 40)) return -1;  // offsets beyond any host pool
  std::vector<uint8_t> guarded(pool_len + 2 * kSeqGuard, 0xEE);
  std::memcpy(guarded.data() + kSeqGuard, pool, pool_len);
  pool = guarded.data() + kSeqGuard;
  std::vector<SwTask> tasks(n);
  int maxq = 0;
  for (size_t i = 0; i < n; ++i) {
    SwTask& t = tasks[i];
    t.q = (uint64_t)(uintptr_t)(pool + q_off[i]);
    t.t = (uint64_t)(uintptr_t)(pool + t_off[i]);
    t.qlen = q_len[i];
    t.thi = std::min(t_len[i], q_len[i] + 2 * sp.bw);
    t.tlo = tlo ? std::min(tlo[i], t.thi) : 0;
    t.flags = rev && rev[i] ? 1u : 0u;
    maxq = std::max(maxq, q_len[i]);
  }
  const int ccap = 2 * maxq + 4;
  std::vector<SwResult> res(n);
  std::vector<uint32_t> cbuf((size_t)n * ccap + 1);
  bool packed_all = force != 1;
  for (size_t i = 0; i < n && packed_all; ++i) packed_all = packed_ok(sp, q_len[i]);
  bool done = false;
  if (packed_all) {
    SwArgs a{};
    a.tasks = tasks.data();
    a.n = (uint32_t)n;
    a.res = res.data();
    a.cigar = cbuf.data();
    a.cigar_cap = ccap;
    a.sp = sp;
    score_table(sp, a.tab0, a.tab1);
    if (lanes > 1)
      done = traceback ? dispatch_split<true>(sp.bw, lanes, a) : dispatch_split<false>(sp.bw, lanes, a);
    else
      done = traceback ? dispatch<true>(sp.bw, a) : dispatch<false>(sp.bw, a);
  }
  if (!done) {
    if (force == 2) return -2;
    const int W = 2 * sp.bw + 1;
    std::vector<int32_t> rows(2 * (W + 1));
    std::vector<uint8_t> dir((size_t)std::max(maxq, 1) * W);
    SwGenericArgs g{};
    g.tasks = tasks.data();
    g.n = (uint32_t)n;
    g.res = res.data();
    g.cigar = cbuf.data();
    g.cigar_cap = ccap;
    g.rows = rows.data();
    g.dir = dir.data();
    g.dir_stride = 0;
    g.sp = sp;
    g.traceback = traceback;
    for (uint32_t i = 0; i < n; ++i) sw_generic_one(g, i, 0);
  }
  size_t used = 0;
  for (size_t i = 0; i < n; ++i) {
    const SwResult& r = res[i];
    int32_t* o = res6 + 6 * i;
    o[0] = r.score;
    o[1] = r.qb;
    o[2] = r.qe;
    o[3] = r.tb;
    o[4] = r.te;
    o[5] = r.nm;
    if (r.n_cigar < 0 || r.status) return -3;
    if (used + (size_t)r.n_cigar > cap) return -1;
    cig_off[i] = (int64_t)used;
    cig_n[i] = r.n_cigar;
    for (int c = 0; c < r.n_cigar; ++c) cig[used + c] = cbuf[(size_t)i * ccap + (r.n_cigar - 1 - c)];
    used += (size_t)r.n_cigar;
  }
  return done ? 1 : 0;
}

// Runs both seed+chain programs on n reads (codes 0..4) and compares their candidates.
// mp = k, w, max_occ, bw, max_cand, min_chain, lookback, max_gap, band, min_anchors, table form
// (0: the 16-byte buckets, 1: the 8-byte probe table).
// Returns the number of reads whose candidates differ (-1 on bad input); stats[0] = reads the
// LDS program sent to the global path (overflow), stats[1] = total anchors, stats[2] = max.
// out_n / out_c (may be null): the global-memory program's candidates, n_cand per read and
// max_cand entries of (rev, contig, diag, chain score) per read.
long emu_seed_chain(const uint8_t* codes, const uint64_t* offs, const int32_t* lens, uint64_t n,
                    const void* table, uint64_t n_buckets, const uint32_t* pos, const uint64_t* contig_off,
                    int n_contigs, const int* mp, double pri_ratio, const int32_t* cost, int cap,
                    int maxmin, uint64_t* stats, int32_t* out_n, int64_t* out_c) {
  if (cap != 32 && cap != 64) return -1;  // the fused shapes
  uint64_t total = 0;
  for (uint64_t r = 0; r < n; ++r) total = std::max<uint64_t>(total, offs[r] + (uint64_t)lens[r]);
  std::vector<uint8_t> guarded(total + 256, 0xEE);
  uint8_t* gcodes = guarded.data() + (64 - ((uintptr_t)guarded.data() & 15));  // 16-aligned, >= 49 bytes before
  std::memcpy(gcodes, codes, total);
  SeedArgs sa{};
  sa.codes = gcodes;
  sa.offs = offs;
  sa.lens = lens;
  // The device's index layout: unique minimizers hold their occurrence in the bucket.
  std::vector<Bucket> tab(static_cast<const Bucket*>(table), static_cast<const Bucket*>(table) + n_buckets);
  for (Bucket& b : tab)
    if (b.key != ~0ull && b.count == 1) b.offset = pos[b.offset];
  sa.table = tab.data();
  sa.mask = n_buckets - 1;
  sa.pos = pos;
  sa.pos_inline = 1;
  sa.k = mp[0];
  sa.w = mp[1];
  sa.max_occ = (uint32_t)mp[2];
  // mp[10] = 1: probes through the 8-byte table (probe_table.hpp), as the device does for k <= 15.
  std::vector<uint64_t> ctab;
  std::vector<uint32_t> pos2;
  if (mp[10]) {
    if (2 * mp[0] > 30) return -1;
    build_probe_table(tab, pos, sa.max_occ, ctab, pos2);
    sa.ctab = ctab.data();
    sa.pos = pos2.data();
  }
  ChainArgs ca{};
  ChainSetup cs;
  cs.k = mp[0];
  cs.bw = mp[3];
  cs.max_cand = mp[4];
  cs.min_score = mp[5];
  cs.lookback = mp[6];
  cs.max_gap = mp[7];
  cs.band = mp[8];
  cs.min_anchors = mp[9];
  cs.pri_ratio = pri_ratio;
  init_chain_args(ca, cs, cost, cost, contig_off, n_contigs);
  const int mc = std::max(mp[4], 1);
  std::vector<Cand> c1((size_t)n * mc), c2((size_t)n * mc);
  std::vector<int32_t> n1(n), n2(n);
  std::vector<uint64_t> rh(sa.w);
  std::vector<uint32_t> rps(sa.w);
  Ring ring{rh.data(), rps.data(), 1};
  std::vector<uint64_t> keys(cap), mins(maxmin), gkeys, order;
  std::vector<int32_t> f32, prev;
  std::vector<ChainRec> chains;
  std::vector<uint32_t> hits;
  long differ = 0;
  stats[0] = stats[1] = stats[2] = 0;
  for (uint64_t r = 0; r < n; ++r) {
    // global path for read r alone
    uint32_t cnt = 0, nh = 0;
    uint64_t off0 = 0;
    SeedArgs s1 = sa;
    const uint32_t rid = (uint32_t)r;
    s1.list = &rid;
    s1.n_reads = 1;
    s1.anchor_count = &cnt;
    // Odd reads: the emit pass replays the count pass's probe results; even reads: it sketches
    // and probes again. The region of read r starts at 3 * offs[r] words.
    if (r & 1) {
      hits.assign(3 * ((size_t)offs[r] + (size_t)std::max(lens[r], 0)) + 3, 0xA5A5A5A5u);
      s1.hits = hits.data();
      s1.n_hits = &nh;
    }
    cnt = seed_count_read(s1, 0, ring);
    gkeys.assign(cnt + 1, 0);
    f32.assign(cnt + 1, 0);
    prev.assign(cnt + 1, 0);
    order.assign(cnt + 1, 0);
    chains.assign(cnt + 1, ChainRec{});
    s1.anchor_off = &off0;
    s1.keys = gkeys.data();
    seed_emit_read(s1, 0, ring);
    stats[1] += cnt;
    stats[2] = std::max<uint64_t>(stats[2], cnt);
    ChainArgs g = ca;
    g.list = &rid;
    g.n_reads = 1;
    g.anchor_count = &cnt;
    g.anchor_off = &off0;
    g.keys = gkeys.data();
    g.f = f32.data();
    g.prev = prev.data();
    g.order = order.data();
    g.chains = chains.data();
    g.cands = c1.data();
    g.n_cand = n1.data();
    chain_read(g, 0);
    if (out_n) {
      out_n[r] = n1[r];
      for (int c = 0; c < n1[r]; ++c) {
        const Cand& x = c1[r * mc + c];
        int64_t* o = out_c + 4 * (r * mc + c);
        o[0] = x.rev;
        o[1] = x.contig;
        o[2] = x.diag;
        o[3] = x.score;
      }
    }
    // LDS path
    ChainArgs l = ca;
    l.cands = c2.data();
    LdsView v{keys.data(), 1};
    int nc;
    if (sa.ctab) nc = cap == 32 ? seed_chain_lds<8, 10, true, 32>(sa, l, (uint32_t)r, v, mins.data(), 1, maxmin, ring)
                                : seed_chain_lds<8, 10, true, 64>(sa, l, (uint32_t)r, v, mins.data(), 1, maxmin, ring);
    else nc = cap == 32 ? seed_chain_lds<8, 10, false, 32>(sa, l, (uint32_t)r, v, mins.data(), 1, maxmin, ring)
                        : seed_chain_lds<8, 10, false, 64>(sa, l, (uint32_t)r, v, mins.data(), 1, maxmin, ring);
    if (nc < 0) {
      ++stats[0];
      continue;
    }
    n2[r] = nc;
    bool same = n1[r] == nc;
    for (int c = 0; same && c < nc; ++c) {
      const Cand& x = c1[r * mc + c];
      const Cand& y = c2[r * mc + c];
      same = x.rev == y.rev && x.contig == y.contig && x.diag == y.diag && x.score == y.score;
    }
    differ += !same;
  }
  return differ;
}

// The GPU index build's occurrence list (k_ix_count / k_ix_emit): every contig cut into
// segments of `seg` k-mer starts, each sketched on its own by sketch_segment. Writes the
// unsorted occurrence keys (hash << 32 | pos << 1 | strand) and returns their number, or minus
// the number needed when cap is short.
long emu_index_occ(const uint8_t* ref, const int64_t* coff, const int64_t* clen, int nc, int k, int w, uint32_t seg,
                   uint64_t* out, size_t cap) {
  uint64_t total = 0;
  for (int c = 0; c < nc; ++c) total = std::max<uint64_t>(total, (uint64_t)(coff[c] + clen[c]));
  std::vector<uint8_t> guarded(total + 2 * kSeqGuard, 4);
  std::memcpy(guarded.data() + kSeqGuard, ref, total);
  const uint8_t* base = guarded.data() + kSeqGuard;
  std::vector<uint64_t> rh(w);
  std::vector<uint32_t> rps(w);
  size_t n = 0;
  for (int c = 0; c < nc; ++c) {
    if (clen[c] < k) continue;
    const uint32_t nk = (uint32_t)clen[c] - (uint32_t)k + 1;
    for (uint32_t s = 0; s < nk; s += seg) {
      const IndexSeg g{(uint64_t)coff[c], (uint32_t)clen[c], s, std::min(nk, s + seg), 0};
      sketch_segment(base, g, k, w, Ring{rh.data(), rps.data(), 1}, [&](uint64_t h, uint32_t p, uint32_t st) {
        if (n < cap) out[n] = occ_key(h, g.base, p, st);
        ++n;
      });
    }
  }
  return n <= cap ? (long)n : -(long)n;
}

// Looks every query hash up through the 16-byte table (the reference's probe, src/index.hpp:24-38)
// and through the 8-byte table built from it, the way the device's probes read each: the
// four-slot load with its serial continuation (c_find4, c_probe_from) and the count and
// occurrences from pos2. table: n buckets (n a power of two, count 1 not inlined). Returns the
// number of queries whose occurrence lists differ (after the max_occ filter of
// src/mapper.cpp:27-28); stats[0] = queries found, stats[1] = queries that needed more than
// four slots.
long emu_probe_check(const void* table, uint64_t n, const uint32_t* pos, uint32_t max_occ, const uint64_t* q,
                     uint64_t nq, uint64_t* stats) {
  std::vector<Bucket> tab(static_cast<const Bucket*>(table), static_cast<const Bucket*>(table) + n);
  for (Bucket& b : tab)
    if (b.key != ~0ull && b.count == 1) b.offset = pos[b.offset];
  std::vector<uint64_t> ctab;
  std::vector<uint32_t> pos2;
  build_probe_table(tab, pos, max_occ, ctab, pos2);
  const uint64_t mask = n - 1;
  long bad = 0;
  stats[0] = stats[1] = 0;
  const Bucket* raw = static_cast<const Bucket*>(table);
  for (uint64_t x = 0; x < nq; ++x) {
    const uint64_t h = q[x];
    std::vector<uint32_t> want, got;
    uint32_t off = 0, cnt = 0;
    probe_from(raw, mask, h, h & mask, off, cnt);
    if (cnt && cnt <= max_occ) want.assign(pos + off, pos + off + cnt);
    uint64_t e[4];
    c_load4(ctab.data(), h & mask, e);
    uint64_t ent = 0;
    int f = c_find4(e, (uint32_t)h, ent);
    if (f < 0) {
      ++stats[1];
      f = c_probe_from(ctab.data(), mask, (uint32_t)h, ((h & mask) + 4) & mask, ent);
    }
    const uint32_t tag = c_key(ent) & 3u;
    if (f > 0 && tag == kCOne) got.push_back(c_val(ent));
    if (f > 0 && tag == kCMany) {
      const uint32_t at = c_val(ent);
      got.assign(pos2.data() + at + 1, pos2.data() + at + 1 + pos2[at]);
    }
    stats[0] += cnt != 0;
    bad += want != got;
  }
  return bad;
}

}  // extern "C"
