// sw_packed.hpp — inter-sequence banded Smith–Waterman, two alignments per 32-bit lane.
//
// Follows the recurrence and tie rules of src/align.cpp:31-84 (scalar spec) and
// src/align_simd.cpp:72-156 (lane, padding and best-cell semantics):
//   H(i,d) = max(0, H(i-1,d) + s(q_i, t_{i+d}), E(i,d), F(i,d))
//   E(i,d) = max(H(i-1,d+1) - (O+E), E(i-1,d+1) - E)
//   F(i,d) = max(H(i,d-1)   - (O+E), F(i,d-1)   - E)
// with row -1 and the out-of-band cell d = W equal to 0 (not -inf), F and the left H reset to
// 0 at every row, and the best cell = first maximum in row-major order.
//
// Layout on gfx950: a wave64 runs 128 alignments; each 32-bit VGPR holds the int16 values of
// alignment A (low half) and B (high half), so every recurrence step is one v_pk_*_i16. The H and
// E band rows (W+1 words each) and the W target selectors of the current row live in VGPRs
// (the band loop is fully unrolled for a compile-time W). Substitution scores come from one
// v_xor_b32 + one v_perm_b32 per cell and pair (see score_table()).
//
// Sequence bytes are streamed with one 8-byte load per stream per 4 rows, a block ahead
// (BlockStreams), from addresses kept within a few bytes of each sequence; every device sequence
// buffer carries guard bytes on both sides.
//
// Traceback storage: 4 flag bits per cell and alignment. Flags are the sign bits of four packed
// differences, gathered by v_perm_b32 into byte masks and accumulated 8 cells to a byte:
//   byte 0/1 extF (A/B), 2/3 extE, 4/5 notE, 6/7 notDiag — bit (d & 7) of the 8-byte word of
// group d >> 3, split by two perms into one 32-bit plane per task half. With four-row steps
// (W <= kGranuleW) the words of rows 4k .. 4k+3 of one group, half and lane form a 16-byte granule
// (flag_granules below), staged in LDS during the step and stored as one coalesced 1-KiB write per
// group and half; wider bands store each row's words interleaved by lane, one coalesced 256-byte
// write per group and half.
#pragma once

#include "common.hpp"

namespace dn {

// Selector encodings (per 16-bit half): query [4|q, 8|q] or [0xF0, 0xF0] for N; target [t, t] or
// [0x30, 0x30] for N (BlockStreams::next, tsel_from below). x = qsel ^ tsel indexes the table below: low
// byte 4|(q^t) picks the score, high byte 8|(q^t) replicates its sign; any N gives a selector
// >= 13, i.e. 0xFFFF = -1 (ambig 1).

// v_perm sources: s0 bytes = {A, -B, -B, -B} (x = 0..3), s1 byte 1 = 0 (sign source of a match),
// s1 byte 3 = -B (sign source of a mismatch).
DN_HD void score_table(const ScoreP& sp, uint32_t& s0, uint32_t& s1) {
  const uint32_t a = (uint32_t)(uint8_t)(int8_t)sp.match, b = (uint32_t)(uint8_t)(int8_t)(-sp.mismatch);
  s0 = a | b << 8 | b << 16 | b << 24;
  s1 = b << 24;
}

// True when the packed int16 kernel reproduces the reference exactly: all scores and gap costs
// fit the byte table and no intermediate leaves int16 (the reference's own int16 limit is
// qlen <= 30000 / match, src/align_simd.cpp:51-52).
DN_HD bool packed_ok(const ScoreP& sp, int qlen) {
  return sp.ambig == 1 && sp.match >= 1 && sp.match <= 127 && sp.mismatch >= 0 && sp.mismatch <= 127 &&
         sp.gap_open >= 0 && sp.gap_ext >= 1 && sp.gap_open + 2 * sp.gap_ext <= 2000 &&
         qlen <= 30000 / sp.match;
}

struct SwArgs {
  const SwTask* tasks;
  const uint32_t* order;   // tasks processed in this order (null: identity)
  uint32_t n;              // entries of order (or tasks)
  SwResult* res;           // indexed by task id
  uint32_t* cigar;         // reversed CIGAR ops, slot of cigar_cap per task id
  int32_t cigar_cap;
  uint64_t* dir;           // direction words, one slot per wave of the launch
  uint64_t dir_slot_words; // words per slot when woff is null
  const uint64_t* woff;    // per-wave slot offsets in words (absolute; null: wave * dir_slot_words)
  uint64_t woff_base;      // woff of the launch's first wave, which sits at dir[0]
  ScoreP sp;
  uint32_t tab0, tab1;
};

// Direction slot of launch-relative wave `wave`: fixed-size slots, or (tasks sorted by length)
// slots sized by each wave's longest query.
DN_HD uint64_t* dir_slot(const SwArgs& a, uint32_t wave) {
  return a.dir + (a.woff ? a.woff[wave] - a.woff_base : (uint64_t)wave * a.dir_slot_words);
}

DN_HD uint8_t q_at(const SwTask& t, int i) {
  if (i < 0 || i >= t.qlen) return 4;
  const uint8_t* q = (const uint8_t*)(uintptr_t)t.q;
  return (t.flags & 1) ? comp_code(q[t.qlen - 1 - i]) : q[i];
}

DN_HD uint8_t t_at(const SwTask& t, int j) {
  if (j < t.tlo || j >= t.thi) return 4;
  return ((const uint8_t*)(uintptr_t)t.t)[j];
}

// Guard bytes kept readable on both sides of every sequence buffer the packed kernel streams:
// it may load up to 4 bytes past a read, and a target window may start or end up to
// qlen + 2*bw bytes outside the reference (those bytes are masked to N, never used).
constexpr uint64_t kSeqGuard = 65536;

// The 4 bytes at addr [j, j+4) as one little-endian word, from two aligned dword loads.
DN_HD uint32_t load4(uint64_t base, int64_t j) {
  const uint64_t a = base + (uint64_t)j;
  const uint32_t* p = (const uint32_t*)(uintptr_t)(a & ~(uint64_t)3);
  const uint32_t sh = (uint32_t)(a & 3);
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_alignbyte(p[1], p[0], sh);  // v_alignbyte_b32 shifts by S2[1:0] bytes
#else
  return (uint32_t)((((uint64_t)p[1] << 32) | p[0]) >> (8 * sh));
#endif
}

// Target codes at window positions j0 .. j0+3; positions outside [tlo, thi) are masked by the
// caller. The address is clamped so it stays next to the valid range.
DN_HD uint32_t tword4(const SwTask& t, int j0) {
  int j = j0 < t.tlo - 4 ? t.tlo - 4 : j0;
  j = j > t.thi ? t.thi : j;
  return load4(t.t, j);
}

DN_HD uint32_t byte_at(uint32_t w, int u) { return (w >> (8 * u)) & 0xffu; }

// The 8 bytes at the 4-aligned address a, as two dwords (one global_load_dwordx2).
DN_HD void load8(uint64_t a, uint32_t& lo, uint32_t& hi) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef const __attribute__((address_space(1))) uint32_t* gptr;
  const gptr p = (gptr)a;
#else
  const uint32_t* p = (const uint32_t*)(uintptr_t)a;
#endif
  lo = p[0];
  hi = p[1];
}

// Bytes sh .. sh+3 (sh = bits 1:0) of hi:lo.
DN_HD uint32_t align4(uint32_t lo, uint32_t hi, uint32_t sh) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_alignbyte(hi, lo, sh);
#else
  return (uint32_t)((((uint64_t)hi << 32) | lo) >> (8 * (sh & 3)));
#endif
}

// Rows per fill slot: the traceback fill runs its rows in steps of up to 4, so a slot holds
// the wave's longest query rounded up to a multiple of 4 (rows past a query are never read).
DN_HD uint64_t fill_rows(int qmax) { return (uint64_t)(((qmax > 1 ? qmax : 1) + 3) & ~3); }

// Traceback fills with four-row steps (W <= kGranuleW with one lane per pair, and every split
// fill) keep a lane's flag words of four consecutive steps together in a 16-byte granule: granule
// (k, group g, half h, segment seg, pair p) at u32 index ((((k * GS + g) * 2 + h) * 64P + seg * 64
// + p) * 4, word r of it step 4k + r (one lane per pair: P = 1, seg = 0, steps = rows). A walk then
// reads a diagonal run of four rows from one line, where the row layout below needs a line per row.
constexpr int kGranuleW = 33;
template <int W, int P>
DN_HD constexpr bool flag_granules() { return P > 1 || W <= kGranuleW; }

// A band of W columns split over P lanes (sw_split.hpp; P = 1 is sw_pair): lane `seg` of a pair
// holds columns seg*S .. seg*S + S-1 and runs row r at step r + seg. A slot holds, per step, the
// flag words of its 128 tasks' segments: word (step s, group g of a segment, half h, segment seg,
// pair p) at u32 index ((s * GS + g) * 2 + h) * 64P + seg * 64 + p, which for P = 1 is the
// sw_pair layout. The 64 pairs of one segment lie side by side, so the walks of a wave's tasks,
// which move through nearby cells, read nearby words. A task's pair p (its index mod 64) runs in
// fill wave p / (64/P) of its group, lanes (p mod 64/P) * P + seg.
DN_HD constexpr int seg_width(int W, int P) { return (W + P - 1) / P; }
DN_HD constexpr int seg_groups(int W, int P) { return (seg_width(W, P) + 7) / 8; }
// u64 direction words per step of a slot, and steps per slot.
DN_HD constexpr uint64_t slot_row_words(int W, int P) { return (uint64_t)seg_groups(W, P) * (uint64_t)P * 64; }
DN_HD uint64_t slot_rows(int qmax, int P) { return fill_rows(qmax + P - 1); }

// The sequence streams of one lane's two alignments (h = 0: A, 1: B), read one block of four
// rows ahead. For block r0 (a multiple of 4) a query stream needs the codes of rows r0..r0+3
// (q[r0..r0+3], or q[qlen-4-r0 .. qlen-1-r0] reversed) and a target stream the codes of
// positions r0+W .. r0+W+3, which enter the band during those rows. Each block is one 8-byte
// load from a 4-aligned address, issued a block before its bytes are used, so the fill never
// waits on it; the byte offset within the 8 bytes is the same for every block of a stream.
// Offsets are clamped to the range where a block with any valid row or position reads its
// bytes: loads stay within a few bytes of the sequence, and a clamped block is fully masked.
// LAG (sw_split.hpp): rows below row0 also read as N, and the target limits may be negative.
template <int W, bool LAG = false>
struct BlockStreams {
  uint64_t q[2], t[2];
  int32_t qlen[2];
  int32_t row0;
  uint32_t tlim[2];                             // tlo | thi << 16, as int16
  uint32_t qlo[2], qhi[2], tlo_w[2], thi_w[2];  // raw bytes of the next block
  uint32_t sh;  // per byte (query A, B, target A, B): bits 1:0 the byte offset; bit 7 of the
                // query bytes: reverse-complement

  DN_HD bool rev(int h) const { return (sh >> (8 * h + 7)) & 1; }
  DN_HD int tlo(int h) const { return LAG ? (int)(int16_t)(tlim[h] & 0xffff) : (int)(tlim[h] & 0xffff); }
  DN_HD int thi(int h) const { return LAG ? (int)(int16_t)(tlim[h] >> 16) : (int)(tlim[h] >> 16); }
  DN_HD int qoff(int h, int r0) const {
    const int o = rev(h) ? qlen[h] - 4 - r0 : r0;
    const int lo = -3, hi = qlen[h] - 1;
    return o < lo ? lo : o > hi ? hi : o;
  }
  DN_HD int toff(int h, int r0) const {
    int o = r0 + W;
    o = o < tlo(h) - 3 ? tlo(h) - 3 : o;
    return o > thi(h) - 1 ? thi(h) - 1 : o;
  }
  DN_HD void fetch(int r0) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      load8((q[h] + (uint64_t)(int64_t)qoff(h, r0)) & ~(uint64_t)3, qlo[h], qhi[h]);
      load8((t[h] + (uint64_t)(int64_t)toff(h, r0)) & ~(uint64_t)3, tlo_w[h], thi_w[h]);
    }
  }
  // tlo and thi lie in [0, 32768) (thi <= qlen + 2*bw); with LAG in (-32768, 32768).
  DN_HD void init(const SwTask& a, const SwTask& b, int first_row = 0) {
    const SwTask* tk[2] = {&a, &b};
    sh = 0;
    row0 = first_row;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      q[h] = tk[h]->q;
      t[h] = tk[h]->t;
      qlen[h] = tk[h]->qlen;
      tlim[h] = (uint32_t)(uint16_t)tk[h]->tlo | (uint32_t)(uint16_t)tk[h]->thi << 16;
      const bool r = tk[h]->flags & 1;
      const uint32_t sq = (uint32_t)(q[h] + (r ? (uint64_t)(int64_t)(qlen[h] - 4) : 0)) & 3;
      const uint32_t st = (uint32_t)(t[h] + W) & 3;
      sh |= (sq | (r ? 0x80u : 0u)) << (8 * h) | st << (16 + 8 * h);
    }
    fetch(0);
  }
  // Block r0 from the fetched bytes, then fetches block r0 + 4. qc: the query codes of rows
  // r0..r0+3 (byte u = row r0+u; A in qc[0], B in qc[1]) as 0..3, complemented for reversed
  // queries, and 0xF0 for N or a row past the query. ts[u]: the target selector of position
  // r0+W+u (both halves).
  DN_HD void next(int r0, uint32_t (&qc)[2], uint32_t (&ts)[4]) {
    uint32_t tt[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t qw = align4(qlo[h], qhi[h], sh >> (8 * h));
      const uint32_t tw = align4(tlo_w[h], thi_w[h], sh >> (16 + 8 * h));
      // Reversed queries read backwards and complement (c ^ 3 swaps A/T and C/G; N = 4 becomes 7).
      const uint32_t c = rev(h) ? perm_b32(0, qw, 0x00010203u) ^ 0x03030303u : qw;
      const uint32_t qbad = bytes_from(qlen[h] - r0) | (LAG ? bytes_below(row0 - r0) : 0u);
      qc[h] = bfi(qbad, 0xF0F0F0F0u, perm_b32(0xF0F0F0F0u, 0x03020100u, c));
      const int j0 = r0 + W;
      const uint32_t tbad = bytes_below(tlo(h) - j0) | bytes_from(thi(h) - j0);
      tt[h] = bfi(tbad, 0x30303030u, perm_b32(0x30303030u, 0x03020100u, tw));
    }
    fetch(r0 + 4);
#pragma unroll
    for (int u = 0; u < 4; ++u) ts[u] = perm_b32(tt[1], tt[0], (uint32_t)u * 0x0101u | (uint32_t)(u + 4) * 0x01010000u);
  }
};

// Query selector of row u of a block (both halves): bytes [4|cA, 8|cA, 4|cB, 8|cB]; the N code
// 0xF0 stays >= 13.
DN_HD uint32_t qsel_row(const uint32_t (&qc)[2], int u) {
  return perm_b32(qc[1], qc[0], (uint32_t)u * 0x01010101u + 0x04040000u) | 0x08040804u;
}

// Target selector half with validity and N.
DN_HD uint32_t tsel_from(uint32_t c, bool valid) { return (!valid || c > 3) ? 0x3030u : c * 0x0101u; }

// ------------------------------------------------------------------------------------ walk
// src/align.cpp:87-143, one cell per step.
struct Walk {
  int i, d, v, nm, run, state;  // state 0: H (diagonal), 1: E (insertion), 2: F (deletion)
  int n, pend_len, status;
  uint32_t pend_op;
  bool done;
};

// The kSeqWindow aligned bytes of one sequence around the walk's position, in memory private to
// the walking thread (LDS on the device): dword x of the window at buf[x * stride], the threads
// of a block interleaved, so every byte read is free of bank conflicts. The walk moves one
// position per step, so a window is refilled (16-byte loads) every kSeqWindow steps. Measured on
// MI355X (2M x 150 bp, W = 33): a dword cached in a register and refilled every 4 steps walks in
// 2.7 ms, since with a wave's worth of streams per SIMD its refills miss L2; 64-byte windows in
// 1.8 ms, 32 bytes in 1.2 ms, 16 bytes in 1.24 ms (the LDS each thread holds sets the occupancy
// that hides the flag loads, which matters more than how often a window refills); prefetching
// the next 64 bytes into registers cost more occupancy than it saved (2.6 ms). The walk is then
// bound by its instruction issue more than by its loads (about 108 VALU instructions per wave
// step kept the SIMDs ~74% busy), which is why a step is written branch-free. Windows reach up to
// kSeqWindow - 1 bytes before and after the bytes read: every sequence buffer carries guard
// bytes.
constexpr int kSeqWindow = 32;  // bytes
static_assert(kSeqWindow == 32, "WalkSeq::bases divides offsets by 32 with a shift");

// The query and target windows of one walk: kWalkSeqDwords dwords at buf, stride apart (query
// window first). begin() takes the task's constants once (where row i's query byte lies, and
// whether to complement it); bases() refreshes a window only when the walk leaves it and then
// issues both LDS reads before using either, with no branch on the strand.
constexpr int kWalkSeqDwords = 2 * kSeqWindow / 4;
struct WalkSeq {
  uint32_t* buf = nullptr;
  int stride = 1;
  // Positions as 32-bit byte offsets from kSeqWindow-aligned origins: the query byte of row i at
  // qorg + qofs + qstep * i, target position j at torg + tofs + j (64-bit only on a refill).
  uint64_t qorg = 0, torg = 0;
  int qofs = 0, tofs = 0, qstep = 1;
  uint32_t qrc = 0;  // 3 for a reversed query: its codes are complemented (c ^ 3 = 3 - c)
  int qlen = 0, tlo = 0, thi = 0;
  int qwin = 0, twin = 0;  // the windows held (offset / kSeqWindow); a begin() sets impossible ones
  WalkSeq() = default;
  DN_HD WalkSeq(uint32_t* b, int s) : buf(b), stride(s) {}
  DN_HD void begin(const SwTask& tk) {
    const bool rev = tk.flags & 1;
    const uint64_t q0 = rev ? tk.q + (uint64_t)(int64_t)(tk.qlen - 1) : tk.q;
    qorg = q0 & ~(uint64_t)(kSeqWindow - 1);
    qofs = (int)(q0 & (kSeqWindow - 1));
    torg = tk.t & ~(uint64_t)(kSeqWindow - 1);
    tofs = (int)(tk.t & (kSeqWindow - 1));
    qstep = rev ? -1 : 1;
    qrc = rev ? 3u : 0u;
    qlen = tk.qlen;
    tlo = tk.tlo;
    thi = tk.thi;
    qwin = twin = -(1 << 30);
  }
  // The kSeqWindow aligned bytes at a into the window at dword offset off.
  DN_HD void fill(int off, uint64_t a) {
    Quad c[kSeqWindow / 16];
#pragma unroll
    for (int x = 0; x < kSeqWindow / 16; ++x) c[x] = load16(a + 16 * x);
#pragma unroll
    for (int x = 0; x < kSeqWindow / 16; ++x) {
      buf[(off + 4 * x) * stride] = c[x].x;
      buf[(off + 4 * x + 1) * stride] = c[x].y;
      buf[(off + 4 * x + 2) * stride] = c[x].z;
      buf[(off + 4 * x + 3) * stride] = c[x].w;
    }
  }
  // The query code of row i and the target code of position j, 4 outside the sequences (a
  // reversed query's codes complemented, as comp_code).
  DN_HD void bases(int i, int j, uint8_t& qc, uint8_t& tc) {
    const bool qin = i >= 0 && i < qlen, tin = j >= tlo && j < thi;
    const int oq = qofs + qstep * i, ot = tofs + j;
    const int wq = oq >> 5, wt = ot >> 5;  // kSeqWindow = 32
    if (qin && wq != qwin) {
      fill(0, qorg + (uint64_t)(int64_t)(wq * kSeqWindow));
      qwin = wq;
    }
    if (tin && wt != twin) {
      fill(kSeqWindow / 4, torg + (uint64_t)(int64_t)(wt * kSeqWindow));
      twin = wt;
    }
    const uint32_t xq = (uint32_t)oq & (kSeqWindow - 1), xt = (uint32_t)ot & (kSeqWindow - 1);
    const uint32_t dq = buf[(xq >> 2) * stride], dt = buf[(kSeqWindow / 4 + (xt >> 2)) * stride];
    const uint32_t cq = (dq >> (8 * (xq & 3))) & 0xffu, ct = (dt >> (8 * (xt & 3))) & 0xffu;
    qc = qin ? (uint8_t)(cq < 4 ? cq ^ qrc : 4u) : (uint8_t)4;
    tc = tin ? (uint8_t)ct : (uint8_t)4;
  }
};

DN_HD void walk_push(Walk& w, uint32_t op, int len, uint32_t* out, int cap) {
  if (w.pend_len > 0 && w.pend_op == op) {
    w.pend_len += len;
    return;
  }
  if (w.pend_len > 0) {
    if (w.n < cap) out[w.n] = (uint32_t)w.pend_len << 4 | w.pend_op;
    ++w.n;
  }
  w.pend_op = op;
  w.pend_len = len;
}

DN_HD void walk_flush(Walk& w, uint32_t* out, int cap) {
  if (w.pend_len > 0) {
    if (w.n < cap) out[w.n] = (uint32_t)w.pend_len << 4 | w.pend_op;
    ++w.n;
  }
  w.pend_len = 0;
}

// One traceback step. plane holds the flag words of the walk's own task (its pair and half of
// the fill group) from step row0 on: column d lies in segment seg = d / S at offset c = d % S
// (S = seg_width(W, P)); the word of row i, group c / 8 at ((i + seg - row0) * GS + c / 8) * 128P
// + seg * 64, bytes extF, extE, E and not-diagonal, bit c & 7 in each (for P = 1: row i, group
// d / 8 at ((i - row0) * G + d / 8) * 128). With flag_granules, plane is the task's granule
// column (half h, pair p: (h * 64P + p) * 4) and the word of row i, column c of segment seg is
// word s & 3 of granule (s >> 2) * GS + c / 8 of that segment, s = i + seg - row0 (row0 a multiple
// of 4): a diagonal run reads four rows from one line. A step reads only row w.i, and none when
// w.i < 0.
template <int W, int P = 1>
DN_HD void walk_step(Walk& w, const SwTask& tk, const uint32_t* plane, const ScoreP& sp, uint32_t* out, int cap,
                     WalkSeq& seq, int row0 = 0) {
  constexpr int S = seg_width(W, P), GS = seg_groups(W, P);
  constexpr bool kGran = flag_granules<W, P>();
  const int oe = sp.gap_open + sp.gap_ext, ge = sp.gap_ext;
  // Branch-free: every lane of a wave computes the same step, whatever state it is in, and the
  // CIGAR pushes are two predicated stores (a branch per state made each wave step run the union
  // of its lanes' paths). The outcome equals walk_task's reference order of cases:
  //   fin   (state H, row < 0 or score <= 0): push the match run, flush, done;
  //   bad   (cell outside the band; never reached by a consistent walk): status 1, done;
  //   diag  (state H, diagonal flag): score and NM of the pair, run + 1, row - 1;
  //   open  (state H, not diagonal): push the match run, enter E or F by the notE flag;
  //   ext / close (state E or F): one gap cell (E: row - 1, column + 1; F: column - 1); a cell
  //         without its extension flag closes the gap: push it, NM += its length, back to H.
  const bool fin = w.state == 0 && (w.i < 0 || w.v <= 0);
  const bool bad = !fin && (w.i < 0 || w.d < 0 || w.d >= W);
  const bool live = !fin && !bad;
  const int seg = P == 1 ? 0 : (live ? w.d : 0) / S, col = (live ? w.d : 0) - seg * S;
  uint32_t word = 0;
  if (live) {
    // 32-bit word indexes: a slot holds far fewer than 2^32 words.
    if (kGran) {
      const int r = w.i + seg - row0;
      word = plane[(uint32_t)(((r >> 2) * GS + (col >> 3)) * (512 * P) + seg * 256 + (r & 3))];
    } else {
      word = plane[(uint32_t)(((w.i + seg - row0) * GS + (col >> 3)) * (128 * P) + seg * 64)];
    }
  }
  // The two bases of the cell (4 outside the sequences; mostly from the windows).
  uint8_t a, c;
  seq.bases(w.i, w.i + w.d, a, c);
  const int b = col & 7;
  const bool s0 = w.state == 0, s1 = w.state == 1;
  const bool diag = live && s0 && !((word >> (24 + b)) & 1);
  const bool open = live && s0 && !diag;
  const bool gap = live && !s0;
  const bool ext = gap && ((s1 ? (word >> (8 + b)) : (word >> b)) & 1);
  const bool close = gap && !ext;
  // Run push (walk_push): the match run on fin / open, the gap on close.
  const int run1 = w.run + 1;
  const bool push = ((fin || open) && w.run > 0) || close;
  const uint32_t pop = close ? (s1 ? 1u : 2u) : 0u;
  const int plen = close ? run1 : w.run;
  const bool merge = push && w.pend_len > 0 && w.pend_op == pop;
  const bool emit = push && !merge && w.pend_len > 0;
  if (emit && w.n < cap) out[w.n] = (uint32_t)w.pend_len << 4 | w.pend_op;
  int n = w.n + (emit ? 1 : 0);
  int pend_len = merge ? w.pend_len + plen : push ? plen : w.pend_len;
  const uint32_t pend_op = push && !merge ? pop : w.pend_op;
  // Flush on fin (walk_flush).
  const bool flush = fin && pend_len > 0;
  if (flush && n < cap) out[n] = (uint32_t)pend_len << 4 | pend_op;
  n += flush ? 1 : 0;
  pend_len = fin ? 0 : pend_len;
  w.n = n;
  w.pend_len = pend_len;
  w.pend_op = pend_op;
  // Score, NM, run, coordinates, state.
  w.v += diag ? -sub_score(sp, a, c) : gap ? (ext ? ge : oe) : 0;
  w.nm += diag ? (int)(a != c || a > 3) : close ? run1 : 0;
  w.run = diag || ext ? run1 : (fin || open || close) ? 0 : w.run;
  w.i -= diag || (gap && s1) ? 1 : 0;
  w.d += gap ? (s1 ? 1 : -1) : 0;
  w.state = open ? (((word >> (16 + b)) & 1) ? 2 : 1) : close ? 0 : w.state;
  w.status = bad ? 1 : w.status;
  w.done = fin || bad;
}

DN_HD Walk walk_init(int i_end, int d_end, int score) {
  Walk w;
  w.i = i_end;
  w.d = d_end;
  w.v = score;
  w.nm = 0;
  w.run = 0;
  w.state = 0;
  w.n = 0;
  w.pend_len = 0;
  w.pend_op = 0;
  w.status = 0;
  w.done = score <= 0;
  return w;
}

DN_HD void walk_finish(const Walk& w, int i_end, int d_end, int score, int cap, SwResult& r) {
  r.score = score;
  if (score <= 0) {
    r.qb = r.qe = r.tb = r.te = r.nm = 0;
    r.n_cigar = 0;
    r.status = 0;
    return;
  }
  r.qb = w.i + 1;
  r.tb = w.i + 1 + w.d;
  r.qe = i_end + 1;
  r.te = i_end + d_end + 1;
  r.nm = w.nm;
  r.n_cigar = w.n <= cap ? w.n : -1;
  r.status = w.status;
}

// ------------------------------------------------------------------------------------ fill
// One lane: tasks ida and idb (va / vb false for an empty half), rows 0..qmax-1 of the wave.
// Score-only: writes the scores. With flags (TB): stores the flag words into the wave's slot
// and writes the score and the best cell (row in qe, diagonal in te) for the walk.
// Traceback fills wider than this keep the saved best row in LDS (bh_lds) instead of VGPRs.
#if !defined(DN_BEST_ROW_IN_REGS)
#define DN_BEST_ROW_IN_REGS 33
#endif
constexpr int kBestRowInRegs = DN_BEST_ROW_IN_REGS;

// Best-row words per fill wave when the row lives in LDS (one u32 per band cell and lane).
DN_HD int best_row_lds_words(int W) { return W > kBestRowInRegs ? W * 64 : 0; }
// LDS words per fill wave of the traceback fill (the saved best row, or the flag granules'
// staging: four rows of every group and half per lane).
static_assert(kBestRowInRegs >= kGranuleW, "a fill stages flag granules or keeps its best row in LDS, not both");
DN_HD int fill_lds_words(int W) { return W > kBestRowInRegs ? W * 64 : W <= kGranuleW ? ((W + 7) / 8) * 8 * 64 : 0; }

// stage (flag_granules): this lane's LDS words for one step's flags, word ((g * 2 + h) * 4 + r)
// at stage[((g * 2 + h) * 4 + r) * stage_stride].
template <int W, bool TB>
DN_HD void sw_pair(const SwArgs& a, uint32_t ida, uint32_t idb, bool va, bool vb, int qmax, uint64_t* slot,
                   int lane, uint16_t* bh_lds = nullptr, int bh_stride = 0, uint32_t* stage = nullptr,
                   int stage_stride = 0) {
  constexpr int G = (W + 7) / 8;
  // An empty half reads task 0's sequences (every launch has a task 0) with no valid row or
  // target position, so its loads stay in bounds and its cells score as N.
  SwTask ta = a.tasks[va ? ida : 0], tb = a.tasks[vb ? idb : 0];
  if (!va) ta.qlen = ta.tlo = ta.thi = 0, ta.flags = 0;
  if (!vb) tb.qlen = tb.tlo = tb.thi = 0, tb.flags = 0;

  const pk16 zero = pk_splat(0);
  const pk16 oe = pk_splat(a.sp.gap_open + a.sp.gap_ext), ge = pk_splat(a.sp.gap_ext);
  const uint32_t tab0 = a.tab0, tab1 = a.tab1;

  // Rows per step of the row loop: the row body is unrolled U times, so the band's target codes
  // (T[d] = target position i + d) shift once per step instead of once per row. Only where the
  // extra registers cost no occupancy: the traceback fill up to W = 33 (2 waves per SIMD either
  // way); the score-only fill would drop from 3 to 2 waves, wider bands would spill more.
  // A step always runs all U rows; rows past qmax score below the best and store their flags
  // into the slot's padding rows (fill_rows).
  constexpr int U = TB && W <= kGranuleW ? 4 : 1;
  constexpr bool kGran = TB && flag_granules<W, 1>();
  pk16 H[W + 1], E[W + 1];
  uint32_t T[W + U - 1];
#pragma unroll
  for (int d = 0; d <= W; ++d) H[d] = E[d] = zero;
  // Target positions 0..W-1, four at a time.
#pragma unroll
  for (int d0 = 0; d0 < W; d0 += 4) {
    const uint32_t wa = tword4(ta, d0), wb = tword4(tb, d0);
#pragma unroll
    for (int u = 0; u < 4 && d0 + u < W; ++u) {
      const int j = d0 + u;
      T[j] = tsel_from(byte_at(wa, u), j >= ta.tlo && j < ta.thi) |
             tsel_from(byte_at(wb, u), j >= tb.tlo && j < tb.thi) << 16;
    }
  }
  // Per block of four rows: the query codes of its rows (QC) and the target selectors of the
  // positions entering the band during them (TS, consumed U per step).
  BlockStreams<W> bs;
  bs.init(ta, tb);
  uint32_t QC[2], TS[4];
  bs.next(0, QC, TS);

  // The saved best row: in VGPRs up to W = kBestRowInRegs; wider, in LDS as half-words
  // (alignment A at bh_lds[2 * d * bh_stride], B beside it), which keeps W registers free for
  // occupancy. It changes only on rows whose maximum beats the best so far.
  constexpr bool kBhLds = TB && W > kBestRowInRegs;
  pk16 best = zero;
  pk16 BH[TB && !kBhLds ? W : 1];
  uint32_t best_i = 0xffffffffu;
  if (TB) {
#pragma unroll
    for (int d = 0; d < W; ++d) {
      if (kBhLds) bh_lds[2 * d * bh_stride] = bh_lds[2 * d * bh_stride + 1] = 0;
      else BH[kBhLds ? 0 : d] = zero;
    }
  }
  uint32_t* slot_lane = reinterpret_cast<uint32_t*>(slot) + lane;
  // Flag bit of band column d within its group of eight: 0x01010101 << (d & 7). Shifted by a
  // value that is 0 at run time but unknown to the compiler, the eight masks live in scalar
  // registers, so each flag update is one three-operand and_or (which takes no literal).
  const uint32_t zs = (uint32_t)a.sp.bw >> 16;  // 0 (bw <= 1000)
  uint32_t bits[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) bits[k] = 0x01010101u << (k + zs);

  for (int i0 = 0; i0 < qmax; i0 += U) {
    // Target positions i0 + W .. i0 + W + U - 2, which rows i0 + 1 .. of this step reach.
#pragma unroll
    for (int v = 0; v + 1 < U; ++v) T[W + v] = TS[v];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u;
      const uint32_t qsel = qsel_row(QC, (i0 & 3) + u);

      pk16 hdiag = H[0], f = zero, hleft = zero, rowmax = zero;
      uint32_t acc12 = 0, acc34 = 0;
#pragma unroll
      for (int d = 0; d < W; ++d) {
        const pk16 hup = H[d + 1], eup = E[d + 1];
        const pk16 s = pk_from(perm_b32(tab0, tab1, T[d + u] ^ qsel));
        const pk16 hd = pk_add(hdiag, s);
        // E and F are kept at max(E, 0) and max(F, 0) (clamped subtractions of non-negative
        // values), which leaves H = max(hd, E, F, 0) unchanged and makes the zero floor free.
        // Their flags then differ from the reference's only in cells with H = 0, which no walk
        // reads: a walk stops at a cell whose score is 0, and every cell of a gap run has a
        // positive gap score, whose larger operand is exact.
        const pk16 eo = pk_subsat(hup, oe), ee = pk_subsat(eup, ge);
        const pk16 e = pk_max(eo, ee);
        const pk16 fo = pk_subsat(hleft, oe), fe = pk_subsat(f, ge);
        f = pk_max(fo, fe);
        const pk16 h = pk_max(pk_max(hd, e), f);
        H[d] = h;
        E[d] = e;
        if (TB) {
          const uint32_t p12 = perm_b32(pk_bits(pk_sub(eo, ee)), pk_bits(pk_sub(fo, fe)), 0x0B0A0908u);
          const uint32_t p34 = perm_b32(pk_bits(pk_sub(hd, h)), pk_bits(pk_sub(e, h)), 0x0B0A0908u);
          const uint32_t bit = bits[d & 7];
          acc12 = and_or(p12, bit, acc12);
          acc34 = and_or(p34, bit, acc34);
          if ((d & 7) == 7 || d == W - 1) {
            // One 32-bit plane per half (bytes: extF, extE, E, not-diagonal), so that a walk,
            // which follows one half, reads only its own plane.
            const uint32_t fa = perm_b32(acc34, acc12, 0x06040200u), fb = perm_b32(acc34, acc12, 0x07050301u);
            if (kGran) {
              stage[(((d >> 3) * 2) * 4 + u) * stage_stride] = fa;
              stage[(((d >> 3) * 2 + 1) * 4 + u) * stage_stride] = fb;
#if defined(__HIP_DEVICE_COMPILE__)
              // Left free, the scheduler sinks these LDS stores into later rows and the fill
              // spills 800 registers; with the barrier it keeps 249 and no scratch.
              __builtin_amdgcn_sched_barrier(0);
#endif
            } else {
              uint32_t* row = slot_lane + (uint64_t)(i * G + (d >> 3)) * 128;
              store_stream(row, fa);
              store_stream(row + 64, fb);
            }
            acc12 = acc34 = 0;
          }
          rowmax = pk_max(rowmax, h);
        } else {
          best = pk_max(best, h);
        }
        hleft = h;
        hdiag = hup;
      }
      if (TB) {
        const uint32_t m = pk_signmask(pk_sub(best, rowmax));  // halves whose row max beats best
        if (m) {
          if (kBhLds) {
            if (m & 0xffffu) {
#pragma unroll
              for (int d = 0; d < W; ++d) bh_lds[2 * d * bh_stride] = (uint16_t)pk_bits(H[d]);
            }
            if (m >> 16) {
#pragma unroll
              for (int d = 0; d < W; ++d) bh_lds[2 * d * bh_stride + 1] = (uint16_t)(pk_bits(H[d]) >> 16);
            }
          } else {
#pragma unroll
            for (int d = 0; d < (TB && !kBhLds ? W : 1); ++d)
              BH[d] = pk_from(bfi(m, pk_bits(H[d]), pk_bits(BH[d])));
          }
          best = pk_max(best, rowmax);
          best_i = bfi(m, (uint32_t)i * 0x10001u, best_i);
        }
      }
    }
    if (kGran) {
      // The step's four rows of each group and half, from LDS, as one 16-byte granule store.
      uint32_t* gran = reinterpret_cast<uint32_t*>(slot) + (uint64_t)(i0 >> 2) * G * 512 + (uint64_t)lane * 4;
#pragma unroll
      for (int x = 0; x < 2 * G; ++x) {
        Quad q;
        q.x = stage[(x * 4 + 0) * stage_stride];
        q.y = stage[(x * 4 + 1) * stage_stride];
        q.z = stage[(x * 4 + 2) * stage_stride];
        q.w = stage[(x * 4 + 3) * stage_stride];
        store_stream16(gran + (uint64_t)x * 256, q);
      }
    }
    // Shift the band U positions and bring in target position i0 + W + U - 1; move the block's
    // remaining target selectors down by U.
#pragma unroll
    for (int d = 0; d + 1 < W; ++d) T[d] = T[d + U];
    T[W - 1] = TS[U - 1];
#pragma unroll
    for (int v = 0; v + U < 4; ++v) TS[v] = TS[v + U];
    // The next block's selectors, from bytes fetched a block ago (the stores of a whole block
    // lie between that load and this use).
    if (((i0 + U) & 3) == 0) bs.next(i0 + U, QC, TS);
  }

  const int sa = pk_lo(best), sb = pk_hi(best);
  if (!TB) {
    if (va) {
      SwResult& r = a.res[ida];
      r.score = sa;
      r.qb = r.qe = r.tb = r.te = r.nm = r.n_cigar = r.status = 0;
    }
    if (vb) {
      SwResult& r = a.res[idb];
      r.score = sb;
      r.qb = r.qe = r.tb = r.te = r.nm = r.n_cigar = r.status = 0;
    }
    return;
  }
  // Best cell: first d of the saved best row that attains the score.
  int da = -1, db = -1;
#pragma unroll
  for (int d = W - 1; d >= 0; --d) {
    const int ha = kBhLds ? (int16_t)bh_lds[2 * d * bh_stride] : pk_lo(BH[kBhLds ? 0 : d]);
    const int hb = kBhLds ? (int16_t)bh_lds[2 * d * bh_stride + 1] : pk_hi(BH[kBhLds ? 0 : d]);
    if (ha == sa) da = d;
    if (hb == sb) db = d;
  }
  if (va) {
    SwResult& r = a.res[ida];
    r.score = sa;
    r.qe = (int16_t)(best_i & 0xffff);
    r.te = da;
    r.qb = r.tb = r.nm = r.n_cigar = r.status = 0;
  }
  if (vb) {
    SwResult& r = a.res[idb];
    r.score = sb;
    r.qe = (int16_t)(best_i >> 16);
    r.te = db;
    r.qb = r.tb = r.nm = r.n_cigar = r.status = 0;
  }
}

// Traceback of one alignment from the end cell sw_pair<W, true> (or sw_split with P lanes per
// pair) left in its result. k is the launch-relative position of the task (fill group k / 128,
// pair k % 64, half (k / 64) % 2): one thread per task reads its flag words straight from the slot.
template <int W, int P = 1>
DN_HD void walk_task(const SwArgs& a, uint32_t k, WalkSeq& seq) {
  const uint32_t id = a.order ? a.order[k] : k;
  SwResult& r = a.res[id];
  const int score = r.score, i_end = r.qe, d_end = r.te;
  if (score <= 0) {
    r.qb = r.qe = r.tb = r.te = r.nm = r.n_cigar = r.status = 0;
    return;
  }
  const SwTask tk = a.tasks[id];
  const uint32_t wave = k >> 7, pair = k & 63, half = (k >> 6) & 1;
  const uint32_t* plane = reinterpret_cast<const uint32_t*>(dir_slot(a, wave)) +
                          (flag_granules<W, P>() ? (half * 64 * P + pair) * 4 : half * 64 * P + pair);
  uint32_t* out = a.cigar + (uint64_t)id * (uint64_t)a.cigar_cap;
  Walk w = walk_init(i_end, d_end, score);
  seq.begin(tk);
  while (!w.done) walk_step<W, P>(w, tk, plane, a.sp, out, a.cigar_cap, seq);
  walk_finish(w, i_end, d_end, score, a.cigar_cap, r);
}

}  // namespace dn
