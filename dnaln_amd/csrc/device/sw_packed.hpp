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
// Traceback storage: 4 flag bits per cell and alignment. Flags are the sign bits of four packed
// differences, gathered by v_perm_b32 into byte masks and accumulated 8 cells to a byte:
//   byte 0/1 extF (A/B), 2/3 extE, 4/5 notE, 6/7 notDiag — bit (d & 7) of the 8-byte word of
// group d >> 3. The words of one wave are interleaved by lane: word(i, g, lane) at
// slot[(i * G + g) * 64 + lane], so every store is one coalesced 512-byte write.
#pragma once

#include "common.hpp"

namespace dn {

// Selector encodings (per 16-bit half): query [4|q, 8|q] or [0xF0, 0xF0] for N; target [t, t] or
// [0x30, 0x30] for N. x = qsel ^ tsel indexes the table below: low byte 4|(q^t) picks the score,
// high byte 8|(q^t) replicates its sign; any N gives a selector >= 13, i.e. 0xFFFF = -1 (ambig 1).
DN_HD uint32_t qsel_half(uint8_t c) { return c < 4 ? (uint32_t)((8u | c) << 8 | (4u | c)) : 0xF0F0u; }
DN_HD uint32_t tsel_half(uint8_t c) { return c < 4 ? (uint32_t)c * 0x0101u : 0x3030u; }

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
  uint64_t* dir;           // direction words: one slot of dir_slot_words per wave
  uint64_t dir_slot_words;
  ScoreP sp;
  uint32_t tab0, tab1;
};

DN_HD uint8_t q_at(const SwTask& t, int i) {
  if (i < 0 || i >= t.qlen) return 4;
  const uint8_t* q = (const uint8_t*)(uintptr_t)t.q;
  return (t.flags & 1) ? comp_code(q[t.qlen - 1 - i]) : q[i];
}

DN_HD uint8_t t_at(const SwTask& t, int j) {
  if (j < t.tlo || j >= t.thi) return 4;
  return ((const uint8_t*)(uintptr_t)t.t)[j];
}

// ------------------------------------------------------------------------------------ walk
// src/align.cpp:87-143, one cell per step so the two alignments of a lane walk in lock-step.
struct Walk {
  int i, d, v, nm, run, state;  // state 0: H (diagonal), 1: E (insertion), 2: F (deletion)
  int n, pend_len, status;
  uint32_t pend_op;
  bool done;
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

template <int W>
DN_HD void walk_step(Walk& w, int half, const SwTask& tk, const uint64_t* slot_lane, const ScoreP& sp,
                     uint32_t* out, int cap) {
  constexpr int G = (W + 7) / 8;
  const int oe = sp.gap_open + sp.gap_ext, ge = sp.gap_ext;
  if (w.state == 0 && (w.i < 0 || w.v <= 0)) {
    if (w.run) walk_push(w, 0, w.run, out, cap);
    w.run = 0;
    walk_flush(w, out, cap);
    w.done = true;
    return;
  }
  if (w.i < 0 || w.d < 0 || w.d >= W) {  // never reached by a consistent walk
    w.status = 1;
    w.done = true;
    return;
  }
  const uint64_t word = slot_lane[(uint64_t)(w.i * G + (w.d >> 3)) * 64];
  const int b = w.d & 7;
  if (w.state == 0) {
    if (!((word >> (48 + 8 * half + b)) & 1)) {  // diagonal
      const uint8_t a = q_at(tk, w.i), c = t_at(tk, w.i + w.d);
      w.v -= sub_score(sp, a, c);
      w.nm += (a != c || a > 3);
      ++w.run;
      --w.i;
      return;
    }
    if (w.run) walk_push(w, 0, w.run, out, cap);
    w.run = 0;
    w.state = ((word >> (32 + 8 * half + b)) & 1) ? 2 : 1;  // notE set: deletion
    return;
  }
  if (w.state == 1) {
    const bool ext = (word >> (16 + 8 * half + b)) & 1;
    w.v += ext ? ge : oe;
    ++w.run;
    --w.i;
    ++w.d;
    if (!ext) {
      walk_push(w, 1, w.run, out, cap);
      w.nm += w.run;
      w.run = 0;
      w.state = 0;
    }
    return;
  }
  const bool ext = (word >> (8 * half + b)) & 1;
  w.v += ext ? ge : oe;
  ++w.run;
  --w.d;
  if (!ext) {
    walk_push(w, 2, w.run, out, cap);
    w.nm += w.run;
    w.run = 0;
    w.state = 0;
  }
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
// One lane: tasks ta and tb (task ids, >= n_tasks for an empty half), rows 0..qmax-1.
template <int W, bool TB>
DN_HD void sw_pair(const SwArgs& a, uint32_t ida, uint32_t idb, bool va, bool vb, int qmax, uint64_t* slot,
                   int lane) {
  constexpr int G = (W + 7) / 8;
  SwTask ta, tb;
  if (va) ta = a.tasks[ida];
  else ta.qlen = ta.tlo = ta.thi = 0, ta.flags = 0, ta.q = ta.t = 0;
  if (vb) tb = a.tasks[idb];
  else tb.qlen = tb.tlo = tb.thi = 0, tb.flags = 0, tb.q = tb.t = 0;

  const pk16 zero = pk_splat(0);
  const pk16 oe = pk_splat(a.sp.gap_open + a.sp.gap_ext), ge = pk_splat(a.sp.gap_ext);
  const uint32_t tab0 = a.tab0, tab1 = a.tab1;

  pk16 H[W + 1], E[W + 1];
  uint32_t T[W];
#pragma unroll
  for (int d = 0; d <= W; ++d) H[d] = E[d] = zero;
#pragma unroll
  for (int d = 0; d < W; ++d) T[d] = tsel_half(t_at(ta, d)) | tsel_half(t_at(tb, d)) << 16;
  uint32_t qsel = qsel_half(q_at(ta, 0)) | qsel_half(q_at(tb, 0)) << 16;

  pk16 best = zero;
  pk16 BH[TB ? W : 1];
  uint32_t best_i = 0xffffffffu;
  if (TB) {
#pragma unroll
    for (int d = 0; d < (TB ? W : 1); ++d) BH[d] = zero;
  }
  uint64_t* slot_lane = slot + lane;

  for (int i = 0; i < qmax; ++i) {
    // Next row's bases, loaded a row ahead.
    const uint8_t qa_n = q_at(ta, i + 1), qb_n = q_at(tb, i + 1);
    const uint8_t ta_n = t_at(ta, i + W), tb_n = t_at(tb, i + W);

    pk16 hdiag = H[0], f = zero, hleft = zero, rowmax = zero;
    uint32_t acc12 = 0, acc34 = 0;
#pragma unroll
    for (int d = 0; d < W; ++d) {
      const pk16 hup = H[d + 1], eup = E[d + 1];
      const pk16 s = pk_from(perm_b32(tab0, tab1, T[d] ^ qsel));
      const pk16 hd = pk_add(hdiag, s);
      const pk16 eo = pk_sub(hup, oe), ee = pk_sub(eup, ge);
      const pk16 e = pk_max(eo, ee);
      const pk16 fo = pk_sub(hleft, oe), fe = pk_sub(f, ge);
      f = pk_max(fo, fe);
      const pk16 h = pk_max(pk_max(pk_max(hd, e), zero), f);
      H[d] = h;
      E[d] = e;
      if (TB) {
        const uint32_t p12 = perm_b32(pk_bits(pk_sub(eo, ee)), pk_bits(pk_sub(fo, fe)), 0x0B0A0908u);
        const uint32_t p34 = perm_b32(pk_bits(pk_sub(hd, h)), pk_bits(pk_sub(e, h)), 0x0B0A0908u);
        const uint32_t bit = 0x01010101u << (d & 7);
        acc12 |= p12 & bit;
        acc34 |= p34 & bit;
        if ((d & 7) == 7 || d == W - 1) {
          slot_lane[(uint64_t)(i * G + (d >> 3)) * 64] = (uint64_t)acc12 | (uint64_t)acc34 << 32;
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
#pragma unroll
        for (int d = 0; d < (TB ? W : 1); ++d) BH[d] = pk_from(bfi(m, pk_bits(H[d]), pk_bits(BH[d])));
        best = pk_max(best, rowmax);
        best_i = bfi(m, (uint32_t)i * 0x10001u, best_i);
      }
    }
#pragma unroll
    for (int d = 0; d + 1 < W; ++d) T[d] = T[d + 1];
    T[W - 1] = tsel_half(ta_n) | tsel_half(tb_n) << 16;
    qsel = qsel_half(qa_n) | qsel_half(qb_n) << 16;
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
  for (int d = (TB ? W : 1) - 1; d >= 0; --d) {
    if (pk_lo(BH[d]) == sa) da = d;
    if (pk_hi(BH[d]) == sb) db = d;
  }
  const int ia = (int16_t)(best_i & 0xffff), ib = (int16_t)(best_i >> 16);
  uint32_t* oa = a.cigar + (uint64_t)(va ? ida : 0) * (uint64_t)a.cigar_cap;
  uint32_t* ob = a.cigar + (uint64_t)(vb ? idb : 0) * (uint64_t)a.cigar_cap;
  Walk wa = walk_init(ia, da, va ? sa : 0), wb = walk_init(ib, db, vb ? sb : 0);
  while (!(wa.done && wb.done)) {
    if (!wa.done) walk_step<W>(wa, 0, ta, slot_lane, a.sp, oa, a.cigar_cap);
    if (!wb.done) walk_step<W>(wb, 1, tb, slot_lane, a.sp, ob, a.cigar_cap);
  }
  if (va) walk_finish(wa, ia, da, sa, a.cigar_cap, a.res[ida]);
  if (vb) walk_finish(wb, ib, db, sb, a.cigar_cap, a.res[idb]);
}

}  // namespace dn
