// sw_split.hpp — the packed int16 fill of sw_packed.hpp with one pair of alignments spread over
// P lanes, for bands too wide for one lane's registers.
//
// Same recurrence, tie rules and flag bits as sw_pair (src/align.cpp:31-84, src/align_simd.cpp:
// 72-156). Lane `seg` (0 .. P-1) of a pair holds the band columns seg*S .. seg*S + S-1
// (S = seg_width(W, P)) and, at step s, computes row s - seg of them: one row behind its left
// neighbour. Its first cell needs the left neighbour's last cell of the same row (H and F), which
// that lane computed a step earlier; its last cell needs the right neighbour's first cell of the
// row above (H and E) as its upper neighbour, which that lane computes in the same step, first.
// So every row is two phases: cell 0 (`cell0`), then the other S-1 cells (`rest`), with one
// neighbour exchange before each. On the device the P lanes are adjacent in a wave and the
// exchanges are DPP row shifts; the CPU emulation calls the same phases lane by lane.
//
// The last segment holds up to S - 1 columns past the band. Its column W - (P-1)*S is the
// out-of-band cell d = W of sw_pair, kept at H = E = 0; the columns right of it only feed each
// other and are left out of the row maximum and the best cell. Rows before 0 (the first steps of
// lanes seg > 0) read as N and stay 0, so nothing reaches the neighbours early.
//
// Flags go to the slot in step order as four-step granules (flag_granules in sw_packed.hpp),
// each segment's 64 pairs side by side: a step's flag words wait in LDS, and every four steps a
// wave writes each group and half as P runs of 64/P granules (256 bytes or more each). Each lane
// keeps the first cell of its own
// segment attaining its maximum (saved best row in LDS, as sw_pair's wide bands); the pair's
// best cell is the lane maximum ordered by score, then lowest row, then lowest column.
#pragma once

#include "sw_packed.hpp"

namespace dn {

// A 32-bit store of both halves of a best-row word (LDS on the device).
DN_HD void bh_store2(uint16_t* p, uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  *reinterpret_cast<uint32_t*>(p) = v;
#else
  p[0] = (uint16_t)v;
  p[1] = (uint16_t)(v >> 16);
#endif
}

// LDS words per wave of a split traceback fill's flag staging: four steps of every group and
// half of a segment, per lane.
DN_HD constexpr int split_stage_words(int W, int P) { return seg_groups(W, P) * 8 * 64; }

template <int W, int P, bool TB>
struct SegFill {
  static constexpr int S = seg_width(W, P);
  static constexpr int GS = seg_groups(W, P);
  // Rows per step of the step loop, as sw_pair: four for the traceback fill (its registers at
  // S = 33 fit two waves per SIMD), one for the score-only fill.
  static constexpr int U = TB ? 4 : 1;
  static constexpr int LW = W - (P - 1) * S;  // in-band columns of the last segment
  static_assert(P >= 2 && LW >= 1 && LW <= S, "split shape");

  pk16 H[S + 1], E[S + 1];
  uint32_t T[S + U - 1];
  BlockStreams<S, true> bs;
  uint32_t QC[2], TS[4];
  pk16 oe, ge;
  uint32_t tab0, tab1;
  uint32_t bits[8];
  int seg;
  bool last;
  // Carried from cell 0 to the rest of a row.
  pk16 f, hleft, hdiag, rowmax, rowmax_in;
  uint32_t acc12, acc34, qsel;
  // Exchanged: this row's cell 0 (for the left neighbour) and last cell (for the right one).
  pk16 x_h0, x_e0, x_hl, x_fl;
  pk16 best;
  uint32_t best_i;
  uint16_t* bh;
  int bh_stride;
  uint32_t* gran;   // this lane's granule column of the slot
  uint32_t* stage;  // this lane's LDS words for one step's flags: ((g * 2 + h) * 4 + r) * stride
  int stage_stride;

  // ta / tb: the pair's tasks (an invalid half is emptied). slot: the fill group's slot; wave:
  // this lane's wave within the group (0 .. P-1); lane: its lane in the wave (pair
  // wave * 64/P + lane / P, segment lane % P). bh: this lane's
  // best-row words, column c at bh[2 * c * stride] (A) and the half-word after it (B).
  DN_HD void init(const SwArgs& a, SwTask ta, SwTask tb, bool va, bool vb, int seg_, uint64_t* slot, int wave,
                  int lane, uint16_t* bh_, int stride, uint32_t* stage_ = nullptr, int stage_stride_ = 0) {
    if (!va) ta.qlen = ta.tlo = ta.thi = 0, ta.flags = 0;
    if (!vb) tb.qlen = tb.tlo = tb.thi = 0, tb.flags = 0;
    seg = seg_;
    last = seg == P - 1;
    // Step s holds row s - seg: the query moves seg rows later (rows below seg read as N) and
    // column c at step s is target position s + seg*(S-1) + c.
    const int off = seg * (S - 1);
    SwTask* tk[2] = {&ta, &tb};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      tk[h]->t += (uint64_t)(int64_t)off;
      tk[h]->tlo -= off;
      tk[h]->thi -= off;
      if (!(tk[h]->flags & 1)) tk[h]->q -= (uint64_t)seg;
      tk[h]->qlen += seg;
    }
    const pk16 zero = pk_splat(0);
    oe = pk_splat(a.sp.gap_open + a.sp.gap_ext);
    ge = pk_splat(a.sp.gap_ext);
    tab0 = a.tab0;
    tab1 = a.tab1;
#pragma unroll
    for (int c = 0; c <= S; ++c) H[c] = E[c] = zero;
#pragma unroll
    for (int c0 = 0; c0 < S; c0 += 4) {
      const uint32_t wa = tword4(ta, c0), wb = tword4(tb, c0);
#pragma unroll
      for (int u = 0; u < 4 && c0 + u < S; ++u) {
        const int j = c0 + u;
        T[j] = tsel_from(byte_at(wa, u), j >= ta.tlo && j < ta.thi) |
               tsel_from(byte_at(wb, u), j >= tb.tlo && j < tb.thi) << 16;
      }
    }
    bs.init(ta, tb, seg);
    bs.next(0, QC, TS);
    x_h0 = x_e0 = x_hl = x_fl = zero;
    best = zero;
    best_i = 0xffffffffu;
    bh = bh_;
    bh_stride = stride;
    if (TB) {
#pragma unroll
      for (int c = 0; c < S; ++c) bh_store2(bh + 2 * c * bh_stride, 0);
      gran = reinterpret_cast<uint32_t*>(slot) + (uint64_t)(seg * 64 + wave * (64 / P) + lane / P) * 4;
      stage = stage_;
      stage_stride = stage_stride_;
    }
    const uint32_t zs = (uint32_t)a.sp.bw >> 16;  // 0 (bw <= 1000): keeps the masks in SGPRs
#pragma unroll
    for (int k = 0; k < 8; ++k) bits[k] = 0x01010101u << (k + zs);
  }

  DN_HD void step_begin() {
#pragma unroll
    for (int v = 0; v + 1 < U; ++v) T[S + v] = TS[v];
  }

  // Cell c of the row of step s (u: the row within the step; hup / eup the cell above-right).
  DN_HD void cell(int u, int c, int s) {
    const pk16 zero = pk_splat(0);
    const pk16 hup = H[c + 1], eup = E[c + 1];
    const pk16 sc = pk_from(perm_b32(tab0, tab1, T[c + u] ^ qsel));
    const pk16 hd = pk_add(hdiag, sc);
    const pk16 eo = pk_subsat(hup, oe), ee = pk_subsat(eup, ge);
    const pk16 e = pk_max(eo, ee);
    const pk16 fo = pk_subsat(hleft, oe), fe = pk_subsat(f, ge);
    f = pk_max(fo, fe);
    const pk16 h = pk_max(pk_max(hd, e), f);
    const bool edge = LW < S && c == LW;  // the last segment's out-of-band column
    H[c] = edge && last ? zero : h;
    E[c] = edge && last ? zero : e;
    if (TB) {
      const uint32_t p12 = perm_b32(pk_bits(pk_sub(eo, ee)), pk_bits(pk_sub(fo, fe)), 0x0B0A0908u);
      const uint32_t p34 = perm_b32(pk_bits(pk_sub(hd, h)), pk_bits(pk_sub(e, h)), 0x0B0A0908u);
      const uint32_t bit = bits[c & 7];
      acc12 = and_or(p12, bit, acc12);
      acc34 = and_or(p34, bit, acc34);
      if ((c & 7) == 7 || c == S - 1) {
        stage[(((c >> 3) * 2) * 4 + u) * stage_stride] = perm_b32(acc34, acc12, 0x06040200u);
        stage[(((c >> 3) * 2 + 1) * 4 + u) * stage_stride] = perm_b32(acc34, acc12, 0x07050301u);
#if defined(__HIP_DEVICE_COMPILE__)
        __builtin_amdgcn_sched_barrier(0);  // as in sw_pair: keeps the stores from sinking into later rows
#endif
        acc12 = acc34 = 0;
      }
    }
    if (LW < S && c == LW) rowmax_in = rowmax;  // the maximum over the in-band columns
    rowmax = c == 0 ? h : pk_max(rowmax, h);
    hleft = h;
    hdiag = hup;
  }

  // Cell 0 of the row of step s. hin / fin: H and F of the cell to its left (0 for segment 0).
  // qrow: the row's index within its block of four.
  DN_HD void cell0(int u, int qrow, int s, pk16 hin, pk16 fin) {
    qsel = qsel_row(QC, qrow);
    hleft = hin;
    f = fin;
    hdiag = H[0];
    acc12 = acc34 = 0;
    cell(u, 0, s);
    x_h0 = H[0];
    x_e0 = E[0];
  }

  // Cells 1 .. S-1 of the row of step s. hs / es: the cell above the right neighbour of the last
  // column (the next segment's cell 0 of the row above; 0 for the last segment).
  DN_HD void rest(int u, int s, pk16 hs, pk16 es) {
    H[S] = hs;
    E[S] = es;
#pragma unroll
    for (int c = 1; c < S; ++c) cell(u, c, s);
    x_hl = hleft;
    x_fl = f;
    pk16 rm = rowmax;
    if (LW < S) rm = pk_from(last ? pk_bits(rowmax_in) : pk_bits(rowmax));
    if (TB) {
      const uint32_t m = pk_signmask(pk_sub(best, rm));  // halves whose row maximum beats the best
      if (m) {
        if (m == 0xffffffffu) {
#pragma unroll
          for (int c = 0; c < S; ++c) bh_store2(bh + 2 * c * bh_stride, pk_bits(H[c]));
        } else if (m & 0xffffu) {
#pragma unroll
          for (int c = 0; c < S; ++c) bh[2 * c * bh_stride] = (uint16_t)pk_bits(H[c]);
        } else {
#pragma unroll
          for (int c = 0; c < S; ++c) bh[2 * c * bh_stride + 1] = (uint16_t)(pk_bits(H[c]) >> 16);
        }
        best_i = bfi(m, (uint32_t)(s - seg) * 0x10001u, best_i);
      }
    }
    best = pk_max(best, rm);
  }

  DN_HD void step_end(int s0) {
    if (TB) {
      // The four steps' flag words of every group and half, as one 16-byte granule store each.
      uint32_t* g0 = gran + (uint64_t)(s0 >> 2) * GS * 2 * 256 * P;
#pragma unroll
      for (int x = 0; x < 2 * GS; ++x) {
        Quad q;
        q.x = stage[(x * 4 + 0) * stage_stride];
        q.y = stage[(x * 4 + 1) * stage_stride];
        q.z = stage[(x * 4 + 2) * stage_stride];
        q.w = stage[(x * 4 + 3) * stage_stride];
        store_stream16(g0 + (uint64_t)x * 256 * P, q);
      }
    }
#pragma unroll
    for (int c = 0; c + 1 < S; ++c) T[c] = T[c + U];
    T[S - 1] = TS[U - 1];
#pragma unroll
    for (int v = 0; v + U < 4; ++v) TS[v] = TS[v + U];
    if (((s0 + U) & 3) == 0) bs.next(s0 + U, QC, TS);
  }

  // This lane's best cell per half as an ordering key: score, then lower row, then lower column
  // (score-only: the score).
  DN_HD void keys(uint64_t& ka, uint64_t& kb) const {
    const int sa = pk_lo(best), sb = pk_hi(best);
    if (!TB) {
      ka = (uint64_t)(uint32_t)sa;
      kb = (uint64_t)(uint32_t)sb;
      return;
    }
    int da = 0, db = 0;
#pragma unroll
    for (int c = S - 1; c >= 0; --c) {
      const bool in_band = !(last && c >= LW);
      if (in_band && (int16_t)bh[2 * c * bh_stride] == sa) da = c;
      if (in_band && (int16_t)bh[2 * c * bh_stride + 1] == sb) db = c;
    }
    const uint32_t ra = best_i & 0xffffu, rb = best_i >> 16;
    ka = (uint64_t)(uint32_t)sa << 32 | (uint64_t)(0xffffu - ra) << 16 | (uint64_t)(0xffffu - (uint32_t)(seg * S + da));
    kb = (uint64_t)(uint32_t)sb << 32 | (uint64_t)(0xffffu - rb) << 16 | (uint64_t)(0xffffu - (uint32_t)(seg * S + db));
  }
};

// The result of one half from the pair's combined key (as sw_pair writes it).
template <bool TB>
DN_HD void split_result(SwResult& r, uint64_t key) {
  if (!TB) {
    r.score = (int)(uint32_t)key;
    r.qb = r.qe = r.tb = r.te = r.nm = r.n_cigar = r.status = 0;
    return;
  }
  r.score = (int)(key >> 32);
  r.qe = (int)(int16_t)(0xffffu - (uint32_t)((key >> 16) & 0xffffu));
  r.te = (int)(0xffffu - (uint32_t)(key & 0xffffu));
  r.qb = r.tb = r.nm = r.n_cigar = r.status = 0;
}

#if defined(__HIP_DEVICE_COMPILE__)
// Value of the lane below / above within the wave's rows of 16 lanes (groups of P lanes never
// straddle a row: P divides 16); the caller replaces what crosses a pair's edge.
DN_DEV uint32_t from_lane_below(uint32_t v) { return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x111, 0xf, 0xf, false); }
DN_DEV uint32_t from_lane_above(uint32_t v) { return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x101, 0xf, 0xf, false); }
#define DN_SPLIT_LANES 1
#elif defined(DN_EMU)
// The CPU run of the whole library: a wave's threads are fibers with wave collectives.
inline uint32_t from_lane_below(uint32_t v) { return __shfl_up(v, 1u); }
inline uint32_t from_lane_above(uint32_t v) {
  const int l = (int)(threadIdx.x & 63);
  return __shfl(v, l < 63 ? l + 1 : l);
}
#define DN_SPLIT_LANES 1
#endif

#if defined(__HIPCC__) || defined(DN_EMU)
// One lane of the split fill on the device: tasks ida / idb of the pair, qmax the wave's longest
// query; wave / lane as in SegFill::init. Every lane of the wave runs the same steps. (The host
// pass of hipcc compiles it without the exchanges; only the device pass runs it.)
template <int W, int P, bool TB>
__device__ void sw_split(const SwArgs& a, uint32_t ida, uint32_t idb, bool va, bool vb, int qmax, uint64_t* slot,
                         int wave, int lane, uint16_t* bh, int stride, uint32_t* stage, int stage_stride) {
  using F = SegFill<W, P, TB>;
  constexpr int U = F::U;
  const int seg = lane % P;
  const bool first = seg == 0, last = seg == P - 1;
  F L;
  L.init(a, a.tasks[va ? ida : 0], a.tasks[vb ? idb : 0], va, vb, seg, slot, wave, lane, bh, stride, stage,
         stage_stride);
  const int steps = qmax + P - 1;
  for (int s0 = 0; s0 < steps; s0 += U) {
    L.step_begin();
#pragma unroll
    for (int u = 0; u < U; ++u) {
#if defined(DN_SPLIT_LANES)
      const uint32_t hl = from_lane_below(pk_bits(L.x_hl)), fl = from_lane_below(pk_bits(L.x_fl));
#else
      const uint32_t hl = 0, fl = 0;
#endif
      L.cell0(u, (s0 & 3) + u, s0 + u, pk_from(first ? 0u : hl), pk_from(first ? 0u : fl));
#if defined(DN_SPLIT_LANES)
      const uint32_t h0 = from_lane_above(pk_bits(L.x_h0)), e0 = from_lane_above(pk_bits(L.x_e0));
#else
      const uint32_t h0 = 0, e0 = 0;
#endif
      L.rest(u, s0 + u, pk_from(last ? 0u : h0), pk_from(last ? 0u : e0));
    }
    L.step_end(s0);
  }
  uint64_t ka, kb;
  L.keys(ka, kb);
#pragma unroll
  for (int o = 1; o < P; o <<= 1) {
    const uint64_t xa = __shfl_xor(ka, o), xb = __shfl_xor(kb, o);
    ka = xa > ka ? xa : ka;
    kb = xb > kb ? xb : kb;
  }
  if (first) {
    if (va) split_result<TB>(a.res[ida], ka);
    if (vb) split_result<TB>(a.res[idb], kb);
  }
}
#endif

}  // namespace dn
