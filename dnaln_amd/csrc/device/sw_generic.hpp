// sw_generic.hpp — int32 banded Smith–Waterman, one alignment per thread, any band width.
//
// The catch-all path for tasks the packed int16 kernel does not take: band widths without a
// compiled specialisation, reads longer than 30000/match, or scores outside the byte table.
// It is the scalar recurrence of src/align.cpp:31-84 with the traceback of align.cpp:87-143.
// Per-thread H/E rows and one direction byte per cell live in device scratch.
#pragma once

#include "common.hpp"
#include "sw_packed.hpp"

namespace dn {

struct SwGenericArgs {
  const SwTask* tasks;
  const uint32_t* order;
  uint32_t n;
  SwResult* res;
  uint32_t* cigar;
  int32_t cigar_cap;
  int32_t* rows;          // per thread: 2 * (W + 1) ints
  uint8_t* dir;           // per thread: dir_stride bytes (>= qlen * W)
  uint64_t dir_stride;
  ScoreP sp;
  int traceback;
};

DN_HD void sw_generic_one(const SwGenericArgs& a, uint32_t id, uint32_t thread) {
  const SwTask tk = a.tasks[id];
  const ScoreP sp = a.sp;
  const int W = 2 * sp.bw + 1, ql = tk.qlen;
  const int oe = sp.gap_open + sp.gap_ext, ge = sp.gap_ext;
  int32_t* H = a.rows + (uint64_t)thread * (uint64_t)(2 * (W + 1));
  int32_t* E = H + (W + 1);
  uint8_t* D = a.dir + (uint64_t)thread * a.dir_stride;
  for (int d = 0; d <= W; ++d) H[d] = E[d] = 0;
  int best = 0, best_i = -1, best_d = -1;
  for (int i = 0; i < ql; ++i) {
    const uint8_t qc = q_at(tk, i);
    int f = 0, hleft = 0, hdiag = H[0];
    for (int d = 0; d < W; ++d) {
      const int hup = H[d + 1], eup = E[d + 1];
      const int hd = hdiag + sub_score(sp, qc, t_at(tk, i + d));
      const int eo = hup - oe, ee = eup - ge;
      const int e = eo > ee ? eo : ee;
      const int fo = hleft - oe, fe = f - ge;
      f = fo > fe ? fo : fe;
      int h = hd > e ? hd : e;
      h = h > 0 ? h : 0;
      h = h > f ? h : f;
      H[d] = h;
      E[d] = e;
      if (a.traceback) {
        D[(uint64_t)i * W + d] =
            (uint8_t)((h == hd ? 0x80 : 0) | (h == e ? 0x40 : 0) | (ee > eo ? 0x20 : 0) | (fe > fo ? 0x10 : 0));
        if (h > best) {
          best = h;
          best_i = i;
          best_d = d;
        }
      } else {
        best = h > best ? h : best;
      }
      hleft = h;
      hdiag = hup;
    }
  }
  SwResult& r = a.res[id];
  r.score = best;
  r.qb = r.qe = r.tb = r.te = r.nm = r.n_cigar = r.status = 0;
  if (!a.traceback || best <= 0) return;

  uint32_t* out = a.cigar + (uint64_t)id * (uint64_t)a.cigar_cap;
  Walk w = walk_init(best_i, best_d, best);
  while (!w.done) {
    if (w.state == 0 && (w.i < 0 || w.v <= 0)) {
      if (w.run) walk_push(w, 0, w.run, out, a.cigar_cap);
      walk_flush(w, out, a.cigar_cap);
      break;
    }
    if (w.i < 0 || w.d < 0 || w.d >= W) {
      w.status = 1;
      break;
    }
    const uint8_t x = D[(uint64_t)w.i * W + w.d];
    if (w.state == 0) {
      if (x & 0x80) {
        const uint8_t qa = q_at(tk, w.i), tc = t_at(tk, w.i + w.d);
        w.v -= sub_score(sp, qa, tc);
        w.nm += (qa != tc || qa > 3);
        ++w.run;
        --w.i;
        continue;
      }
      if (w.run) walk_push(w, 0, w.run, out, a.cigar_cap);
      w.run = 0;
      w.state = (x & 0x40) ? 1 : 2;
      continue;
    }
    const bool ext = w.state == 1 ? (x & 0x20) != 0 : (x & 0x10) != 0;
    w.v += ext ? ge : oe;
    ++w.run;
    if (w.state == 1) {
      --w.i;
      ++w.d;
    } else {
      --w.d;
    }
    if (!ext) {
      walk_push(w, (uint32_t)w.state, w.run, out, a.cigar_cap);
      w.nm += w.run;
      w.run = 0;
      w.state = 0;
    }
  }
  walk_finish(w, best_i, best_d, best, a.cigar_cap, r);
}

}  // namespace dn
