// emu.cpp — CPU build of the device lane programs (test infrastructure).
//
// Runs sw_pair / sw_generic_one exactly as a wave64 would, lane by lane, with the same
// task pairing (lane l holds tasks l and l + 64 of each 128-task group), the same
// direction-slot layout and the same dispatch rules, so the kernels' integer tricks can be
// checked against the oracle and the reference vectors without a GPU.
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "device/bw_list.hpp"
#include "device/sw_generic.hpp"
#include "device/sw_packed.hpp"

using namespace dn;

namespace {

template <int W, bool TB>
void run_packed(const SwArgs& a) {
  constexpr int G = (W + 7) / 8;
  for (uint32_t base = 0; base < a.n; base += 128) {
    int qmax = 0;
    for (uint32_t k = base; k < a.n && k < base + 128; ++k) qmax = std::max(qmax, a.tasks[k].qlen);
    std::vector<uint64_t> slot((size_t)std::max(qmax, 1) * G * 64, 0xdeadbeefdeadbeefULL);
    for (int lane = 0; lane < 64; ++lane) {
      const uint32_t ia = base + lane, ib = base + 64 + lane;
      sw_pair<W, TB>(a, ia, ib, ia < a.n, ib < a.n, qmax, slot.data(), lane);
    }
  }
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

}  // namespace

extern "C" {

// Same contract as the reference shim's ref_align (res6 = score qb qe tb te nm), CIGAR in
// forward order. force: 0 auto, 1 generic int32 kernel, 2 packed (fails if unavailable).
int emu_align(const int* sp6, const uint8_t* pool, const int64_t* q_off, const int32_t* q_len, const int64_t* t_off,
              const int32_t* t_len, size_t n, int traceback, int32_t* res6, uint32_t* cig, int64_t* cig_off,
              int32_t* cig_n, size_t cap, int force) {
  ScoreP sp{sp6[0], sp6[1], sp6[2], sp6[3], sp6[4], sp6[5]};
  std::vector<SwTask> tasks(n);
  int maxq = 0;
  for (size_t i = 0; i < n; ++i) {
    SwTask& t = tasks[i];
    t.q = (uint64_t)(uintptr_t)(pool + q_off[i]);
    t.t = (uint64_t)(uintptr_t)(pool + t_off[i]);
    t.qlen = q_len[i];
    t.tlo = 0;
    t.thi = std::min(t_len[i], q_len[i] + 2 * sp.bw);
    t.flags = 0;
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

}  // extern "C"
