// evaluate.cpp — accuracy against simulated truth, on the device's records and alignments.
//
//   dnaln_evaluate     Evaluator::add (src/evaluate.cpp:40-84): a primary record is correct on
//                      the true contig and strand with its unclipped start within `tol` of the
//                      true start; base concordance over correct records; MAPQ histograms.
//   dnaln_truth_check  the true-locus checks of src/bench.cpp:272-314 (tasks built as
//                      src/groundtruth.cpp:11-44): the CIGAR re-scores to the score, the score
//                      is at least the true path's when that path stays in the band, and the
//                      read bases placed on their true target column.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dnaln_host.h"
#include "host/c_guard.hpp"

namespace {

enum : uint32_t { kM = 0, kI = 1, kD = 2, kN = 3, kS = 4, kH = 5, kEq = 7, kX = 8 };

// Reference position of every read base in CIGAR order; -1 for inserted or clipped bases
// (src/evaluate.cpp:40-54).
void base_positions(int64_t pos, const uint32_t* cig, uint32_t n, std::vector<int64_t>& out) {
  out.clear();
  for (uint32_t c = 0; c < n; ++c) {
    const uint32_t op = cig[c] & 0xf, len = cig[c] >> 4;
    if (op == kM || op == kEq || op == kX) {
      for (uint32_t k = 0; k < len; ++k) out.push_back(pos++);
    } else if (op == kI || op == kS) {
      out.insert(out.end(), len, -1);
    } else if (op == kD || op == kN) {
      pos += len;
    }
  }
}

int sub_score(const dnaln_score_params& sp, uint8_t a, uint8_t b) {  // src/align.hpp:39
  return (a > 3 || b > 3) ? -sp.ambig : (a == b ? sp.match : -sp.mismatch);
}

}  // namespace

extern "C" {

int dnaln_evaluate(const dnaln_record* recs, const uint32_t* cigar_pool, uint64_t n, const int32_t* truth_contig,
                   const int64_t* truth_start, const uint8_t* truth_rev, const uint32_t* truth_cigar,
                   const uint64_t* truth_cigar_off, int32_t tol, dnaln_eval_stats* out) {
  return dnaln_amd::c_guard([&]() -> int {
    if (!out || (n && (!recs || !truth_contig || !truth_start || !truth_rev || !truth_cigar_off))) return DNALN_E_INVALID;
    std::memset(out, 0, sizeof *out);
    std::vector<int64_t> tp, pp;
    for (uint64_t i = 0; i < n; ++i) {
      if (truth_start[i] < 0) continue;  // not in the truth set
      ++out->n_truth;
      ++out->records;
      const dnaln_record& m = recs[i];
      if (m.contig < 0 || m.cigar_len == 0) continue;
      ++out->mapped;
      const uint32_t* cig = cigar_pool + m.cigar_off;
      int64_t clip = 0;
      const uint32_t op0 = cig[0] & 0xf;
      if (op0 == kS || op0 == kH) clip = cig[0] >> 4;
      const bool ok = m.contig == truth_contig[i] && (m.rev != 0) == (truth_rev[i] != 0) &&
                      std::llabs(m.pos - clip - truth_start[i]) <= tol;
      const int q = std::clamp<int>(m.mapq, 0, 60);
      ++out->mq_mapped[q];
      if (!ok) {
        ++out->mq_wrong[q];
        continue;
      }
      ++out->correct;
      if (!truth_cigar) continue;
      base_positions(truth_start[i], truth_cigar + truth_cigar_off[i],
                     (uint32_t)(truth_cigar_off[i + 1] - truth_cigar_off[i]), tp);
      base_positions(m.pos, cig, m.cigar_len, pp);
      const size_t k = std::min(tp.size(), pp.size());
      for (size_t b = 0; b < k; ++b) {
        if (tp[b] < 0) continue;
        ++out->bases;
        out->bases_ok += pp[b] == tp[b];
      }
    }
    return DNALN_OK;
  });
}

int dnaln_truth_check(const dnaln_score_params* sp, const uint8_t* pool, const dnaln_task* tasks, uint64_t n,
                      const dnaln_result* res, const uint32_t* cigar_pool, const uint32_t* truth_cigar,
                      const uint64_t* truth_cigar_off, dnaln_truth_stats* out) {
  return dnaln_amd::c_guard([&]() -> int {
    if (!out || !sp || (n && (!pool || !tasks || !res || !truth_cigar || !truth_cigar_off))) return DNALN_E_INVALID;
    std::memset(out, 0, sizeof *out);
    const int W = 2 * sp->bw + 1;
    std::vector<int> tcols, pcols;
    for (uint64_t x = 0; x < n; ++x) {
      const dnaln_task& tk = tasks[x];
      const uint8_t* q = pool + tk.q_off;
      const uint8_t* t = pool + tk.t_off;
      const dnaln_result& r = res[x];
      const uint32_t* cig = cigar_pool + r.cigar_off;
      ++out->n;
      // score_cigar (src/align.cpp:145-162)
      int s = 0;
      for (uint32_t c = 0, i = (uint32_t)r.qb, j = (uint32_t)r.tb; c < r.cigar_len; ++c) {
        const uint32_t op = cig[c] & 0xf, len = cig[c] >> 4;
        if (op == kM || op == kEq || op == kX) {
          for (uint32_t k = 0; k < len; ++k) s += sub_score(*sp, q[i++], t[j++]);
        } else if (op == kI) {
          s -= sp->gap_open + (int)len * sp->gap_ext;
          i += len;
        } else if (op == kD) {
          s -= sp->gap_open + (int)len * sp->gap_ext;
          j += len;
        }
      }
      out->rescored_ok += s == r.score;
      // the true path from (0, bw) (src/bench.cpp:102-131)
      int i = 0, j = sp->bw, ts = 0;
      bool in_band = true;
      tcols.clear();
      for (uint64_t c = truth_cigar_off[x]; c < truth_cigar_off[x + 1]; ++c) {
        const uint32_t op = truth_cigar[c] & 0xf;
        const int len = (int)(truth_cigar[c] >> 4);
        if (op == kM) {
          for (int k = 0; k < len; ++k, ++i, ++j) {
            if (j - i < 0 || j - i >= W || i >= tk.qlen || j >= tk.tlen) {
              in_band = false;
              continue;
            }
            ts += sub_score(*sp, q[i], t[j]);
            tcols.push_back(j);
          }
        } else if (op == kI) {
          ts -= sp->gap_open + len * sp->gap_ext;
          tcols.insert(tcols.end(), (size_t)len, -1);
          i += len;
        } else if (op == kD) {
          ts -= sp->gap_open + len * sp->gap_ext;
          j += len;
        }
      }
      if (in_band) {
        ++out->in_band;
        out->optimal_ok += r.score >= ts;
      }
      // target column of each query base in the result (src/bench.cpp:133-147)
      pcols.assign((size_t)tk.qlen, -1);
      for (uint32_t c = 0, pi = (uint32_t)r.qb, pj = (uint32_t)r.tb; c < r.cigar_len; ++c) {
        const uint32_t op = cig[c] & 0xf, len = cig[c] >> 4;
        if (op == kM) {
          for (uint32_t k = 0; k < len && pi < (uint32_t)tk.qlen; ++k) pcols[pi++] = (int)pj++;
        } else if (op == kI) {
          pi += len;
        } else if (op == kD) {
          pj += len;
        }
      }
      uint64_t ok = 0, tot = 0;
      for (size_t b = 0; b < tcols.size() && b < pcols.size(); ++b) {
        if (tcols[b] < 0) continue;
        ++tot;
        ok += pcols[b] == tcols[b];
      }
      out->bases += tot;
      out->bases_ok += ok;
      out->majority_ok += tot && ok >= 0.95 * tot;
    }
    return DNALN_OK;
  });
}

}  // extern "C"
