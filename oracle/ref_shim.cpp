// ref_shim.cpp — C entry points over the reference dnaln classes (test infrastructure only).
//
// Compiled by oracle/ref.mk against the reference's own headers and sources into
// oracle/_ref/libdnaln_ref.so. Tests use it to pin the oracle on random inputs; bench.py may
// time it as the CPU baseline. Nothing in the product links it.
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "align.hpp"
#include "chain.hpp"
#include "index.hpp"
#include "mapper.hpp"
#include "minimizer.hpp"
#include "reference.hpp"

using namespace dnaln;

namespace {
struct RefCtx {
  Reference ref;
  MinimizerIndex idx;
};

ScoreParams score_params(const int* s) {
  ScoreParams sp;
  sp.match = s[0];
  sp.mismatch = s[1];
  sp.gap_open = s[2];
  sp.gap_ext = s[3];
  sp.ambig = s[4];
  sp.bw = s[5];
  return sp;
}
}  // namespace

extern "C" {

// Aligns n tasks; res6[i*6..] = score, qb, qe, tb, te, nm. CIGARs are written to cig at
// cig_off[i] with cig_n[i] ops. Returns 0, or -1 when cap is too small.
int ref_align(const int* sp6, int simd, const uint8_t* pool, const int64_t* q_off, const int32_t* q_len,
              const int64_t* t_off, const int32_t* t_len, size_t n, int traceback, int32_t* res6, uint32_t* cig,
              int64_t* cig_off, int32_t* cig_n, size_t cap) {
  const ScoreParams sp = score_params(sp6);
  std::unique_ptr<Aligner> al;
  if (simd) al = std::make_unique<SimdAligner>(sp);
  else al = std::make_unique<ScalarAligner>(sp);
  std::vector<AlignTask> tasks(n);
  for (size_t i = 0; i < n; ++i) tasks[i] = {pool + q_off[i], q_len[i], pool + t_off[i], t_len[i]};
  std::vector<AlignResult> out(n);
  al->align(tasks.data(), n, traceback != 0, out.data());
  size_t used = 0;
  for (size_t i = 0; i < n; ++i) {
    const AlignResult& r = out[i];
    int32_t* o = res6 + 6 * i;
    o[0] = r.score;
    o[1] = r.qb;
    o[2] = r.qe;
    o[3] = r.tb;
    o[4] = r.te;
    o[5] = r.nm;
    if (used + r.cigar.size() > cap) return -1;
    cig_off[i] = static_cast<int64_t>(used);
    cig_n[i] = static_cast<int32_t>(r.cigar.size());
    std::memcpy(cig + used, r.cigar.data(), r.cigar.size() * sizeof(uint32_t));
    used += r.cigar.size();
  }
  return 0;
}

size_t ref_sketch(const uint8_t* seq, uint32_t len, int k, int w, uint32_t offset, uint64_t* hash, uint32_t* pos,
                  uint32_t* strand) {
  std::vector<Minimizer> v;
  sketch(seq, len, k, w, offset, v);
  for (size_t i = 0; i < v.size(); ++i) {
    hash[i] = v[i].hash;
    pos[i] = v[i].pos;
    strand[i] = v[i].strand;
  }
  return v.size();
}

// cp5 = max_gap, band, lookback, min_score, min_anchors. Anchors are sorted in place.
// out7 per chain: score, strand, n_anchors, dmin(lo32), dmin(hi32), dmax(lo32), dmax(hi32); out_rpos.
size_t ref_chain(uint32_t* rpos, int32_t* qpos, uint32_t* strand, size_t n, int k, const int* cp5, int64_t* out6) {
  std::vector<Anchor> a(n);
  for (size_t i = 0; i < n; ++i) a[i] = {rpos[i], qpos[i], strand[i]};
  ChainParams p;
  p.max_gap = cp5[0];
  p.band = cp5[1];
  p.lookback = cp5[2];
  p.min_score = cp5[3];
  p.min_anchors = cp5[4];
  Chainer ch;
  std::vector<Chain> out;
  ch.run(a, k, p, out);
  for (size_t i = 0; i < n; ++i) {
    rpos[i] = a[i].rpos;
    qpos[i] = a[i].qpos;
    strand[i] = a[i].strand;
  }
  for (size_t i = 0; i < out.size(); ++i) {
    int64_t* o = out6 + 6 * i;
    o[0] = out[i].score;
    o[1] = out[i].strand;
    o[2] = out[i].n_anchors;
    o[3] = out[i].dmin;
    o[4] = out[i].dmax;
    o[5] = out[i].rpos;
  }
  return out.size();
}

void* ref_ctx_create(const uint8_t* codes, const uint64_t* contig_off, const uint32_t* contig_len, int n_contigs,
                     int k, int w) {
  RefCtx* c = new RefCtx;
  uint64_t total = 0;
  for (int i = 0; i < n_contigs; ++i) {
    c->ref.contigs.push_back({"chr" + std::to_string(i + 1), contig_off[i], contig_len[i]});
    total = contig_off[i] + contig_len[i];
  }
  c->ref.seq.assign(codes, codes + total);
  c->idx.build(c->ref, k, w);
  return c;
}

void ref_ctx_destroy(void* ctx) { delete static_cast<RefCtx*>(ctx); }

// Copies the occurrence list of hash h (at most cap entries); returns n.
uint32_t ref_index_lookup(void* ctx, uint64_t h, uint32_t* out, uint32_t cap) {
  uint32_t n = 0;
  const uint32_t* p = static_cast<RefCtx*>(ctx)->idx.lookup(h, n);
  for (uint32_t i = 0; i < n && i < cap; ++i) out[i] = p[i];
  return n;
}

uint64_t ref_index_stats(void* ctx, uint64_t* n_positions) {
  *n_positions = static_cast<RefCtx*>(ctx)->idx.n_positions();
  return static_cast<RefCtx*>(ctx)->idx.n_keys();
}

// mp = k, w, max_occ, A, B, O, E, bw, max_cand, min_chain, min_score, simd.
// rec8 per read: contig, pos(lo32), pos(hi32), rev, mapq, score, nm, n_cand. CIGARs as in ref_align.
int ref_map(void* ctx, const int* mp, const char* seq, const uint64_t* seq_off, const int32_t* seq_len,
            const char* names, const uint64_t* name_off, const int32_t* name_len, size_t n, int32_t* rec8,
            uint32_t* cig, int64_t* cig_off, int32_t* cig_n, size_t cap, uint64_t* cells) {
  RefCtx* c = static_cast<RefCtx*>(ctx);
  MapParams p;
  p.k = mp[0];
  p.w = mp[1];
  p.max_occ = static_cast<uint32_t>(mp[2]);
  p.sc.match = mp[3];
  p.sc.mismatch = mp[4];
  p.sc.gap_open = mp[5];
  p.sc.gap_ext = mp[6];
  p.sc.bw = mp[7];
  p.max_cand = mp[8];
  p.ch.min_score = mp[9];
  p.min_aln_score = mp[10];
  p.simd = mp[11] != 0;
  Mapper m(c->ref, c->idx, p);
  std::vector<SeqRecord> reads(n);
  for (size_t i = 0; i < n; ++i) {
    reads[i].name.assign(names + name_off[i], static_cast<size_t>(name_len[i]));
    reads[i].seq.assign(seq + seq_off[i], static_cast<size_t>(seq_len[i]));
    reads[i].qual.assign(static_cast<size_t>(seq_len[i]), 'I');
  }
  std::vector<MapRecord> recs;
  m.map(reads.data(), n, recs);
  size_t used = 0;
  for (size_t i = 0; i < n; ++i) {
    const MapRecord& r = recs[i];
    int32_t* o = rec8 + 8 * i;
    o[0] = r.contig;
    o[1] = static_cast<int32_t>(r.pos & 0xffffffff);
    o[2] = static_cast<int32_t>(r.pos >> 32);
    o[3] = r.rev;
    o[4] = r.mapq;
    o[5] = r.score;
    o[6] = r.nm;
    o[7] = r.n_cand;
    if (used + r.cigar.size() > cap) return -1;
    cig_off[i] = static_cast<int64_t>(used);
    cig_n[i] = static_cast<int32_t>(r.cigar.size());
    std::memcpy(cig + used, r.cigar.data(), r.cigar.size() * sizeof(uint32_t));
    used += r.cigar.size();
  }
  if (cells) *cells = m.stats().cells;
  return 0;
}

}  // extern "C"
