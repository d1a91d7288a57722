// dnaln_amd.hpp — the reference's C++ mapping interface, served by the MI355X path.
//
// Same class and member names, argument meanings and results as dnaln's own headers, so code
// written against them (src/main.cpp:12-75, src/bench.cpp:160-199) runs unchanged:
//   ScoreParams, AlignTask, AlignResult, Aligner        src/align.hpp:34-68
//   ChainParams                                         src/chain.hpp:30-36
//   Contig, Reference                                   src/reference.hpp:10-30
//   MinimizerIndex                                      src/index.hpp:19-60
//   SeqRecord, SeqReader                                src/seqio.hpp:12-45
//   MapParams, MapRecord, MapStats, Mapper              src/mapper.hpp:25-94
// Everything on the hot path goes through the C ABI in dnaln_hip.h; this layer owns host
// buffers, encodes reads, hashes names and formats SAM. Errors throw dnaln_amd::Error (the
// reference calls die(); the CLI turns the exception into the same exit).
#pragma once

#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include "dnaln_hip.h"
#include "dnaln_host.h"
#include "host/fastx_record.hpp"

namespace dnaln_amd {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct ScoreParams {
  int match = 2, mismatch = 4, gap_open = 4, gap_ext = 2, ambig = 1;
  int bw = 16;
  int band_cols() const { return 2 * bw + 1; }
  int sub(uint8_t a, uint8_t b) const { return (a > 3 || b > 3) ? -ambig : (a == b ? match : -mismatch); }
};

struct AlignTask {
  const uint8_t* q = nullptr;  // query codes
  int32_t qlen = 0;
  const uint8_t* t = nullptr;  // target window codes (qlen + 2*bw; shorter is N-padded)
  int32_t tlen = 0;
};

struct AlignResult {
  int32_t score = 0;
  int32_t qb = 0, qe = 0;
  int32_t tb = 0, te = 0;
  int32_t nm = 0;
  std::vector<uint32_t> cigar;  // M/I/D, len << 4 | op
  void clear() {
    score = qb = qe = tb = te = nm = 0;
    cigar.clear();
  }
};

class Aligner {
 public:
  virtual ~Aligner() = default;
  virtual void align(const AlignTask* tasks, size_t n, bool traceback, AlignResult* out) = 0;
  virtual std::string name() const = 0;
};

// Aligner on the GPU: every task's query and target are copied into one pool and aligned by
// the packed int16 kernels (dnaln_hip_align).
class HipAligner final : public Aligner {
 public:
  explicit HipAligner(const ScoreParams& sp, int device = 0);
  ~HipAligner() override;
  HipAligner(const HipAligner&) = delete;
  HipAligner& operator=(const HipAligner&) = delete;
  void align(const AlignTask* tasks, size_t n, bool traceback, AlignResult* out) override;
  std::string name() const override { return "hip-gfx950"; }

 private:
  ScoreParams sp_;
  dnaln_hip_ctx* ctx_ = nullptr;
  std::vector<uint8_t> pool_;
  std::vector<dnaln_task> tasks_;
  std::vector<dnaln_result> res_;
  std::vector<uint32_t> cig_;
};

struct ChainParams {
  int max_gap = 5000;
  int band = 16;
  int lookback = 50;
  int min_score = 20;
  int min_anchors = 2;
};

struct SeqRecord {
  std::string name, seq, qual;  // qual is empty for FASTA input
};

// FASTA/FASTQ reader over plain or gzipped files ("-" reads stdin), with the reference's
// record rules (src/seqio.cpp:59-87; the rule list is in host/fastx_record.hpp): the name ends
// at the first space or tab, sequence lines are joined, a FASTQ quality is read until it is as
// long as the sequence. Whole lines are cut from a growing window of decompressed bytes and fed
// to the shared record state machine.
class SeqReader {
 public:
  explicit SeqReader(const std::string& path);
  ~SeqReader();
  SeqReader(const SeqReader&) = delete;
  SeqReader& operator=(const SeqReader&) = delete;
  bool next(SeqRecord& r);
  size_t next_batch(std::vector<SeqRecord>& out, size_t max);

  // The line source of the record machine: next() cuts the next line out of the window,
  // unread() puts the last one back.
  bool next(std::string_view& ln);
  void unread() { head_ = last_; }

 private:
  void* fp_ = nullptr;
  std::vector<char> win_;          // decompressed bytes; [head_, tail_) not yet cut into lines
  size_t head_ = 0, tail_ = 0, last_ = 0;
  bool eof_ = false;
  fastx::RecordMachine machine_;
};

struct Contig {
  std::string name;
  uint64_t offset = 0;
  uint32_t len = 0;
};

class Reference {
 public:
  std::vector<Contig> contigs;
  std::vector<uint8_t> seq;  // concatenated codes (A=0 C=1 G=2 T=3 N=4)
  static Reference load(const std::string& path);
  int contig_of(uint64_t gpos) const;
};

class MinimizerIndex {
 public:
  MinimizerIndex() = default;
  ~MinimizerIndex();
  MinimizerIndex(const MinimizerIndex&) = delete;
  MinimizerIndex& operator=(const MinimizerIndex&) = delete;
  void build(const Reference& ref, int k, int w);
  // The same index built on HIP device `device` (k <= 16) and downloaded; returns the device
  // build's wall time in ms.
  double build_gpu(const Reference& ref, int k, int w, int device);
  // The index file of include/dnaln_host.h, checked against `ref` on load.
  void save(const Reference& ref, const std::string& path) const;
  void load(const Reference& ref, const std::string& path);
  // Occurrences (global_pos << 1 | strand) of hash h, ascending; n = 0 when absent.
  const uint32_t* lookup(uint64_t h, uint32_t& n) const;
  int k() const { return idx_.view.k; }
  int w() const { return idx_.view.w; }
  size_t n_keys() const { return idx_.n_keys; }
  size_t n_positions() const { return idx_.view.n_pos; }
  size_t memory_bytes() const { return idx_.view.n_buckets * sizeof(dnaln_bucket) + idx_.view.n_pos * 4; }
  const dnaln_index_view& view() const { return idx_.view; }

 private:
  dnaln_index idx_{};
  bool built_ = false;
};

struct MapParams {
  int k = 15, w = 10;
  uint32_t max_occ = 500;
  ScoreParams sc;
  ChainParams ch;
  int max_cand = 10;
  double pri_ratio = 0.5;
  int min_aln_score = 40;
  bool simd = true;      // no meaning on the GPU (kept for source compatibility)
  bool prefetch = true;  // idem
  // GPU-only: batched form (length buckets, wave chaining) or plain per-chunk launches, and
  // whether the first batch runs the autotuner.
  bool batched = true;
  bool tune = false;
  int device = 0;
};

struct MapRecord {
  int32_t contig = -1;
  int64_t pos = 0;
  uint8_t rev = 0, mapq = 0;
  int32_t score = 0, nm = 0, n_cand = 0;
  std::vector<uint32_t> cigar;  // including soft clips
  bool mapped() const { return contig >= 0; }
};

struct MapStats {
  uint64_t reads = 0, mapped = 0, anchors = 0, candidates = 0, cells = 0;
  double t_seed = 0, t_align = 0;  // seconds (device time of the seed+chain and SW stages)
};

class Mapper {
 public:
  Mapper(const Reference& ref, const MinimizerIndex& idx, const MapParams& p);
  ~Mapper();
  Mapper(const Mapper&) = delete;
  Mapper& operator=(const Mapper&) = delete;
  void map(const SeqRecord* reads, size_t n, std::vector<MapRecord>& out);
  void append_sam(const SeqRecord& r, const MapRecord& m, std::string& out) const;
  void write_sam_header(FILE* fp, const std::string& cmdline) const;
  const MapStats& stats() const { return stats_; }
  std::string kernel_name() const { return "hip-gfx950"; }

  // The device context (for dnaln_map_file and other C-ABI calls on the same resident index).
  dnaln_hip_ctx* context() const { return ctx_; }

 private:
  const Reference& ref_;
  MapParams p_;
  dnaln_hip_ctx* ctx_ = nullptr;
  dnaln_hip_batch* batch_ = nullptr;
  MapStats stats_;
  std::vector<uint8_t> codes_;
  std::vector<uint64_t> offs_;
  std::vector<int32_t> lens_;
  std::vector<uint32_t> fnv_;
  std::vector<dnaln_record> rec_;
  std::vector<uint32_t> cig_;
};

// C-ABI parameter block from MapParams.
dnaln_map_params to_c(const MapParams& p);

}  // namespace dnaln_amd
