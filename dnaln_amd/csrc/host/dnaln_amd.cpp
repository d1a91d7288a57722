// dnaln_amd.cpp — host classes of dnaln_amd.hpp over the C ABI.
#include "host/dnaln_amd.hpp"

#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <cstring>

namespace dnaln_amd {

namespace {

void check(int rc, const char* what) {
  if (rc != DNALN_OK) throw Error(std::string(what) + ": " + dnaln_hip_last_error());
}

dnaln_score_params to_c(const ScoreParams& s) {
  return {s.match, s.mismatch, s.gap_open, s.gap_ext, s.ambig, s.bw};
}

void append_uint(std::string& o, uint64_t v) {
  char tmp[24];
  int n = 0;
  do {
    tmp[n++] = (char)('0' + v % 10);
    v /= 10;
  } while (v);
  while (n) o += tmp[--n];
}

}  // namespace

dnaln_map_params to_c(const MapParams& p) {
  dnaln_map_params c;
  dnaln_hip_default_params(&c);
  c.k = p.k;
  c.w = p.w;
  c.max_occ = p.max_occ;
  c.sc = to_c(p.sc);
  c.ch = {p.ch.max_gap, p.ch.band, p.ch.lookback, p.ch.min_score, p.ch.min_anchors};
  c.max_cand = p.max_cand;
  c.pri_ratio = p.pri_ratio;
  c.min_aln_score = p.min_aln_score;
  return c;
}

// ------------------------------------------------------------------------------ HipAligner
HipAligner::HipAligner(const ScoreParams& sp, int device) : sp_(sp) {
  // An aligner needs no reference or index: the context holds a one-base reference and an
  // empty one-bucket table.
  MapParams mp;
  mp.sc = sp;
  const dnaln_map_params c = to_c(mp);
  static const uint8_t kN = 4;
  static const dnaln_bucket kEmpty = {~0ULL, 0, 0};
  const dnaln_contig contig = {0, 1, 0};
  dnaln_index_view view{};
  view.table = &kEmpty;
  view.n_buckets = 1;
  view.pos = nullptr;
  view.n_pos = 0;
  view.k = c.k;
  view.w = c.w;
  check(dnaln_hip_create(&c, &kN, 1, &contig, 1, &view, device, &ctx_), "dnaln_hip_create");
}

HipAligner::~HipAligner() { dnaln_hip_destroy(ctx_); }

void HipAligner::align(const AlignTask* tasks, size_t n, bool traceback, AlignResult* out) {
  pool_.clear();
  tasks_.resize(n);
  uint64_t cig_cap = 0;
  for (size_t i = 0; i < n; ++i) {
    const AlignTask& t = tasks[i];
    dnaln_task& d = tasks_[i];
    d.q_off = pool_.size();
    d.qlen = t.qlen;
    pool_.insert(pool_.end(), t.q, t.q + t.qlen);
    d.t_off = pool_.size();
    d.tlen = t.tlen;
    pool_.insert(pool_.end(), t.t, t.t + t.tlen);
    cig_cap += 2 * (uint64_t)std::max(t.qlen, 0) + 4;
  }
  res_.resize(n);
  cig_.resize(traceback ? cig_cap + 1 : 1);
  const dnaln_score_params sp = to_c(sp_);
  check(dnaln_hip_align(ctx_, &sp, tasks_.data(), n, pool_.data(), pool_.size(), traceback ? 1 : 0, res_.data(),
                        cig_.data(), cig_.size()),
        "dnaln_hip_align");
  for (size_t i = 0; i < n; ++i) {
    const dnaln_result& r = res_[i];
    AlignResult& o = out[i];
    o.score = r.score;
    o.qb = r.qb;
    o.qe = r.qe;
    o.tb = r.tb;
    o.te = r.te;
    o.nm = r.nm;
    o.cigar.clear();
    if (traceback) o.cigar.assign(cig_.begin() + r.cigar_off, cig_.begin() + r.cigar_off + r.cigar_len);
  }
}

// ------------------------------------------------------------------------------ SeqReader
SeqReader::SeqReader(const std::string& path) : win_(4u << 20) {
  gzFile f = path == "-" ? gzdopen(dup(fileno(stdin)), "rb") : gzopen(path.c_str(), "rb");
  if (!f) throw Error("cannot open '" + path + "'");
  gzbuffer(f, 1u << 20);
  fp_ = f;
}

SeqReader::~SeqReader() {
  if (fp_) gzclose(static_cast<gzFile>(fp_));
}

bool SeqReader::next(std::string_view& ln) {
  for (;;) {
    const char* p = win_.data() + head_;
    if (fastx::cut_line(p, win_.data() + tail_, eof_, ln)) {
      last_ = head_;
      head_ = (size_t)(p - win_.data());
      return true;
    }
    if (eof_) return false;
    // No whole line is left in the window: keep the partial line, read more behind it.
    std::memmove(win_.data(), win_.data() + head_, tail_ - head_);
    tail_ -= head_;
    head_ = last_ = 0;
    if (tail_ == win_.size()) win_.resize(win_.size() * 2);
    const int got = gzread(static_cast<gzFile>(fp_), win_.data() + tail_, (unsigned)(win_.size() - tail_));
    if (got < 0) throw Error("read error");
    if (got == 0) eof_ = true;
    tail_ += (size_t)got;
  }
}

bool SeqReader::next(SeqRecord& r) {
  fastx::RecordMachine& m = machine_;
  try {
    if (!fastx::read_record(*this, m)) return false;
  } catch (const fastx::RecordError& e) {
    throw Error(e.what());
  }
  r.name.swap(m.name);
  r.seq.swap(m.seq);
  r.qual.swap(m.qual);
  return true;
}

size_t SeqReader::next_batch(std::vector<SeqRecord>& out, size_t max) {
  if (out.size() < max) out.resize(max);
  size_t n = 0;
  while (n < max && next(out[n])) ++n;
  return n;
}

// ------------------------------------------------------------------------------ Reference
Reference Reference::load(const std::string& path) {
  Reference ref;
  SeqReader rd(path);
  SeqRecord rec;
  while (rd.next(rec)) {
    if (rec.seq.size() > UINT32_MAX) throw Error("contig '" + rec.name + "' is too long");
    const size_t off = ref.seq.size();
    ref.contigs.push_back({rec.name, off, (uint32_t)rec.seq.size()});
    ref.seq.resize(off + rec.seq.size());
    dnaln_encode(rec.seq.data(), rec.seq.size(), ref.seq.data() + off);
  }
  if (ref.contigs.empty()) throw Error("no sequences found in '" + path + "'");
  if (ref.seq.size() >= (1ull << 31)) throw Error("reference is too large (limit: 2^31 bp)");
  return ref;
}

int Reference::contig_of(uint64_t gpos) const {
  size_t lo = 0, hi = contigs.size();
  while (hi - lo > 1) {
    const size_t mid = (lo + hi) / 2;
    if (contigs[mid].offset <= gpos) lo = mid;
    else hi = mid;
  }
  return (int)lo;
}

// ------------------------------------------------------------------------------ MinimizerIndex
MinimizerIndex::~MinimizerIndex() {
  if (built_) dnaln_index_free(&idx_);
}

void MinimizerIndex::build(const Reference& ref, int k, int w) {
  if (built_) dnaln_index_free(&idx_);
  built_ = false;
  std::vector<dnaln_contig> cs;
  for (const Contig& c : ref.contigs) cs.push_back({c.offset, c.len, 0});
  if (dnaln_index_build(ref.seq.data(), cs.data(), (int32_t)cs.size(), k, w, &idx_) != DNALN_OK)
    throw Error("index build failed (k in [4, 28], w in [1, 255], reference < 2^31 bp)");
  built_ = true;
}

namespace {
std::vector<dnaln_contig> contig_table(const Reference& ref) {
  std::vector<dnaln_contig> cs;
  for (const Contig& c : ref.contigs) cs.push_back({c.offset, c.len, 0});
  return cs;
}
}  // namespace

double MinimizerIndex::build_gpu(const Reference& ref, int k, int w, int device) {
  if (built_) dnaln_index_free(&idx_);
  built_ = false;
  const std::vector<dnaln_contig> cs = contig_table(ref);
  double ms = 0;
  if (dnaln_index_build_gpu(ref.seq.data(), cs.data(), (int32_t)cs.size(), k, w, device, &idx_, &ms) != DNALN_OK)
    throw Error(std::string("GPU index build failed: ") + dnaln_hip_last_error());
  built_ = true;
  return ms;
}

void MinimizerIndex::save(const Reference& ref, const std::string& path) const {
  if (!built_) throw Error("no index to save");
  const std::vector<dnaln_contig> cs = contig_table(ref);
  if (dnaln_index_save(&idx_, ref.seq.data(), cs.data(), (int32_t)cs.size(), path.c_str()) != DNALN_OK)
    throw Error("cannot write index file '" + path + "'");
}

void MinimizerIndex::load(const Reference& ref, const std::string& path) {
  if (built_) dnaln_index_free(&idx_);
  built_ = false;
  const std::vector<dnaln_contig> cs = contig_table(ref);
  if (dnaln_index_load(path.c_str(), ref.seq.data(), cs.data(), (int32_t)cs.size(), &idx_) != DNALN_OK)
    throw Error("index file '" + path + "' is unreadable, corrupt, or was built from another reference");
  built_ = true;
}

const uint32_t* MinimizerIndex::lookup(uint64_t h, uint32_t& n) const {
  n = 0;
  if (!built_) return nullptr;
  const uint64_t mask = idx_.view.n_buckets - 1;
  for (uint64_t i = h & mask;; i = (i + 1) & mask) {
    const dnaln_bucket& b = idx_.view.table[i];
    if (b.key == h) {
      n = b.count;
      return idx_.view.pos + b.offset;
    }
    if (b.key == ~0ULL) return nullptr;
  }
}

// ------------------------------------------------------------------------------ Mapper
Mapper::Mapper(const Reference& ref, const MinimizerIndex& idx, const MapParams& p) : ref_(ref), p_(p) {
  p_.ch.band = std::min(p_.ch.band, p_.sc.bw);  // src/mapper.cpp:9
  std::vector<dnaln_contig> cs;
  for (const Contig& c : ref.contigs) cs.push_back({c.offset, c.len, 0});
  const dnaln_map_params c = to_c(p_);
  check(dnaln_hip_create(&c, ref.seq.data(), ref.seq.size(), cs.data(), (int32_t)cs.size(), &idx.view(), p.device,
                         &ctx_),
        "dnaln_hip_create");
  dnaln_batch_config cfg;
  dnaln_hip_default_batch_config(&cfg);
  cfg.mode = p.batched ? DNALN_MODE_BATCHED : DNALN_MODE_PLAIN;
  cfg.tune = p.tune ? 1 : 0;
  check(dnaln_hip_batch_create(ctx_, &cfg, &batch_), "dnaln_hip_batch_create");
}

Mapper::~Mapper() {
  dnaln_hip_batch_destroy(batch_);
  dnaln_hip_destroy(ctx_);
}

void Mapper::map(const SeqRecord* reads, size_t n, std::vector<MapRecord>& out) {
  codes_.clear();
  offs_.resize(n);
  lens_.resize(n);
  fnv_.resize(n);
  uint64_t cig_cap = 1;
  for (size_t i = 0; i < n; ++i) {
    const SeqRecord& r = reads[i];
    if (r.seq.size() > (size_t)INT32_MAX) throw Error("read '" + r.name + "' is too long");
    offs_[i] = codes_.size();
    lens_[i] = (int32_t)r.seq.size();
    codes_.resize(codes_.size() + r.seq.size());
    dnaln_encode(r.seq.data(), r.seq.size(), codes_.data() + offs_[i]);
    fnv_[i] = dnaln_fnv1a(r.name.data(), r.name.size());
    cig_cap += 2 * (uint64_t)r.seq.size() + 4;
  }
  rec_.resize(n);
  cig_.resize(cig_cap);
  dnaln_map_stats st{};
  check(dnaln_hip_batch_upload(batch_, codes_.data(), offs_.data(), lens_.data(), fnv_.data(), n),
        "dnaln_hip_batch_upload");
  check(dnaln_hip_batch_run(batch_), "dnaln_hip_batch_run");
  check(dnaln_hip_batch_download(batch_, rec_.data(), cig_.data(), cig_.size(), &st), "dnaln_hip_batch_download");
  out.resize(n);
  for (size_t i = 0; i < n; ++i) {
    const dnaln_record& r = rec_[i];
    MapRecord& m = out[i];
    m.contig = r.contig;
    m.pos = r.pos;
    m.rev = r.rev;
    m.mapq = r.mapq;
    m.score = r.score;
    m.nm = r.nm;
    m.n_cand = r.n_cand;
    m.cigar.assign(cig_.begin() + r.cigar_off, cig_.begin() + r.cigar_off + r.cigar_len);
  }
  stats_.reads += st.reads;
  stats_.mapped += st.mapped;
  stats_.anchors += st.anchors;
  stats_.candidates += st.candidates;
  stats_.cells += st.cells;
  stats_.t_seed += (st.ms_seed + st.ms_chain) * 1e-3;
  stats_.t_align += (st.ms_rank + st.ms_traceback) * 1e-3;
}

void Mapper::append_sam(const SeqRecord& r, const MapRecord& m, std::string& out) const {
  dnaln_record rec{};
  rec.contig = m.contig;
  rec.rev = m.rev;
  rec.mapq = m.mapq;
  rec.pos = m.pos;
  rec.score = m.score;
  rec.nm = m.nm;
  rec.n_cand = m.n_cand;
  rec.cigar_len = (uint32_t)m.cigar.size();
  rec.cigar_off = 0;
  const char* cname = m.mapped() ? ref_.contigs[m.contig].name.c_str() : "*";
  std::string line(r.name.size() + 2 * r.seq.size() + 16 * m.cigar.size() + strlen(cname) + 96, '\0');
  const long len = dnaln_sam_format(r.name.data(), r.name.size(), r.seq.data(), r.qual.empty() ? nullptr : r.qual.data(),
                                    r.seq.size(), &rec, m.cigar.data(), cname, &line[0], line.size());
  if (len < 0) throw Error("SAM line buffer too small");
  out.append(line.data(), (size_t)len);
}

void Mapper::write_sam_header(FILE* fp, const std::string& cmdline) const {
  std::string h = "@HD\tVN:1.6\tSO:unsorted\n";
  for (const Contig& c : ref_.contigs) {
    h += "@SQ\tSN:" + c.name + "\tLN:";
    append_uint(h, c.len);
    h += '\n';
  }
  h += "@PG\tID:dnaln\tPN:dnaln\tVN:0.1.0\tCL:" + cmdline + "\n";
  std::fwrite(h.data(), 1, h.size(), fp);
}

}  // namespace dnaln_amd
