// graph_emu_main.cpp — the library's batch pipeline under the CPU emulation of HIP
// (hip/hip_runtime.h beside it), built with AddressSanitizer: a driver for tests/test_emu_graph.py.
// Test infrastructure only.
//
//   graph_emu DIR MODE GRAPH SPLIT0 SPLIT1 ANCHORS_PER_READ DIR_BUDGET WALK LO:HI [LO:HI ...]
//
// MODE and GRAPH are dnaln_batch_config's mode (0 plain, 1 batched) and graph (0 launches one by
// one, 1 one captured graph per plan; this build defines DN_GRAPH_CAPTURE).
//
// DIR holds the inputs as raw little-endian arrays: genome.u8 (codes), contigs.u64 (offset, len
// pairs), codes.u8 / offs.u64 / lens.i32 / fnv.u32 (the reads), params.i32 (k, w, max_occ,
// match, mismatch, gap_open, gap_ext, bw, max_cand, min_chain, min_score). For each range, the
// reads [LO, HI) are uploaded, run and downloaded through one batch object; records and CIGARs
// go to DIR/rec_<i>.bin and DIR/cig_<i>.u32, the candidates to DIR/ncand_<i>.i32 and
// DIR/cand_<i>.bin. Prints one line of statistics per range.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "dnaln_hip.h"
#include "dnaln_host.h"

template <class T>
static std::vector<T> slurp(const std::string& path) {
  std::vector<T> v;
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    std::fprintf(stderr, "cannot open %s\n", path.c_str());
    std::exit(2);
  }
  std::fseek(f, 0, SEEK_END);
  const long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  v.resize((size_t)n / sizeof(T));
  if (n && std::fread(v.data(), 1, (size_t)n, f) != (size_t)n) std::exit(2);
  std::fclose(f);
  return v;
}

template <class T>
static void spill(const std::string& path, const T* p, size_t n) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f || (n && std::fwrite(p, sizeof(T), n, f) != n)) std::exit(2);
  std::fclose(f);
}

static void check(int rc, const char* what) {
  if (rc != DNALN_OK) {
    std::fprintf(stderr, "%s: %d %s\n", what, rc, dnaln_hip_last_error());
    std::exit(1);
  }
}

int main(int argc, char** argv) {
  if (argc < 10) return 2;
  const std::string dir = argv[1];
  const auto genome = slurp<uint8_t>(dir + "/genome.u8");
  const auto cpairs = slurp<uint64_t>(dir + "/contigs.u64");
  const auto codes = slurp<uint8_t>(dir + "/codes.u8");
  const auto offs = slurp<uint64_t>(dir + "/offs.u64");
  const auto lens = slurp<int32_t>(dir + "/lens.i32");
  const auto fnv = slurp<uint32_t>(dir + "/fnv.u32");
  const auto prm = slurp<int32_t>(dir + "/params.i32");
  std::vector<dnaln_contig> contigs(cpairs.size() / 2);
  for (size_t i = 0; i < contigs.size(); ++i) contigs[i] = dnaln_contig{cpairs[2 * i], (uint32_t)cpairs[2 * i + 1], 0};
  dnaln_map_params p;
  dnaln_hip_default_params(&p);
  p.k = prm[0];
  p.w = prm[1];
  p.max_occ = (uint32_t)prm[2];
  p.sc.match = prm[3];
  p.sc.mismatch = prm[4];
  p.sc.gap_open = prm[5];
  p.sc.gap_ext = prm[6];
  p.sc.bw = prm[7];
  p.max_cand = prm[8];
  p.ch.min_score = prm[9];
  p.min_aln_score = prm[10];
  dnaln_index idx;
  check(dnaln_index_build(genome.data(), contigs.data(), (int)contigs.size(), p.k, p.w, &idx), "index");
  dnaln_hip_ctx* ctx = nullptr;
  check(dnaln_hip_create(&p, genome.data(), genome.size(), contigs.data(), (int)contigs.size(), &idx.view, 0, &ctx),
        "create");
  dnaln_batch_config cfg;
  dnaln_hip_default_batch_config(&cfg);
  cfg.mode = std::atoi(argv[2]);
  cfg.graph = std::atoi(argv[3]);
  cfg.len_split[0] = std::atoi(argv[4]);
  cfg.len_split[1] = std::atoi(argv[5]);
  cfg.anchors_per_read = (uint32_t)std::atoi(argv[6]);
  cfg.dir_budget = std::strtoull(argv[7], nullptr, 10);
  cfg.walk_kind = std::atoi(argv[8]);
  dnaln_hip_batch* b = nullptr;
  check(dnaln_hip_batch_create(ctx, &cfg, &b), "batch_create");
  const int mc = p.max_cand > 0 ? p.max_cand : 1;
  for (int a = 9; a < argc; ++a) {
    const std::string r = argv[a];
    const size_t colon = r.find(':');
    const uint64_t lo = std::strtoull(r.c_str(), nullptr, 10), hi = std::strtoull(r.c_str() + colon + 1, nullptr, 10);
    const uint64_t n = hi - lo;
    std::vector<uint64_t> o(offs.begin() + (long)lo, offs.begin() + (long)hi);
    const uint64_t base = n ? o[0] : 0;
    for (auto& x : o) x -= base;
    check(dnaln_hip_batch_upload(b, codes.data() + base, o.data(), lens.data() + lo, fnv.data() + lo, n), "upload");
    check(dnaln_hip_batch_run(b), "run");
    std::vector<dnaln_record> rec(n + 1);
    uint64_t cap = 4;
    for (uint64_t i = lo; i < hi; ++i) cap += 2 * (uint64_t)lens[i] + 4;
    std::vector<uint32_t> cig(cap);
    dnaln_map_stats st;
    check(dnaln_hip_batch_download(b, rec.data(), cig.data(), cap, &st), "download");
    std::vector<int32_t> nc(n + 1);
    std::vector<dnaln_candidate> cd((n + 1) * mc);
    if (cfg.mode == DNALN_MODE_BATCHED || n <= cfg.chunk_reads)
      check(dnaln_hip_batch_candidates(b, nc.data(), cd.data(), (uint64_t)mc), "candidates");
    uint64_t used = 0;
    for (uint64_t i = 0; i < n; ++i) used = std::max<uint64_t>(used, rec[i].cigar_off + rec[i].cigar_len);
    const std::string tag = std::to_string(a - 9);
    spill(dir + "/rec_" + tag + ".bin", rec.data(), n);
    spill(dir + "/cig_" + tag + ".u32", cig.data(), used);
    spill(dir + "/ncand_" + tag + ".i32", nc.data(), n);
    spill(dir + "/cand_" + tag + ".bin", cd.data(), n * mc);
    std::printf("range %s reads %llu global %llu retry %llu anchors %llu rank %llu tb %llu\n", r.c_str(),
                (unsigned long long)st.reads, (unsigned long long)st.global_path_reads,
                (unsigned long long)st.fused_retry_reads, (unsigned long long)st.anchors,
                (unsigned long long)st.rank_tasks, (unsigned long long)st.tb_tasks);
    std::fflush(stdout);
  }
  dnaln_hip_batch_destroy(b);
  dnaln_hip_destroy(ctx);
  dnaln_index_free(&idx);
  return 0;
}
