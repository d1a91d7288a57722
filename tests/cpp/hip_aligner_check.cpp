// hip_aligner_check.cpp — drives dnaln_amd::HipAligner (the Aligner interface of src/align.hpp:62-68)
// from a task file and prints the results, for tests/test_gpu_cli.py.
//
// Input (binary, little-endian): i32 scoring[6] (A, B, O, E, ambig, bw), u64 n, u64 pool_len,
// pool bytes, then n x {u64 q_off, u64 t_off, i32 qlen, i32 tlen}. Output: one line per task,
// "score qb qe tb te nm cigar..." for traceback 1, then the same for traceback 0.
#include <cstdio>
#include <vector>

#include "host/dnaln_amd.hpp"

int main(int argc, char** argv) {
  if (argc != 2) return 2;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int32_t s[6];
  uint64_t n = 0, len = 0;
  if (std::fread(s, 4, 6, f) != 6 || std::fread(&n, 8, 1, f) != 1 || std::fread(&len, 8, 1, f) != 1) return 2;
  std::vector<uint8_t> pool(len);
  if (len && std::fread(pool.data(), 1, len, f) != len) return 2;
  std::vector<dnaln_task> raw(n);
  if (n && std::fread(raw.data(), sizeof(dnaln_task), n, f) != n) return 2;
  std::fclose(f);
  dnaln_amd::ScoreParams sp;
  sp.match = s[0];
  sp.mismatch = s[1];
  sp.gap_open = s[2];
  sp.gap_ext = s[3];
  sp.ambig = s[4];
  sp.bw = s[5];
  std::vector<dnaln_amd::AlignTask> tasks(n);
  for (uint64_t i = 0; i < n; ++i) {
    tasks[i].q = pool.data() + raw[i].q_off;
    tasks[i].qlen = raw[i].qlen;
    tasks[i].t = pool.data() + raw[i].t_off;
    tasks[i].tlen = raw[i].tlen;
  }
  try {
    dnaln_amd::HipAligner al(sp);
    std::vector<dnaln_amd::AlignResult> out(n);
    for (int tb = 1; tb >= 0; --tb) {
      al.align(tasks.data(), n, tb != 0, out.data());
      for (const auto& r : out) {
        std::printf("%d %d %d %d %d %d", r.score, r.qb, r.qe, r.tb, r.te, r.nm);
        for (uint32_t c : r.cigar) std::printf(" %u", c);
        std::printf("\n");
      }
    }
    std::printf("name %s\n", al.name().c_str());
  } catch (const dnaln_amd::Error& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
