// fastx_check — the parallel input stage of dnaln_map_file against the sequential reader.
//
// Usage: fastx_check <reads.fa|fq[.gz]> <batch_bytes> <threads>
//        fastx_check --dump <reads.fa|fq[.gz]>
// Reads the file with fastx::Batcher (batches of batch_bytes, `threads` parts) and with
// dnaln_amd::SeqReader (the reference's rules, src/seqio.cpp:59-87) and compares every
// record's name, sequence, quality and FASTQ flag. Prints "ok <records> parallel=<0|1>" and exits
// 0, or prints the first difference and exits 1. When the sequential reader fails, the batcher
// must fail with the same message (records of the failing batch are not delivered).
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "host/dnaln_amd.hpp"
#include "host/fastx_input.hpp"

// --dump <reads>: every record of dnaln_amd::SeqReader as "name TAB seq TAB qual" ('*' for an
// empty field, as the reference's SAM writes an unmapped read), then "ERROR <message>" if the
// reader failed.
static int dump(const char* path) {
  try {
    dnaln_amd::SeqReader rd(path);
    dnaln_amd::SeqRecord r;
    while (rd.next(r))
      std::printf("%s\t%s\t%s\n", r.name.c_str(), r.seq.empty() ? "*" : r.seq.c_str(),
                  r.qual.empty() ? "*" : r.qual.c_str());
  } catch (const std::exception& e) {
    std::printf("ERROR %s\n", e.what());
  }
  return 0;
}

int main(int argc, char** argv) {
  if (argc == 3 && std::string(argv[1]) == "--dump") return dump(argv[2]);
  if (argc != 4) {
    std::fprintf(stderr, "usage: fastx_check <reads> <batch_bytes> <threads>\n");
    return 2;
  }
  std::vector<dnaln_amd::SeqRecord> want;
  std::string want_err;
  try {
    dnaln_amd::SeqReader rd(argv[1]);
    dnaln_amd::SeqRecord r;
    while (rd.next(r)) want.push_back(r);
  } catch (const std::exception& e) {
    want_err = e.what();
  }
  std::vector<dnaln_amd::fastx::Part> parts;
  size_t k = 0;
  std::string got_err;
  bool parallel = false;
  try {
    dnaln_amd::fastx::Batcher b(argv[1], std::strtoull(argv[2], nullptr, 10), std::atoi(argv[3]));
    while (b.next(parts)) {
      parallel = b.parallel_fastq();
      for (const auto& p : parts)
        for (size_t i = 0; i < p.n(); ++i, ++k) {
          const std::string name(p.names.data() + p.name_off[i], p.name_off[i + 1] - p.name_off[i]);
          const std::string seq(p.seq.data() + p.seq_off[i], p.seq_off[i + 1] - p.seq_off[i]);
          const std::string qual = p.fq[i] ? std::string(p.qual.data() + p.seq_off[i], seq.size()) : std::string();
          if (k >= want.size()) {
            std::printf("extra record %zu '%s'\n", k, name.c_str());
            return 1;
          }
          const auto& w = want[k];
          if (name != w.name || seq != w.seq || qual != w.qual || (p.fq[i] != 0) != !w.qual.empty()) {
            if (!(p.fq[i] && seq.empty() && w.qual.empty() && name == w.name && seq == w.seq)) {
              std::printf("record %zu differs: '%s' vs '%s' (seq %zu/%zu, qual %zu/%zu, fq %d)\n", k, name.c_str(),
                          w.name.c_str(), seq.size(), w.seq.size(), qual.size(), w.qual.size(), (int)p.fq[i]);
              return 1;
            }
          }
        }
    }
  } catch (const std::exception& e) {
    got_err = e.what();
  }
  if (!want_err.empty() || !got_err.empty()) {
    if (want_err != got_err || k > want.size()) {
      std::printf("errors differ: sequential '%s' (%zu records), batcher '%s' (%zu records)\n", want_err.c_str(),
                  want.size(), got_err.c_str(), k);
      return 1;
    }
    std::printf("ok-error %zu '%s'\n", k, got_err.c_str());
    return 0;
  }
  if (k != want.size()) {
    std::printf("missing records: %zu of %zu\n", k, want.size());
    return 1;
  }
  std::printf("ok %zu parallel=%d\n", k, parallel || k == 0 ? 1 : 0);
  return 0;
}
