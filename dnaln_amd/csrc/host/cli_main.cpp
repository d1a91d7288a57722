// cli_main.cpp — `dnaln-amd map`: the reference's `dnaln map` (src/main.cpp:12-75) on one MI355X,
// and `dnaln-amd index`, which writes the index to a file that `map --index` loads.
//
// Same arguments, defaults, validation and SAM output; reads go to the GPU in batches of
// --batch reads (larger than the reference's 8192 by default, since a batch fills the GPU).
#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "host/dnaln_amd.hpp"

using namespace dnaln_amd;

namespace {

[[noreturn]] void die(const std::string& msg) {
  std::fprintf(stderr, "[dnaln-amd] error: %s\n", msg.c_str());
  std::exit(1);
}

void log_msg(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void log_msg(const char* fmt, ...) {
  std::fprintf(stderr, "[dnaln-amd] ");
  va_list ap;
  va_start(ap, fmt);
  std::vfprintf(stderr, fmt, ap);
  va_end(ap);
  std::fputc('\n', stderr);
}

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Options as src/cli.hpp:12-70 parses them: value options take the next argument, flags
// stand alone, anything else is positional; unknown options are errors.
struct Args {
  std::vector<std::string> pos;
  std::map<std::string, std::string> opt;
  Args(int argc, char** argv, const std::set<std::string>& values, const std::set<std::string>& flags) {
    for (int i = 0; i < argc; ++i) {
      const std::string a = argv[i];
      if (a.size() > 1 && a[0] == '-') {
        if (values.count(a)) {
          if (i + 1 >= argc) die("option " + a + " needs a value");
          opt[a] = argv[++i];
        } else if (flags.count(a)) {
          opt[a] = "";
        } else {
          die("unknown option " + a);
        }
      } else {
        pos.push_back(a);
      }
    }
  }
  bool has(const std::string& k) const { return opt.count(k) != 0; }
  long long num(const std::string& k, long long def) const {
    const auto it = opt.find(k);
    if (it == opt.end()) return def;
    char* end = nullptr;
    const long long v = std::strtoll(it->second.c_str(), &end, 10);
    if (it->second.empty() || *end) die("option " + k + " expects an integer, got '" + it->second + "'");
    return v;
  }
};

const char* kHelp =
    "Usage: dnaln-amd map [options] <ref.fa[.gz]> <reads.fq[.gz]>  > out.sam\n\n"
    "  -o FILE        write SAM to FILE instead of stdout\n"
    "  --batch INT    about INT reads per GPU batch: the input is cut into batches of INT x 256\n"
    "                 bytes [262144]\n"
    "  --threads INT  host threads for parsing and SAM formatting [16]\n"
    "  -k INT         minimizer k-mer size [15]\n"
    "  -w INT         minimizer window size [10]\n"
    "  --bw INT       band half-width of banded Smith-Waterman [16]\n"
    "  -A/-B INT      match score / mismatch penalty [2/4]\n"
    "  -O/-E INT      gap open / extension penalty; gap of length L costs O+L*E [4/2]\n"
    "  --max-occ INT  ignore minimizers with more occurrences [500]\n"
    "  --max-cand INT candidate loci aligned per read [10]\n"
    "  --min-chain INT minimum chain score [20]\n"
    "  --min-score INT minimum alignment score to report [40]\n"
    "  --scalar, --no-prefetch   accepted for compatibility (no effect on the GPU)\n"
    "  --plain        plain GPU form (one launch sequence per 65,536-read chunk)\n"
    "  --tune         run the autotuner on the first batch\n"
    "  --device INT   HIP device [0]\n"
    "  --index FILE   load the index from FILE (written by `dnaln-amd index`) instead of building it\n"
    "  --gpu-index    build the index on the GPU instead of the host (k <= 16)\n";

const char* kIndexHelp =
    "Usage: dnaln-amd index [options] <ref.fa[.gz]> -o <out.dnx>\n\n"
    "  -k INT         minimizer k-mer size [15]\n"
    "  -w INT         minimizer window size [10]\n"
    "  --gpu          build on the GPU (k <= 16)\n"
    "  --device INT   HIP device [0]\n";

// Builds the index of a reference and writes it to a file that `map --index` loads (the
// reference rebuilds its index on every run, src/main.cpp:34-36).
int cmd_index(int argc, char** argv) {
  const Args a(argc, argv, {"-k", "-w", "-o", "--device"}, {"--gpu", "-h", "--help"});
  if (a.pos.size() != 1 || !a.has("-o") || a.has("-h") || a.has("--help")) {
    std::fputs(kIndexHelp, stderr);
    return a.has("-h") || a.has("--help") ? 0 : 1;
  }
  const int k = (int)a.num("-k", 15), w = (int)a.num("-w", 10);
  double t = now();
  const Reference ref = Reference::load(a.pos[0]);
  log_msg("reference: %zu contig(s), %zu bp (loaded in %.2f s)", ref.contigs.size(), ref.seq.size(), now() - t);
  t = now();
  MinimizerIndex idx;
  if (a.has("--gpu")) idx.build_gpu(ref, k, w, (int)a.num("--device", 0));
  else idx.build(ref, k, w);
  log_msg("index: k=%d w=%d, %zu distinct minimizers, %zu positions, %.1f MB, built on the %s in %.2f s", k, w,
          idx.n_keys(), idx.n_positions(), idx.memory_bytes() / 1048576.0, a.has("--gpu") ? "GPU" : "host",
          now() - t);
  t = now();
  idx.save(ref, a.opt.at("-o"));
  log_msg("wrote %s in %.2f s", a.opt.at("-o").c_str(), now() - t);
  return 0;
}

int cmd_map(int argc, char** argv) {
  const std::set<std::string> values = {"-k", "-w", "--bw", "-A", "-B", "-O", "-E", "--max-occ", "--min-score",
                                        "--max-cand", "--min-chain", "-o", "--batch", "--device", "--index",
                                        "--threads"};
  const Args a(argc, argv, values,
               {"--scalar", "--no-prefetch", "--plain", "--tune", "--gpu-index", "-h", "--help"});
  if (a.pos.size() != 2 || a.has("-h") || a.has("--help")) {
    std::fputs(kHelp, stderr);
    return a.has("-h") || a.has("--help") ? 0 : 1;
  }
  MapParams p;
  p.k = (int)a.num("-k", p.k);
  p.w = (int)a.num("-w", p.w);
  p.sc.bw = (int)a.num("--bw", p.sc.bw);
  p.sc.match = (int)a.num("-A", p.sc.match);
  p.sc.mismatch = (int)a.num("-B", p.sc.mismatch);
  p.sc.gap_open = (int)a.num("-O", p.sc.gap_open);
  p.sc.gap_ext = (int)a.num("-E", p.sc.gap_ext);
  p.max_occ = (uint32_t)a.num("--max-occ", p.max_occ);
  p.max_cand = (int)a.num("--max-cand", p.max_cand);
  p.ch.min_score = (int)a.num("--min-chain", p.ch.min_score);
  p.min_aln_score = (int)a.num("--min-score", p.min_aln_score);
  p.batched = !a.has("--plain");
  p.tune = a.has("--tune");
  p.device = (int)a.num("--device", 0);
  // src/cli.hpp:93-95
  if (p.sc.bw < 1 || p.sc.bw > 1000) die("--bw must be in [1, 1000]");
  if (p.sc.match < 1 || p.sc.mismatch < 0 || p.sc.gap_open < 0 || p.sc.gap_ext < 1)
    die("scores must satisfy A >= 1, B >= 0, O >= 0, E >= 1");
  const long long batch = a.num("--batch", 1 << 18);
  if (batch < 1) die("--batch must be positive");
  const long long threads = a.num("--threads", 16);
  if (threads < 1 || threads > 64) die("--threads must be in [1, 64]");
  std::string cmdline = "dnaln map";
  for (int i = 0; i < argc; ++i) cmdline += std::string(" ") + argv[i];

  const double t_total = now();
  double t = now();
  const Reference ref = Reference::load(a.pos[0]);
  log_msg("reference: %zu contig(s), %zu bp (loaded in %.2f s)", ref.contigs.size(), ref.seq.size(), now() - t);
  t = now();
  MinimizerIndex idx;
  const char* how = "built";
  if (a.has("--index")) {
    idx.load(ref, a.opt.at("--index"));
    if (idx.k() != p.k || idx.w() != p.w)
      die("index file has k=" + std::to_string(idx.k()) + " w=" + std::to_string(idx.w()) +
          "; pass the same -k / -w");
    how = "loaded";
  } else if (a.has("--gpu-index")) {
    idx.build_gpu(ref, p.k, p.w, p.device);
    how = "built on the GPU";
  } else {
    idx.build(ref, p.k, p.w);
  }
  log_msg("index: k=%d w=%d, %zu distinct minimizers, %zu positions, %.1f MB, %s in %.2f s", p.k, p.w,
          idx.n_keys(), idx.n_positions(), idx.memory_bytes() / 1048576.0, how, now() - t);

  FILE* out = stdout;
  if (a.has("-o")) {
    out = std::fopen(a.opt.at("-o").c_str(), "w");
    if (!out) die("cannot open '" + a.opt.at("-o") + "' for writing");
  }
  Mapper mapper(ref, idx, p);
  mapper.write_sam_header(out, cmdline);
  log_msg("alignment kernel: %s (%s form)", mapper.kernel_name().c_str(), p.batched ? "batched" : "plain");

  // Input, device and output overlap on consecutive batches (dnaln_map_file): the records and
  // their SAM text are those of Mapper::map + append_sam, batch after batch.
  std::fflush(out);
  std::vector<const char*> names;
  for (const Contig& c : ref.contigs) names.push_back(c.name.c_str());
  dnaln_batch_config cfg;
  dnaln_hip_default_batch_config(&cfg);
  cfg.mode = p.batched ? DNALN_MODE_BATCHED : DNALN_MODE_PLAIN;
  cfg.tune = p.tune ? 1 : 0;
  dnaln_pipeline_stats st{};
  const int rc = dnaln_map_file(mapper.context(), &cfg, a.pos[1].c_str(), fileno(out), names.data(),
                                (int32_t)names.size(), (uint64_t)batch * 256, (int32_t)threads, &st);
  if (out != stdout) {
    if (std::fclose(out) != 0) die("write error");
  }
  if (rc != DNALN_OK) die(dnaln_host_last_error());
  log_msg("mapped %llu / %llu reads (%.2f%%) in %.3f s -> %.0f reads/s (excl. index build)",
          (unsigned long long)st.mapped, (unsigned long long)st.reads, st.reads ? 100.0 * st.mapped / st.reads : 0.0,
          st.s_wall, st.reads / std::max(st.s_wall, 1e-9));
  log_msg("%llu batches, %s input; busy seconds: input %.3f, device %.3f, SAM formatting %.3f, write %.3f; total "
          "wall %.2f s",
          (unsigned long long)st.batches, st.parallel_fastq ? "4-line FASTQ (parallel)" : "sequential",
          st.s_input, st.s_device, st.s_format, st.s_write, now() - t_total);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2 || !std::strcmp(argv[1], "-h") || !std::strcmp(argv[1], "--help")) {
    std::fputs("Usage: dnaln-amd map [options] <ref.fa> <reads.fq>\n"
               "       dnaln-amd index [options] <ref.fa> -o <out.dnx>\n",
               stderr);
    return argc < 2 ? 1 : 0;
  }
  const bool map = !std::strcmp(argv[1], "map"), index = !std::strcmp(argv[1], "index");
  if (!map && !index) {
    std::fprintf(stderr, "[dnaln-amd] error: unknown command '%s' (this build provides `map` and `index`)\n",
                 argv[1]);
    return 1;
  }
  try {
    return map ? cmd_map(argc - 2, argv + 2) : cmd_index(argc - 2, argv + 2);
  } catch (const Error& e) {
    die(e.what());
  }
}
