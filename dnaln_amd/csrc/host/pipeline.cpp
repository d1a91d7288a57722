// pipeline.cpp — the host stages around the device path, overlapped (dnaln_map_file).
//
// The reference's driver (src/main.cpp:47-61) reads a batch, maps it and formats it in turn on
// one thread. Here three stages run at once on consecutive batches:
//   input   parse FASTA/FASTQ (src/seqio.cpp:59-87), encode (src/common.hpp:36-42) and hash the
//           names (src/common.hpp:136-140) into page-locked staging, on `threads` threads;
//   device  upload, run and download, one thread per device batch (two), so one batch's copies
//           overlap the other's kernels;
//   output  format the SAM records (src/mapper.cpp:231-270) on `threads` threads and write them
//           in input order.
// Plain 4-line FASTQ is cut into pieces at record starts and parsed in parallel; any other input
// (FASTA, wrapped FASTQ, blank or stray lines) goes through a sequential reader with the
// reference's rules from the first batch that needs it.
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "dnaln_host.h"
#include "host/fastx_input.hpp"
#include "host/sam_text.hpp"

namespace {

using namespace dnaln_amd::fastx;

thread_local std::string g_host_err;

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Page-locked array from the HIP library, grown by reallocation.
template <class T>
struct Pinned {
  T* p = nullptr;
  size_t cap = 0;
  ~Pinned() { dnaln_hip_host_free(p); }
  void reserve(size_t n) {
    if (n <= cap) return;
    dnaln_hip_host_free(p);
    p = nullptr;
    cap = 0;
    const size_t c = std::max<size_t>(n + n / 4, 1024);
    void* q = nullptr;
    if (dnaln_hip_host_alloc(c * sizeof(T), &q) != DNALN_OK) throw Failure(DNALN_E_NOMEM, dnaln_hip_last_error());
    p = static_cast<T*>(q);
    cap = c;
  }
};

// One batch from read to write.
struct Slot {
  std::vector<Part> parts;
  size_t n = 0, bases = 0;
  Pinned<uint8_t> codes;
  Pinned<uint64_t> offs;
  Pinned<int32_t> lens;
  Pinned<uint32_t> fnv;
  Pinned<dnaln_record> rec;
  Pinned<uint32_t> cig;
  std::vector<std::string> text;  // formatted SAM, one string per part
  uint64_t mapped = 0;
  uint64_t seq = 0;               // position of the batch in the input
};

// Slots outlive the call that used them: their page-locked buffers (hipHostMalloc pins pages,
// which takes milliseconds per batch-sized buffer) are reused by later calls, so a warm call
// allocates nothing. Concurrent calls take different slots. The memory stays with the process.
class SlotPool {
 public:
  std::vector<std::unique_ptr<Slot>> take(size_t n) {
    std::lock_guard<std::mutex> g(m_);
    std::vector<std::unique_ptr<Slot>> out;
    while (out.size() < n && !free_.empty()) {
      out.push_back(std::move(free_.back()));
      free_.pop_back();
    }
    while (out.size() < n) out.emplace_back(new Slot);
    return out;
  }
  void give(std::vector<std::unique_ptr<Slot>>& slots) {
    std::lock_guard<std::mutex> g(m_);
    for (auto& s : slots) free_.push_back(std::move(s));
    slots.clear();
  }

 private:
  std::mutex m_;
  std::vector<std::unique_ptr<Slot>> free_;
};
SlotPool& slot_pool() {
  static SlotPool* p = new SlotPool;  // never destroyed: pinned memory is freed with the process
  return *p;
}

template <class T>
class Queue {
 public:
  void push(T v) {
    {
      std::lock_guard<std::mutex> g(m_);
      q_.push_back(v);
    }
    cv_.notify_one();
  }
  T pop() {
    std::unique_lock<std::mutex> g(m_);
    cv_.wait(g, [&] { return !q_.empty(); });
    T v = q_.front();
    q_.pop_front();
    return v;
  }

 private:
  std::mutex m_;
  std::condition_variable cv_;
  std::deque<T> q_;
};

struct Pipe {
  dnaln_hip_batch* b[2] = {nullptr, nullptr};
  const char* const* contig_names;
  std::vector<size_t> contig_len;
  int n_contigs;
  uint64_t batch_bytes;
  int threads;
  int out_fd;
  std::unique_ptr<Batcher> batcher;
  dnaln_pipeline_stats st{};

  // Parses the next batch into the slot's parts and encodes it; false at the end of the input.
  bool fill(Slot& s) {
    if (!batcher->next(s.parts)) return false;
    encode(s);
    return true;
  }

  // Codes, offsets, lengths and name hashes of every read, part by part in parallel.
  void encode(Slot& s) {
    const int np = (int)s.parts.size();
    std::vector<size_t> r0(np + 1, 0), b0(np + 1, 0);
    for (int t = 0; t < np; ++t) {
      r0[t + 1] = r0[t] + s.parts[t].n();
      b0[t + 1] = b0[t] + s.parts[t].seq.n;
    }
    s.n = r0[np];
    s.bases = b0[np];
    for (const Part& p : s.parts)
      for (size_t i = 0; i < p.n(); ++i)
        if (p.seq_off[i + 1] - p.seq_off[i] > (uint64_t)INT32_MAX) throw Failure(DNALN_E_INVALID, "read too long");
    s.codes.reserve(s.bases + 1);
    s.offs.reserve(s.n + 1);
    s.lens.reserve(s.n + 1);
    s.fnv.reserve(s.n + 1);
    parallel(threads, np, [&](int t) {
      const Part& p = s.parts[t];
      const unsigned char* src = reinterpret_cast<const unsigned char*>(p.seq.data());
      uint8_t* dst = s.codes.p + b0[t];
      for (size_t i = 0; i < p.seq.n; ++i) dst[i] = kNt4.t[src[i]];
      for (size_t i = 0; i < p.n(); ++i) {
        s.offs.p[r0[t] + i] = b0[t] + p.seq_off[i];
        s.lens.p[r0[t] + i] = (int32_t)(p.seq_off[i + 1] - p.seq_off[i]);
        s.fnv.p[r0[t] + i] = dnaln_fnv1a(p.names.data() + p.name_off[i], p.name_off[i + 1] - p.name_off[i]);
      }
    });
  }

  void format(Slot& s) {
    const int np = (int)s.parts.size();
    s.text.resize(np);
    std::vector<size_t> r0(np + 1, 0);
    for (int t = 0; t < np; ++t) r0[t + 1] = r0[t] + s.parts[t].n();
    parallel(threads, np, [&](int t) {
      const Part& p = s.parts[t];
      size_t bound = 0;
      for (size_t i = 0; i < p.n(); ++i) {
        const dnaln_record& m = s.rec.p[r0[t] + i];
        bound += dnaln_amd::sam_bound(p.name_off[i + 1] - p.name_off[i], p.seq_off[i + 1] - p.seq_off[i],
                                      m.cigar_len, m.contig >= 0 ? contig_len[m.contig] : 1);
      }
      std::string& out = s.text[t];
      out.resize(bound);
      char* o = &out[0];
      for (size_t i = 0; i < p.n(); ++i) {
        const dnaln_record& m = s.rec.p[r0[t] + i];
        const size_t len = p.seq_off[i + 1] - p.seq_off[i];
        const char* cname = m.contig >= 0 ? contig_names[m.contig] : "*";
        o = dnaln_amd::sam_record(o, p.names.data() + p.name_off[i], p.name_off[i + 1] - p.name_off[i],
                                  p.seq.data() + p.seq_off[i], p.fq[i] ? p.qual.data() + p.seq_off[i] : nullptr,
                                  len, m, s.cig.p, cname, m.contig >= 0 ? contig_len[m.contig] : 1);
      }
      out.resize((size_t)(o - out.data()));
    });
  }

  void write_out(const Slot& s) {
    for (const std::string& t : s.text) {
      size_t done = 0;
      while (done < t.size()) {
        const ssize_t w = ::write(out_fd, t.data() + done, t.size() - done);
        if (w < 0) {
          if (errno == EINTR) continue;
          throw Failure(DNALN_E_IO, "write error");
        }
        done += (size_t)w;
      }
      st.sam_bytes += t.size();
    }
  }

  void check(int rc, const char* what) {
    if (rc != DNALN_OK) throw Failure(rc, std::string(what) + ": " + dnaln_hip_last_error());
  }

  void run() {
    const double t0 = now();
    constexpr int kSlots = 6;  // parsing, two on the device, formatting, and one spare each side
    std::vector<std::unique_ptr<Slot>> slots = slot_pool().take(kSlots);
    struct GiveBack {
      std::vector<std::unique_ptr<Slot>>& s;
      ~GiveBack() { slot_pool().give(s); }
    } give_back{slots};
    Queue<Slot*> free_q, ready_q, done_q;
    for (auto& sl : slots) free_q.push(sl.get());
    std::string err;
    std::atomic<int> err_code{DNALN_OK};
    std::mutex err_m;
    auto note = [&](int code, const std::string& msg) {
      std::lock_guard<std::mutex> g(err_m);
      if (err_code == DNALN_OK) err = msg, err_code = code;
    };
    double s_input = 0, s_format = 0, s_write = 0;
    uint64_t seq = 0;  // input order of the slots the reader fills
    // input
    std::thread reader([&] {
      try {
        for (;;) {
          Slot* s = free_q.pop();
          if (!s) break;
          if (err_code != DNALN_OK) {  // another stage failed: stop reading
            free_q.push(s);
            break;
          }
          const double t = now();
          const bool more = fill(*s);
          s_input += now() - t;
          if (!more) {
            free_q.push(s);
            break;
          }
          s->seq = seq++;
          ready_q.push(s);
        }
      } catch (const Failure& f) {
        note(f.code, f.what());
      } catch (const std::exception& x) {
        note(DNALN_E_INTERNAL, x.what());
      }
      ready_q.push(nullptr);
    });
    // output
    std::thread writer([&] {
      bool failed = false;
      for (;;) {
        Slot* s = done_q.pop();
        if (!s) break;
        if (!failed) {
          try {
            double t = now();
            format(*s);
            s_format += now() - t;
            t = now();
            write_out(*s);
            s_write += now() - t;
          } catch (const Failure& f) {
            note(f.code, f.what());
            failed = true;
          } catch (const std::exception& x) {
            note(DNALN_E_INTERNAL, x.what());
            failed = true;
          }
        }
        free_q.push(s);
      }
    });
    // device: one thread per batch, each taking the next parsed slot through upload, run and
    // download on its own stream, so one batch's copies overlap the other's kernels (a run
    // blocks its thread on a few reads of counts). Slots reach the writer in input order: a
    // slot waits for its predecessor's hand-off. Busy time is the time at least one device
    // thread works on a slot.
    std::mutex turn_m;
    std::condition_variable turn_cv;
    uint64_t next_out = 0;
    int busy = 0;
    double busy_since = 0, busy_s = 0;
    auto set_busy = [&](int delta) {
      std::lock_guard<std::mutex> g(turn_m);
      const double t = now();
      if (busy == 0 && delta > 0) busy_since = t;
      busy += delta;
      if (busy == 0) busy_s += t - busy_since;
    };
    auto hand_off = [&](Slot* s, bool done, const dnaln_map_stats* ms) {
      {
        std::unique_lock<std::mutex> g(turn_m);
        turn_cv.wait(g, [&] { return next_out == s->seq; });
        if (done) {
          st.reads += ms->reads;
          st.mapped += ms->mapped;
          st.bases += s->bases;
          ++st.batches;
          done_q.push(s);
        } else {
          free_q.push(s);
        }
        ++next_out;
      }
      turn_cv.notify_all();
    };
    auto device = [&](int j) {
      for (;;) {
        Slot* s = ready_q.pop();
        if (!s) {
          ready_q.push(nullptr);  // the end, for the other device thread too
          return;
        }
        if (err_code != DNALN_OK) {
          hand_off(s, false, nullptr);
          continue;
        }
        dnaln_map_stats ms{};
        bool ok = false;
        set_busy(1);
        try {
          check(dnaln_hip_batch_upload(b[j], s->codes.p, s->offs.p, s->lens.p, s->fnv.p, s->n), "dnaln_hip_batch_upload");
          check(dnaln_hip_batch_run(b[j]), "dnaln_hip_batch_run");
          check(dnaln_hip_batch_poll(b[j], 1), "dnaln_hip_batch_poll");
          s->rec.reserve(s->n + 1);
          s->cig.reserve(8 * s->n + 64);  // a few CIGAR words per read; grown to the bound if short
          int rc = dnaln_hip_batch_download(b[j], s->rec.p, s->cig.p, s->cig.cap, &ms);
          if (rc == DNALN_E_CAPACITY) {
            s->cig.reserve(2 * s->bases + 4 * s->n + 4);
            rc = dnaln_hip_batch_download(b[j], s->rec.p, s->cig.p, s->cig.cap, &ms);
          }
          check(rc, "dnaln_hip_batch_download");
          ok = true;
        } catch (const Failure& f) {
          note(f.code, f.what());  // the reader stops at its next slot
        } catch (const std::exception& x) {
          note(DNALN_E_INTERNAL, x.what());
        }
        if (!ok) (void)dnaln_hip_batch_poll(b[j], 1);
        set_busy(-1);
        hand_off(s, ok, &ms);
      }
    };
    std::thread dev1(device, 1);
    device(0);
    dev1.join();
    st.s_device = busy_s;
    free_q.push(nullptr);
    reader.join();
    done_q.push(nullptr);
    writer.join();
    st.s_input = s_input;
    st.s_format = s_format;
    st.s_write = s_write;
    st.s_wall = now() - t0;
    st.parallel_fastq = batcher->parallel_fastq() ? 1 : 0;
    if (err_code != DNALN_OK) throw Failure(err_code.load(), err);
  }
};

}  // namespace

extern "C" {

const char* dnaln_host_last_error(void) { return g_host_err.c_str(); }

int dnaln_map_file_on(dnaln_hip_batch* b0, dnaln_hip_batch* b1, const char* reads_path, int out_fd,
                      const char* const* contig_names, int32_t n_contigs, uint64_t batch_bytes, int32_t threads,
                      dnaln_pipeline_stats* stats) {
  g_host_err.clear();
  if (!b0 || !b1 || b0 == b1 || !reads_path || !contig_names || n_contigs < 1 || out_fd < 0) {
    g_host_err = "null argument (or the same batch twice)";
    return DNALN_E_INVALID;
  }
  Pipe p;
  p.b[0] = b0;
  p.b[1] = b1;
  p.contig_names = contig_names;
  p.n_contigs = n_contigs;
  for (int i = 0; i < n_contigs; ++i) p.contig_len.push_back(std::strlen(contig_names[i]));
  p.batch_bytes = std::max<uint64_t>(batch_bytes ? batch_bytes : (64ull << 20), 1 << 16);
  p.threads = std::max(1, std::min(threads > 0 ? threads : 8, 64));
  p.out_fd = out_fd;
  try {
    p.batcher.reset(new Batcher(reads_path, p.batch_bytes, p.threads));
    p.run();
  } catch (const Failure& f) {
    g_host_err = f.what();
    if (stats) *stats = p.st;
    return f.code;
  } catch (const std::exception& x) {
    g_host_err = x.what();
    return DNALN_E_INTERNAL;
  }
  if (stats) *stats = p.st;
  return DNALN_OK;
}

int dnaln_map_encoded(dnaln_hip_batch* const* batches, int32_t n_batches, const uint8_t* codes, const uint64_t* offs,
                      const int32_t* lens, const uint32_t* fnv, uint64_t n, uint64_t chunk_reads, dnaln_record* out,
                      uint32_t* cigar_pool, uint64_t cigar_cap, uint64_t* cigar_used, dnaln_pipeline_stats* stats) {
  if (!batches || n_batches < 1 || n_batches > 64 || !cigar_used ||
      (n && (!codes || !offs || !lens || !fnv || !out || !cigar_pool))) {
    g_host_err = "null argument or batch count outside 1..64";
    return DNALN_E_INVALID;
  }
  for (int j = 0; j < n_batches; ++j)
    for (int x = 0; x < j; ++x)
      if (!batches[j] || batches[j] == batches[x]) {
        g_host_err = "null batch, or the same batch twice";
        return DNALN_E_INVALID;
      }
  const uint64_t R = std::max<uint64_t>(chunk_reads ? chunk_reads : (1ull << 18), 1);
  const uint64_t n_chunks = (n + R - 1) / R;
  const double t0 = now();
  std::atomic<uint64_t> next{0};
  std::mutex turn_m;
  std::condition_variable turn_cv;
  uint64_t next_out = 0, used = 0, mapped = 0;
  int err_code = DNALN_OK;
  std::string err;
  // Chunk c: reads [c R, min(n, (c + 1) R)), uploaded from the caller's arrays (offsets rebased
  // to the chunk's first base), records downloaded straight into out. The chunk's CIGAR words,
  // staged per thread, are appended to the pool in chunk order, and its records' offsets moved
  // with them.
  auto device = [&](int j) {
    Pinned<uint64_t> roffs;
    Pinned<uint32_t> cig;
    for (;;) {
      const uint64_t c = next.fetch_add(1);
      if (c >= n_chunks) return;
      const uint64_t r0 = c * R, m = std::min(n, r0 + R) - r0;
      int rc = DNALN_OK;
      std::string e;
      dnaln_map_stats ms{};
      try {
        uint64_t lo = ~0ull;
        for (uint64_t i = 0; i < m; ++i) lo = std::min(lo, offs[r0 + i]);
        roffs.reserve(m + 1);
        uint64_t bases = 0;
        for (uint64_t i = 0; i < m; ++i) {
          roffs.p[i] = offs[r0 + i] - lo;
          bases += (uint64_t)std::max(lens[r0 + i], 0);
        }
        cig.reserve(8 * m + 64);
        rc = dnaln_hip_batch_upload(batches[j], codes + lo, roffs.p, lens + r0, fnv + r0, m);
        if (rc == DNALN_OK) rc = dnaln_hip_batch_run(batches[j]);
        if (rc == DNALN_OK) rc = dnaln_hip_batch_poll(batches[j], 1);
        if (rc == DNALN_OK) rc = dnaln_hip_batch_download(batches[j], out + r0, cig.p, cig.cap, &ms);
        if (rc == DNALN_E_CAPACITY) {
          cig.reserve(2 * bases + 4 * m + 4);
          rc = dnaln_hip_batch_download(batches[j], out + r0, cig.p, cig.cap, &ms);
        }
        if (rc != DNALN_OK) e = dnaln_hip_last_error();
      } catch (const std::exception& x) {
        rc = DNALN_E_NOMEM;
        e = x.what();
      }
      std::unique_lock<std::mutex> g(turn_m);
      turn_cv.wait(g, [&] { return next_out == c; });
      if (rc != DNALN_OK && err_code == DNALN_OK) err_code = rc, err = e;
      if (rc == DNALN_OK && err_code == DNALN_OK) {
        uint64_t words = 0;
        for (uint64_t i = 0; i < m; ++i) words = std::max<uint64_t>(words, out[r0 + i].cigar_off + out[r0 + i].cigar_len);
        if (used + words > cigar_cap) {
          err_code = DNALN_E_CAPACITY;
          err = "cigar_pool too small";
        } else {
          if (words) std::memcpy(cigar_pool + used, cig.p, words * 4);
          for (uint64_t i = 0; i < m; ++i) out[r0 + i].cigar_off += used;
          used += words;
          mapped += ms.mapped;
        }
      }
      ++next_out;
      g.unlock();
      turn_cv.notify_all();
    }
  };
  {
    // A thread that cannot be started leaves its chunks to the others (the calling thread takes
    // chunks too), so the records are the same with fewer threads.
    std::vector<std::thread> th;
    try {
      th.reserve((size_t)n_batches);
      for (int j = 1; j < n_batches; ++j) th.emplace_back(device, j);
    } catch (const std::exception&) {
    }
    device(0);
    for (auto& t : th) t.join();
  }
  *cigar_used = used;
  if (stats) {
    std::memset(stats, 0, sizeof *stats);
    stats->reads = n;
    stats->mapped = mapped;
    stats->batches = n_chunks;
    stats->s_device = stats->s_wall = now() - t0;
  }
  if (err_code != DNALN_OK) g_host_err = err;
  return err_code;
}

int dnaln_map_file(dnaln_hip_ctx* ctx, const dnaln_batch_config* cfg, const char* reads_path, int out_fd,
                   const char* const* contig_names, int32_t n_contigs, uint64_t batch_bytes, int32_t threads,
                   dnaln_pipeline_stats* stats) {
  g_host_err.clear();
  if (!ctx) {
    g_host_err = "null argument";
    return DNALN_E_INVALID;
  }
  dnaln_batch_config c;
  if (cfg) c = *cfg;
  else dnaln_hip_default_batch_config(&c);
  dnaln_hip_batch* b[2] = {nullptr, nullptr};
  int rc = DNALN_OK;
  for (int j = 0; j < 2 && rc == DNALN_OK; ++j) rc = dnaln_hip_batch_create(ctx, &c, &b[j]);
  if (rc == DNALN_OK)
    rc = dnaln_map_file_on(b[0], b[1], reads_path, out_fd, contig_names, n_contigs, batch_bytes, threads, stats);
  else
    g_host_err = std::string("dnaln_hip_batch_create: ") + dnaln_hip_last_error();
  for (int j = 0; j < 2; ++j) dnaln_hip_batch_destroy(b[j]);
  return rc;
}

}  // extern "C"
