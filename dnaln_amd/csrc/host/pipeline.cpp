// pipeline.cpp — the host stages around the device path, overlapped (dnaln_map_file).
//
// The reference's driver (src/main.cpp:47-61) reads a batch, maps it and formats it in turn on
// one thread. Here three stages run at once on consecutive batches:
//   input   parse FASTA/FASTQ (src/seqio.cpp:59-87), encode (src/common.hpp:36-42) and hash the
//           names (src/common.hpp:136-140) into page-locked staging, on `threads` threads;
//   device  upload, run and download on two device batches in turn, so one batch's copies
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
    constexpr int kSlots = 4;
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
    // device (this thread): batch k takes slot i while batch k ^ 1 finishes slot i - 1
    Slot* pending[2] = {nullptr, nullptr};
    int k = 0;
    auto finish = [&](int j) {
      Slot* s = pending[j];
      const double t = now();
      check(dnaln_hip_batch_poll(b[j], 1), "dnaln_hip_batch_poll");
      s->rec.reserve(s->n + 1);
      s->cig.reserve(8 * s->n + 64);  // a few CIGAR words per read; grown to the bound if short
      dnaln_map_stats ms{};
      int rc = dnaln_hip_batch_download(b[j], s->rec.p, s->cig.p, s->cig.cap, &ms);
      if (rc == DNALN_E_CAPACITY) {
        s->cig.reserve(2 * s->bases + 4 * s->n + 4);
        rc = dnaln_hip_batch_download(b[j], s->rec.p, s->cig.p, s->cig.cap, &ms);
      }
      check(rc, "dnaln_hip_batch_download");
      st.s_device += now() - t;
      st.reads += ms.reads;
      st.mapped += ms.mapped;
      st.bases += s->bases;
      ++st.batches;
      pending[j] = nullptr;
      done_q.push(s);
    };
    try {
      for (;;) {
        Slot* s = ready_q.pop();
        if (!s) break;
        if (err_code != DNALN_OK) {
          free_q.push(s);
          continue;
        }
        const double t = now();
        check(dnaln_hip_batch_upload(b[k], s->codes.p, s->offs.p, s->lens.p, s->fnv.p, s->n), "dnaln_hip_batch_upload");
        check(dnaln_hip_batch_run(b[k]), "dnaln_hip_batch_run");
        st.s_device += now() - t;
        pending[k] = s;
        if (pending[k ^ 1]) finish(k ^ 1);
        k ^= 1;
      }
      for (int j : {k, k ^ 1})
        if (pending[j]) finish(j);
    } catch (const Failure& f) {
      note(f.code, f.what());  // the reader stops at its next slot
      for (int j = 0; j < 2; ++j)
        if (pending[j]) {
          (void)dnaln_hip_batch_poll(b[j], 1);
          free_q.push(pending[j]);
        }
      for (;;) {
        Slot* s = ready_q.pop();
        if (!s) break;
        free_q.push(s);
      }
    }
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
