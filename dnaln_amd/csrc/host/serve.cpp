// serve.cpp — continuous batching of mapping requests on the resident index (dnaln_serve.h).
#include "dnaln_serve.h"
#include "device/bw_list.hpp"
#include "host/c_guard.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// A request's reads: offsets relative to `codes`. Either its own copy of the caller's arrays
// (dnaln_serve_submit) or a view of arrays that outlive it (the trace replay).
struct Request {
  uint64_t id = 0;
  std::vector<uint8_t> own_codes;
  std::vector<uint64_t> own_offs;
  std::vector<int32_t> own_lens;
  std::vector<uint32_t> own_fnv;
  const uint8_t* codes = nullptr;
  const uint64_t* offs = nullptr;
  const int32_t* lens = nullptr;
  const uint32_t* fnv = nullptr;
  uint64_t n = 0;
  uint64_t code_bytes = 0;  // bytes of codes the reads span
  uint64_t bases = 0;
  int32_t maxlen = 0;
  bool view = false;
};

// Grow-only host array whose new elements are left uninitialized: a std::vector's resize
// zero-fills every element it exposes, hundreds of MB per batch for the CIGAR bound, and that
// stalled the worker whenever batches first grew (at each rise of the offered load).
template <class T>
struct HostBuf {
  std::unique_ptr<T[]> p;
  size_t n = 0, cap = 0;
  void resize(size_t k) {
    if (k > cap) {
      const size_t c = std::max(k, cap + cap / 2);
      std::unique_ptr<T[]> q(new T[c]);
      if (n) std::memcpy(q.get(), p.get(), n * sizeof(T));
      p.swap(q);
      cap = c;
    }
    n = k;
  }
  T* data() { return p.get(); }
  const T* data() const { return p.get(); }
  size_t size() const { return n; }
  // Allocates room for k elements and writes it once, so that the pages are mapped before the
  // first batch instead of during it.
  void prefault(size_t k) {
    resize(k);
    if (k) std::memset(static_cast<void*>(p.get()), 0, k * sizeof(T));
    n = 0;
  }
  T& operator[](size_t i) { return p[i]; }
  const T& operator[](size_t i) const { return p[i]; }
  const T* begin() const { return p.get(); }
  const T* end() const { return p.get() + n; }
};

struct Done {
  uint64_t id = 0;
  double t_done = 0;
  std::vector<dnaln_record> rec;  // cigar_off relative to cig
  std::vector<uint32_t> cig;
};

// A batch slot and its worker thread. The main thread hands a batch of requests over (state 1)
// and takes it back finished (state 2); in between only the worker touches the slot: it gathers
// the requests into one upload, runs the batch, and splits the records back into requests, so
// the copying of every batch runs on its own thread.
struct Slot {
  dnaln_hip_batch* b = nullptr;
  std::mutex m;
  std::condition_variable cv;
  int state = 0;  // 0 free, 1 handed to the worker, 2 finished
  bool quit = false;
  int rc = 0;
  std::string err;
  std::thread worker;
  uint64_t est = 0;
  std::vector<Request> reqs;  // the batch
  std::vector<Done> out;      // its requests, finished
  HostBuf<uint8_t> codes;
  HostBuf<uint64_t> offs;
  HostBuf<int32_t> lens;
  HostBuf<uint32_t> fnv;
  HostBuf<dnaln_record> rec;
  HostBuf<uint32_t> cig;

  void gather() {
    uint64_t n = 0, nb = 0;
    for (const Request& q : reqs) n += q.n, nb += q.code_bytes;
    codes.resize(nb);
    offs.resize(n);
    lens.resize(n);
    fnv.resize(n);
    uint64_t r0 = 0, c0 = 0;
    for (const Request& q : reqs) {
      if (q.code_bytes) std::memcpy(codes.data() + c0, q.codes, q.code_bytes);
      for (uint64_t i = 0; i < q.n; ++i) offs[r0 + i] = c0 + q.offs[i];
      if (q.n) {
        std::memcpy(lens.data() + r0, q.lens, q.n * sizeof(int32_t));
        std::memcpy(fnv.data() + r0, q.fnv, q.n * sizeof(uint32_t));
      }
      r0 += q.n;
      c0 += q.code_bytes;
    }
  }

  void split() {
    out.clear();
    out.resize(reqs.size());
    uint64_t r0 = 0;
    for (size_t x = 0; x < reqs.size(); ++x) {
      Done& d = out[x];
      d.id = reqs[x].id;
      const uint64_t n = reqs[x].n;
      d.rec.assign(rec.begin() + r0, rec.begin() + r0 + n);
      uint64_t nc = 0;
      for (const dnaln_record& m : d.rec) nc += m.cigar_len;
      d.cig.resize(nc);
      uint64_t o = 0;
      for (dnaln_record& m : d.rec) {
        std::copy(cig.begin() + m.cigar_off, cig.begin() + m.cigar_off + m.cigar_len, d.cig.begin() + o);
        m.cigar_off = o;
        o += m.cigar_len;
      }
      r0 += n;
    }
    const double t = now_s();
    for (Done& d : out) d.t_done = t;
    reqs.clear();
  }

  void work() {
    for (;;) {
      std::unique_lock<std::mutex> lk(m);
      cv.wait(lk, [&] { return state == 1 || quit; });
      if (state != 1) return;
      lk.unlock();
      // The batch's whole turn runs under one try: an exception on this thread (a host buffer
      // that cannot grow) comes back through collect() as the slot's status, like a device error.
      int r = DNALN_OK;
      std::string e;
      try {
        gather();
        r = dnaln_hip_batch_upload(b, codes.data(), offs.data(), lens.data(), fnv.data(), lens.size());
        if (r == DNALN_OK) r = dnaln_hip_batch_run(b);
        if (r == DNALN_OK) r = dnaln_hip_batch_poll(b, 1);
        if (r == DNALN_OK) {
          // A few CIGAR words per read; the bound (two per base) only when the device holds more.
          rec.resize(lens.size());
          cig.resize(std::max<size_t>(cig.cap, 8 * lens.size() + 64));
          r = dnaln_hip_batch_download(b, rec.data(), cig.data(), cig.size(), nullptr);
          if (r == DNALN_E_CAPACITY) {
            uint64_t cap = 1;
            for (const int32_t l : lens) cap += 2 * (uint64_t)l + 4;
            cig.resize(cap);
            r = dnaln_hip_batch_download(b, rec.data(), cig.data(), cig.size(), nullptr);
          }
        }
        if (r != DNALN_OK) e = dnaln_hip_last_error();
        if (r == DNALN_OK) split();
      } catch (const std::bad_alloc&) {
        r = DNALN_E_NOMEM;
        e = "serving slot: host allocation failed";
      } catch (...) {
        r = DNALN_E_INTERNAL;
        e = "serving slot: unexpected C++ exception";
      }
      if (r != DNALN_OK) {
        out.clear();
        reqs.clear();
      }
      lk.lock();
      rc = r;
      err = e;
      state = 2;
    }
  }
};

}  // namespace

struct dnaln_server {
  dnaln_hip_ctx* ctx = nullptr;
  dnaln_serve_config cfg{};
  int bw = 16, max_cand = 10;
  std::vector<std::unique_ptr<Slot>> slots;
  std::deque<Request> queue;
  std::deque<Done> done;
  dnaln_serve_stats st{};

  std::string last_err;

  ~dnaln_server() {
    for (auto& s : slots) {
      {
        std::lock_guard<std::mutex> lk(s->m);
        s->quit = true;
      }
      s->cv.notify_one();
      if (s->worker.joinable()) s->worker.join();
      if (s->b) dnaln_hip_batch_destroy(s->b);
    }
  }

  uint64_t batch_bytes(uint64_t n, uint64_t bases, int32_t maxlen) const {
    // Inputs, per-read state, candidates and tasks, CIGAR slots, anchors of reads on the
    // global path, minimizer lists of the fused kernel, and the traceback directions of one SW
    // launch (4 bytes per group of 8 band cells per row per alignment), bounded by the
    // per-launch direction budget.
    // Where the fill may split a pair over P lanes, a row holds P segments of ceil(ceil(W/P)/8)
    // groups, and each group of 128 tasks holds P - 1 more rows.
    const int W = 2 * bw + 1;
    uint64_t G = (uint64_t)(W + 7) / 8, lag = 0;
#define DN_CASE(B, P)                                                                           \
  if (bw == B) {                                                                                \
    G = std::max<uint64_t>(G, (uint64_t)(((W + P - 1) / P + 7) / 8) * P);                      \
    lag = std::max<uint64_t>(lag, P - 1);                                                      \
  }
    DN_SPLIT_LIST(DN_CASE)
#undef DN_CASE
    const uint64_t per_read = 16 + sizeof(dnaln_record) + 64 + (uint64_t)(max_cand + 1) * 96 + 64 * 48 + 64 * 8;
    const uint64_t dirs = std::min<uint64_t>(
        bases * 4 * G + ((uint64_t)maxlen + 4 + lag) * G * 512 + (n + 127) / 128 * (lag + 3) * G * 512, cfg.dir_budget);
    // (12 bytes per base: the global path's probe results between its count and emit passes)
    return bases * 13 + 256 + n * per_read + n * (2 * (uint64_t)maxlen + 4) * 4 + dirs;
  }

  int collect() {
    int finished = 0;
    for (auto& sp : slots) {
      Slot& s = *sp;
      {
        std::lock_guard<std::mutex> lk(s.m);
        if (s.state != 2) continue;
        s.state = 0;  // the worker is back waiting for state 1; its results stay untouched
      }
      st.inflight_batches -= 1;
      st.inflight_bytes -= s.est;
      if (s.rc != DNALN_OK) {
        last_err = s.err;
        s.reqs.clear();
        return s.rc;
      }
      for (Done& d : s.out) {
        done.push_back(std::move(d));
        ++finished;
        ++st.completed;
      }
      s.out.clear();
    }
    return finished;
  }

  // Drops the queued requests that are views of a caller's arrays and waits until no batch is
  // running, so that nothing refers to those arrays any more (a trace replay ending early).
  void drop_views() {
    std::deque<Request> keep;
    for (Request& q : queue)
      if (!q.view) keep.push_back(std::move(q));
    queue.swap(keep);
    while (st.inflight_batches > 0) {
      (void)collect();
      std::this_thread::yield();
    }
  }

  int admit() {
    for (auto& sp : slots) {
      if (queue.empty()) return DNALN_OK;
      Slot& s = *sp;
      {
        std::lock_guard<std::mutex> lk(s.m);
        if (s.state != 0) continue;
      }
      // Pack from the head of the queue while the batch limits and the budget allow.
      uint64_t n = 0, bases = 0, k = 0;
      int32_t maxlen = 0;
      uint64_t est = 0;
      while (k < queue.size()) {
        const Request& q = queue[k];
        const uint64_t n2 = n + q.n, b2 = bases + q.bases;
        const int32_t m2 = std::max(maxlen, q.maxlen);
        if (k > 0 && (n2 > cfg.max_batch_reads || b2 > cfg.max_batch_bases)) break;
        const uint64_t e2 = batch_bytes(n2, b2, m2);
        if (st.inflight_bytes + e2 > st.budget_bytes) break;
        n = n2;
        bases = b2;
        maxlen = m2;
        est = e2;
        ++k;
      }
      if (k == 0) {  // the head does not fit next to the batches in flight: wait for them
        ++st.deferred_by_budget;
        return DNALN_OK;
      }
      s.reqs.clear();
      uint64_t reads = 0;
      for (uint64_t r = 0; r < k; ++r) {
        reads += queue.front().n;
        s.reqs.push_back(std::move(queue.front()));
        queue.pop_front();
      }
      s.est = est;
      {
        std::lock_guard<std::mutex> lk(s.m);
        s.state = 1;
      }
      s.cv.notify_one();
      st.batches += 1;
      st.reads += reads;
      st.inflight_batches += 1;
      st.inflight_bytes += est;
      st.max_inflight_batches = std::max(st.max_inflight_batches, st.inflight_batches);
      st.max_inflight_bytes = std::max(st.max_inflight_bytes, st.inflight_bytes);
    }
    return DNALN_OK;
  }
};

extern "C" {

void dnaln_serve_default_config(dnaln_serve_config* cfg) {
  std::memset(cfg, 0, sizeof *cfg);
  cfg->slots = 6;
  cfg->mode = DNALN_MODE_BATCHED;
  cfg->hbm_budget = 0;
  cfg->max_batch_reads = 1 << 17;
  cfg->max_batch_bases = 1ull << 26;
  cfg->dir_budget = 0;
  cfg->max_read_len = 1000;
}

int dnaln_serve_create(dnaln_hip_ctx* ctx, const dnaln_serve_config* cfg, dnaln_server** out) {
  return dnaln_amd::c_guard([&]() -> int {
    if (!ctx || !out) return DNALN_E_INVALID;
    std::unique_ptr<dnaln_server> s(new dnaln_server);
    s->ctx = ctx;
    if (cfg) s->cfg = *cfg;
    else dnaln_serve_default_config(&s->cfg);
    if (s->cfg.slots < 1 || s->cfg.slots > 64 || s->cfg.max_batch_reads == 0 || s->cfg.max_batch_bases == 0 ||
        s->cfg.max_read_len < 1)
      return DNALN_E_INVALID;
    dnaln_map_params p;
    int rc = dnaln_hip_context_params(ctx, &p);
    if (rc != DNALN_OK) return rc;
    s->bw = p.sc.bw;
    s->max_cand = std::max(p.max_cand, 1);
    uint64_t free_b = 0, total_b = 0;
    rc = dnaln_hip_device_memory(ctx, &free_b, &total_b);
    if (rc != DNALN_OK) return rc;
    s->st.budget_bytes = s->cfg.hbm_budget ? s->cfg.hbm_budget : free_b / 2;
    if (!s->cfg.dir_budget) s->cfg.dir_budget = std::max<uint64_t>(s->st.budget_bytes / s->cfg.slots / 4, 1 << 24);
    for (int i = 0; i < s->cfg.slots; ++i) {
      // The server owns the slot from the start, so its destructor releases the batch and joins
      // the worker whichever step below fails.
      s->slots.push_back(std::make_unique<Slot>());
      Slot* slot = s->slots.back().get();
      dnaln_batch_config bc;
      dnaln_hip_default_batch_config(&bc);
      bc.mode = s->cfg.mode;
      bc.dir_budget = s->cfg.dir_budget;
      rc = dnaln_hip_batch_create(ctx, &bc, &slot->b);
      if (rc != DNALN_OK) return rc;
      rc = dnaln_hip_batch_reserve(slot->b, s->cfg.max_batch_reads, s->cfg.max_batch_bases, s->cfg.max_read_len, 64);
      if (rc != DNALN_OK) return rc;
      // host staging for the largest batch admitted (gather, download), mapped up front
      const uint64_t nr = s->cfg.max_batch_reads, nb = s->cfg.max_batch_bases;
      slot->codes.prefault(nb);
      slot->offs.prefault(nr);
      slot->lens.prefault(nr);
      slot->fnv.prefault(nr);
      slot->rec.prefault(nr);
      slot->cig.prefault(8 * nr + 64);
      slot->worker = std::thread([slot] { slot->work(); });
    }
    *out = s.release();
    return DNALN_OK;
  });
}

void dnaln_serve_destroy(dnaln_server* s) { delete s; }

uint64_t dnaln_serve_batch_bytes(const dnaln_server* s, uint64_t n, uint64_t bases, int32_t maxlen) {
  return s ? s->batch_bytes(n, bases, maxlen) : 0;
}

namespace {

// Checks a request and queues it, as a copy of the caller's arrays or as a view of them.
int submit_request(dnaln_server* s, uint64_t id, const uint8_t* codes, const uint64_t* offs, const int32_t* lens,
                   const uint32_t* fnv, uint64_t n, bool copy) {
  if (!s || (n && (!codes || !offs || !lens || !fnv))) return DNALN_E_INVALID;
  Request q;
  q.id = id;
  uint64_t end = 0;
  for (uint64_t i = 0; i < n; ++i) {
    if (lens[i] < 0) return DNALN_E_INVALID;
    end = std::max(end, offs[i] + (uint64_t)lens[i]);
    q.bases += (uint64_t)lens[i];
    q.maxlen = std::max(q.maxlen, lens[i]);
  }
  if (n > s->cfg.max_batch_reads || q.bases > s->cfg.max_batch_bases ||
      s->batch_bytes(n, q.bases, q.maxlen) > s->st.budget_bytes)
    return DNALN_E_CAPACITY;
  q.n = n;
  q.code_bytes = end;
  if (copy) {
    q.own_codes.assign(codes, codes + end);
    q.own_offs.assign(offs, offs + n);
    q.own_lens.assign(lens, lens + n);
    q.own_fnv.assign(fnv, fnv + n);
    q.codes = q.own_codes.data();
    q.offs = q.own_offs.data();
    q.lens = q.own_lens.data();
    q.fnv = q.own_fnv.data();
  } else {
    q.codes = codes;
    q.offs = offs;
    q.lens = lens;
    q.fnv = fnv;
    q.view = true;
  }
  s->queue.push_back(std::move(q));
  s->st.submitted += 1;
  return DNALN_OK;
}

}  // namespace

int dnaln_serve_submit(dnaln_server* s, uint64_t id, const uint8_t* codes, const uint64_t* offs, const int32_t* lens,
                       const uint32_t* fnv, uint64_t n) {
  return dnaln_amd::c_guard([&]() -> int {
    return submit_request(s, id, codes, offs, lens, fnv, n, true);
  });
}

int dnaln_serve_progress(dnaln_server* s) {
  return dnaln_amd::c_guard([&]() -> int {
    if (!s) return DNALN_E_INVALID;
    const int fin = s->collect();
    if (fin < 0) return fin;
    const int rc = s->admit();
    if (rc != DNALN_OK) return rc;
    return fin;
  });
}

int dnaln_serve_take(dnaln_server* s, uint64_t* id, dnaln_record* out, uint64_t out_cap, uint32_t* cigar_pool,
                     uint64_t cigar_cap, uint64_t* n_out, uint64_t* cigar_len_out) {
  if (!s || !id || !n_out) return DNALN_E_INVALID;
  if (s->done.empty()) return DNALN_E_PENDING;
  const Done& d = s->done.front();
  *id = d.id;
  *n_out = d.rec.size();
  if (cigar_len_out) *cigar_len_out = d.cig.size();
  if (!out) return DNALN_OK;
  if (out_cap < d.rec.size() || cigar_cap < d.cig.size() || (d.cig.size() && !cigar_pool)) return DNALN_E_CAPACITY;
  std::copy(d.rec.begin(), d.rec.end(), out);
  std::copy(d.cig.begin(), d.cig.end(), cigar_pool);
  s->done.pop_front();
  return DNALN_OK;
}

const char* dnaln_serve_last_error(const dnaln_server* s) { return s ? s->last_err.c_str() : ""; }

int dnaln_serve_stats_get(const dnaln_server* s, dnaln_serve_stats* out) {
  if (!s || !out) return DNALN_E_INVALID;
  *out = s->st;
  out->queued_requests = s->queue.size();
  return DNALN_OK;
}

int dnaln_serve_run_trace(dnaln_server* s, const uint8_t* codes, const uint64_t* offs, const int32_t* lens,
                          const uint32_t* fnv, uint64_t n_reads, const uint64_t* req_first, const uint64_t* req_count,
                          const double* arrival_s, uint64_t n_req, double* latency_s, dnaln_record* records,
                          uint32_t* cigar_pool, uint64_t cigar_cap, double* wall_s) {
  return dnaln_amd::c_guard([&]() -> int {
    if (!s || !req_first || !req_count || !arrival_s || !latency_s ||
        (n_reads && (!codes || !offs || !lens || !fnv)))
      return DNALN_E_INVALID;
    for (uint64_t r = 0; r < n_req; ++r)
      if (req_first[r] + req_count[r] > n_reads || (r && arrival_s[r] < arrival_s[r - 1])) return DNALN_E_INVALID;
    // Request offsets relative to each request's first base, for all requests up front: the
    // requests are queued as views of these arrays and of the caller's, which outlive the replay.
    std::vector<uint64_t> local_offs;
    std::vector<uint64_t> local_at(n_req + 1, 0);
    for (uint64_t r = 0; r < n_req; ++r) local_at[r + 1] = local_at[r] + req_count[r];
    local_offs.resize(local_at[n_req]);
    for (uint64_t r = 0; r < n_req; ++r) {
      const uint64_t a = req_first[r], n = req_count[r], base = n ? offs[a] : 0;
      for (uint64_t i = 0; i < n; ++i) local_offs[local_at[r] + i] = offs[a + i] - base;
    }
    uint64_t next = 0, finished = 0, cig_used = 0;
    const double t0 = now_s();
    double t_last = t0;
    while (finished < n_req) {
      const double t = now_s() - t0;
      while (next < n_req && arrival_s[next] <= t) {
        const uint64_t a = req_first[next], n = req_count[next];
        const uint64_t base = n ? offs[a] : 0;
        const int rc = submit_request(s, next, codes + base, local_offs.data() + local_at[next], lens + a, fnv + a, n,
                                      false);
        if (rc != DNALN_OK) {
          s->drop_views();
          return rc;
        }
        ++next;
      }
      const int rc = dnaln_serve_progress(s);
      if (rc < 0) {
        s->drop_views();
        return rc;
      }
      while (!s->done.empty()) {
        Done& d = s->done.front();
        const uint64_t r = d.id;
        latency_s[r] = (d.t_done - t0) - arrival_s[r];
        t_last = std::max(t_last, d.t_done);
        if (records) {
          const uint64_t a = req_first[r];
          for (size_t i = 0; i < d.rec.size(); ++i) {
            dnaln_record m = d.rec[i];
            if (cig_used + m.cigar_len > cigar_cap) {
              s->drop_views();
              return DNALN_E_CAPACITY;
            }
            std::copy(d.cig.begin() + m.cigar_off, d.cig.begin() + m.cigar_off + m.cigar_len, cigar_pool + cig_used);
            m.cigar_off = cig_used;
            cig_used += m.cigar_len;
            records[a + i] = m;
          }
        }
        s->done.pop_front();
        ++finished;
      }
      // Nothing running and the next arrival is later: wait for it without spinning hard.
      if (s->st.inflight_batches == 0 && s->queue.empty() && next < n_req) {
        const double wait = arrival_s[next] - (now_s() - t0);
        if (wait > 2e-4) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 1e-4));
      } else {
        std::this_thread::yield();
      }
    }
    if (wall_s) *wall_s = t_last - (t0 + (n_req ? arrival_s[0] : 0.0));
    return DNALN_OK;
  });
}

}  // extern "C"
