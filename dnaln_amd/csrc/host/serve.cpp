// serve.cpp — continuous batching of mapping requests on the resident index (dnaln_serve.h).
#include "dnaln_serve.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
  uint64_t id = 0;
  std::vector<uint8_t> codes;
  std::vector<uint64_t> offs;
  std::vector<int32_t> lens;
  std::vector<uint32_t> fnv;
  uint64_t bases = 0;
  int32_t maxlen = 0;
};

struct Done {
  uint64_t id = 0;
  double t_done = 0;
  std::vector<dnaln_record> rec;  // cigar_off relative to cig
  std::vector<uint32_t> cig;
};

// A batch slot and its worker thread. The main thread hands a batch over (state 1) and takes
// it back finished (state 2); only the worker touches the device batch in between.
struct Slot {
  dnaln_hip_batch* b = nullptr;
  std::mutex m;
  std::condition_variable cv;
  int state = 0;  // 0 free, 1 handed to the worker, 2 finished
  bool quit = false;
  int rc = 0;
  std::string err;
  double t_done = 0;
  std::thread worker;
  uint64_t est = 0;
  std::vector<uint64_t> req_id, req_first, req_n;
  std::vector<uint8_t> codes;
  std::vector<uint64_t> offs;
  std::vector<int32_t> lens;
  std::vector<uint32_t> fnv;
  std::vector<dnaln_record> rec;
  std::vector<uint32_t> cig;

  void work() {
    for (;;) {
      std::unique_lock<std::mutex> lk(m);
      cv.wait(lk, [&] { return state == 1 || quit; });
      if (state != 1) return;
      lk.unlock();
      int r = dnaln_hip_batch_upload(b, codes.data(), offs.data(), lens.data(), fnv.data(), lens.size());
      if (r == DNALN_OK) r = dnaln_hip_batch_run(b);
      if (r == DNALN_OK) r = dnaln_hip_batch_poll(b, 1);
      if (r == DNALN_OK) {
        uint64_t cap = 1;
        for (const int32_t l : lens) cap += 2 * (uint64_t)l + 4;
        rec.resize(lens.size());
        cig.resize(cap);
        r = dnaln_hip_batch_download(b, rec.data(), cig.data(), cig.size(), nullptr);
      }
      const std::string e = r == DNALN_OK ? std::string() : std::string(dnaln_hip_last_error());
      lk.lock();
      rc = r;
      err = e;
      t_done = now_s();
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
    const uint64_t G = (uint64_t)(2 * bw + 1 + 7) / 8;
    const uint64_t per_read = 16 + sizeof(dnaln_record) + 64 + (uint64_t)(max_cand + 1) * 96 + 64 * 48 + 64 * 8;
    const uint64_t dirs = std::min<uint64_t>(bases * 4 * G + (uint64_t)maxlen * G * 512, cfg.dir_budget);
    return bases + 256 + n * per_read + n * (2 * (uint64_t)maxlen + 4) * 4 + dirs;
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
        return s.rc;
      }
      for (size_t r = 0; r < s.req_id.size(); ++r) {
        Done out;
        out.id = s.req_id[r];
        out.t_done = s.t_done;
        out.rec.assign(s.rec.begin() + s.req_first[r], s.rec.begin() + s.req_first[r] + s.req_n[r]);
        for (dnaln_record& m : out.rec) {
          const uint64_t o = out.cig.size();
          out.cig.insert(out.cig.end(), s.cig.begin() + m.cigar_off, s.cig.begin() + m.cigar_off + m.cigar_len);
          m.cigar_off = o;
        }
        done.push_back(std::move(out));
        ++finished;
        ++st.completed;
      }
    }
    return finished;
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
        const uint64_t n2 = n + q.lens.size(), b2 = bases + q.bases;
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
      s.codes.clear();
      s.offs.clear();
      s.lens.clear();
      s.fnv.clear();
      s.req_id.clear();
      s.req_first.clear();
      s.req_n.clear();
      for (uint64_t r = 0; r < k; ++r) {
        Request& q = queue.front();
        s.req_id.push_back(q.id);
        s.req_first.push_back(s.lens.size());
        s.req_n.push_back(q.lens.size());
        const uint64_t base = s.codes.size();
        s.codes.insert(s.codes.end(), q.codes.begin(), q.codes.end());
        for (uint64_t o : q.offs) s.offs.push_back(base + o);
        s.lens.insert(s.lens.end(), q.lens.begin(), q.lens.end());
        s.fnv.insert(s.fnv.end(), q.fnv.begin(), q.fnv.end());
        queue.pop_front();
      }
      s.est = est;
      {
        std::lock_guard<std::mutex> lk(s.m);
        s.state = 1;
      }
      s.cv.notify_one();
      st.batches += 1;
      st.reads += s.lens.size();
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
  cfg->slots = 4;
  cfg->mode = DNALN_MODE_BATCHED;
  cfg->hbm_budget = 0;
  cfg->max_batch_reads = 1 << 17;
  cfg->max_batch_bases = 1ull << 26;
  cfg->dir_budget = 0;
  cfg->max_read_len = 1000;
}

int dnaln_serve_create(dnaln_hip_ctx* ctx, const dnaln_serve_config* cfg, dnaln_server** out) {
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
    auto slot = std::make_unique<Slot>();
    dnaln_batch_config bc;
    dnaln_hip_default_batch_config(&bc);
    bc.mode = s->cfg.mode;
    bc.dir_budget = s->cfg.dir_budget;
    rc = dnaln_hip_batch_create(ctx, &bc, &slot->b);
    if (rc != DNALN_OK) return rc;
    rc = dnaln_hip_batch_reserve(slot->b, s->cfg.max_batch_reads, s->cfg.max_batch_bases, s->cfg.max_read_len, 64);
    if (rc != DNALN_OK) {
      dnaln_hip_batch_destroy(slot->b);
      return rc;
    }
    Slot* raw = slot.get();
    slot->worker = std::thread([raw] { raw->work(); });
    s->slots.push_back(std::move(slot));
  }
  *out = s.release();
  return DNALN_OK;
}

void dnaln_serve_destroy(dnaln_server* s) { delete s; }

uint64_t dnaln_serve_batch_bytes(const dnaln_server* s, uint64_t n, uint64_t bases, int32_t maxlen) {
  return s ? s->batch_bytes(n, bases, maxlen) : 0;
}

int dnaln_serve_submit(dnaln_server* s, uint64_t id, const uint8_t* codes, const uint64_t* offs, const int32_t* lens,
                       const uint32_t* fnv, uint64_t n) {
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
  q.codes.assign(codes, codes + end);
  q.offs.assign(offs, offs + n);
  q.lens.assign(lens, lens + n);
  q.fnv.assign(fnv, fnv + n);
  s->queue.push_back(std::move(q));
  s->st.submitted += 1;
  return DNALN_OK;
}

int dnaln_serve_progress(dnaln_server* s) {
  if (!s) return DNALN_E_INVALID;
  const int fin = s->collect();
  if (fin < 0) return fin;
  const int rc = s->admit();
  if (rc != DNALN_OK) return rc;
  return fin;
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
  if (!s || !req_first || !req_count || !arrival_s || !latency_s ||
      (n_reads && (!codes || !offs || !lens || !fnv)))
    return DNALN_E_INVALID;
  for (uint64_t r = 0; r < n_req; ++r)
    if (req_first[r] + req_count[r] > n_reads || (r && arrival_s[r] < arrival_s[r - 1])) return DNALN_E_INVALID;
  std::vector<uint64_t> local_offs;
  uint64_t next = 0, finished = 0, cig_used = 0;
  const double t0 = now_s();
  double t_last = t0;
  while (finished < n_req) {
    const double t = now_s() - t0;
    while (next < n_req && arrival_s[next] <= t) {
      const uint64_t a = req_first[next], n = req_count[next];
      // A request's reads as their own buffer: offsets relative to its first base.
      const uint64_t base = n ? offs[a] : 0;
      local_offs.resize(n);
      for (uint64_t i = 0; i < n; ++i) local_offs[i] = offs[a + i] - base;
      const int rc = dnaln_serve_submit(s, next, codes + base, local_offs.data(), lens + a, fnv + a, n);
      if (rc != DNALN_OK) return rc;
      ++next;
    }
    const int rc = dnaln_serve_progress(s);
    if (rc < 0) return rc;
    while (!s->done.empty()) {
      Done& d = s->done.front();
      const uint64_t r = d.id;
      latency_s[r] = (d.t_done - t0) - arrival_s[r];
      t_last = std::max(t_last, d.t_done);
      if (records) {
        const uint64_t a = req_first[r];
        for (size_t i = 0; i < d.rec.size(); ++i) {
          dnaln_record m = d.rec[i];
          if (cig_used + m.cigar_len > cigar_cap) return DNALN_E_CAPACITY;
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
}

}  // extern "C"
