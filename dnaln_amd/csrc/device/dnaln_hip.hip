// dnaln_hip.hip — gfx950 kernels and the C ABI of include/dnaln_hip.h.
//
// Pipeline of one batch (src/mapper.cpp:75-223 restated as device stages):
//   seed_count -> scan -> seed_emit -> chain (+candidates) -> rank tasks -> packed SW score-only
//   -> select -> traceback tasks -> packed SW with flags + walk -> records -> scan -> CIGAR gather
// The reference codes, the index and all parameters stay resident in HBM for the life of the
// context; reads arrive per batch.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "dnaln_hip.h"
#include "bw_list.hpp"
#include "map_steps.hpp"
#include "seed_chain.hpp"
#include "sw_generic.hpp"
#include "sw_packed.hpp"
#include "sw_split.hpp"
#include "fused_seed_chain.hpp"
#include "chain_wave.hpp"
#include "index_build.hpp"

using namespace dn;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define DN_HIP(call)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail(DNALN_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));         \
  } while (0)

// Entry points run their bodies under guarded(): a C++ exception (a failed host allocation in a
// std::vector, say) becomes a status and a message instead of crossing the C ABI.
template <class F>
int guarded(F&& body) noexcept {
  const char* what = nullptr;
  int code = DNALN_E_INTERNAL;
  try {
    return body();
  } catch (const std::bad_alloc&) {
    code = DNALN_E_NOMEM;
    what = "host allocation failed";
  } catch (...) {
    what = "unexpected C++ exception";
  }
  try {
    g_err = what;
  } catch (...) {
  }
  return code;
}

#define DN_TRY(expr)          \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != DNALN_OK) return rc_; \
  } while (0)

// Grow-only device buffer.
// Device allocations made by DevBuf in this process (count, bytes).
std::atomic<uint64_t> g_alloc_events{0}, g_alloc_bytes{0};

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  // Grows the buffer, discarding its contents. Work still queued on the buffer must finish
  // first: `user` names the one stream that uses it (waited for), else the whole device is.
  int reserve(size_t bytes, hipStream_t user = nullptr) {
    if (bytes <= cap) return DNALN_OK;
    if (p) {
      (void)(user ? hipStreamSynchronize(user) : hipDeviceSynchronize());
      (void)hipFree(p);
      p = nullptr;
      cap = 0;
    }
#if defined(DN_EMU)
    const size_t want = std::max<size_t>(bytes, 1);  // exact in the CPU emulation: ASan sees the request
#else
    const size_t want = std::max<size_t>(bytes + bytes / 4, 256);
#endif
    if (hipMalloc(&p, want) != hipSuccess) {
      (void)hipGetLastError();
      return fail(DNALN_E_NOMEM, "hipMalloc of " + std::to_string(want) + " bytes failed");
    }
    cap = want;
    g_alloc_events.fetch_add(1, std::memory_order_relaxed);
    g_alloc_bytes.fetch_add(want, std::memory_order_relaxed);
    return DNALN_OK;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Receiver of timing spans (event pairs on the launching stream) around kernel launches.
struct SpanSink {
  virtual ~SpanSink() = default;
  virtual int begin(int family, size_t& idx) = 0;
  virtual int end(size_t idx) = 0;
};

// Timing families: 0 seed, 1 chain, 2 rank fill, 3 traceback fill, 4 traceback walk, 5 run.
constexpr int kFamilies = 6;

// Traceback directions per SW launch: enough fill waves to occupy every SIMD several times at
// W = 129 and 1000-bp reads (8.7 MB per wave), within an eighth of the free HBM.
uint64_t default_dir_budget() {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    (void)hipGetLastError();
    return 1ull << 30;
  }
  return std::min<uint64_t>(16ull << 30, std::max<uint64_t>(1ull << 30, free_b / 8));
}

// Grow-only device buffers of run_sw, and the host copy of the per-wave slot offsets.
struct SwWork {
  DevBuf dir, rows, keys, keys2, ids, ids2, woff, tmp;
  std::vector<uint64_t> woff_h;
  uint64_t dir_budget = 1ull << 30;  // bytes of direction words per launch
  int walk_kind = 0;                 // launch_walk_bw
  int lanes = 1;                     // lanes per pair of the packed fill (split_available)
};

// Layouts shared with other languages' bindings.
static_assert(sizeof(dnaln_map_params) == 80, "dnaln_map_params layout");
static_assert(sizeof(dnaln_batch_config) == 64, "dnaln_batch_config layout");
static_assert(sizeof(dnaln_map_stats) == 112, "dnaln_map_stats layout");
static_assert(sizeof(dnaln_record) == 40 && sizeof(dnaln_result) == 40, "record/result layout");

constexpr int kSeedBlock = 128;
constexpr int kFusedBlock = 64;  // one wave per block: LDS is the occupancy limit
// Guard bytes before and after read and task pools (the packed kernel loads aligned dwords
// that may reach up to 4 bytes past a sequence).
constexpr uint64_t kReadGuard = 64;
constexpr int kSwBlock = 256;  // 4 waves, 512 alignments per block

// ------------------------------------------------------------------------------------ kernels
__device__ inline int wave_max(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = max(v, __shfl_xor(v, m, 64));
  return v;
}

__global__ __launch_bounds__(kSeedBlock) void k_seed_count(SeedArgs a, uint64_t* ring_h_g, uint32_t* ring_ps_g,
                                                           int use_lds) {
  DN_DYN_LDS(uint64_t, lds);
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.n_reads) return;
  Ring ring;
  if (use_lds) {
    ring.h = lds + threadIdx.x;
    ring.ps = reinterpret_cast<uint32_t*>(lds + (size_t)a.w * blockDim.x) + threadIdx.x;
    ring.stride = blockDim.x;
  } else {
    ring.h = ring_h_g + r;
    ring.ps = ring_ps_g + r;
    ring.stride = (int)a.n_reads;
  }
  a.anchor_count[r] = seed_count_read(a, r, ring);
}

__global__ __launch_bounds__(kSeedBlock) void k_seed_emit(SeedArgs a, uint64_t* ring_h_g, uint32_t* ring_ps_g,
                                                          int use_lds) {
  DN_DYN_LDS(uint64_t, lds);
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.n_reads) return;
  Ring ring;
  if (use_lds) {
    ring.h = lds + threadIdx.x;
    ring.ps = reinterpret_cast<uint32_t*>(lds + (size_t)a.w * blockDim.x) + threadIdx.x;
    ring.stride = blockDim.x;
  } else {
    ring.h = ring_h_g + r;
    ring.ps = ring_ps_g + r;
    ring.stride = (int)a.n_reads;
  }
  seed_emit_read(a, r, ring);
}

__global__ __launch_bounds__(kSeedBlock) void k_chain(ChainArgs a) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.n_reads) return;
  chain_read(a, r);
}

// One wave per read (chain_wave.hpp), four waves per block. LDS: the gap-cost table, then per
// wave kWaveChainLds anchors' keys, sort keys, f, prev and chains. Reads with more
// anchors, or min_anchors < 2 (more chains than the LDS list holds), work in the global
// scratch arrays of the chain stage instead. The stage is bound by VALU issue and LDS latency,
// so occupancy pays: 128 anchors keep a block at the default band near 16 KiB, so a CU's
// 160 KiB holds the LDS of nine blocks; reads with more anchors use the global arrays.
constexpr int kWaveChainLds = 128;
constexpr size_t kWaveChainBytes = (size_t)32 * kWaveChainLds;

size_t wave_chain_lds_bytes(int band) { return (size_t)((band + 4) & ~3) * 4 + 4 * kWaveChainBytes; }

__global__ __launch_bounds__(256) void k_chain_wave(ChainArgs a) {
  DN_DYN_LDS(uint64_t, lds);
  int32_t* cost = reinterpret_cast<int32_t*>(lds);
  for (int x = threadIdx.x; x <= a.band; x += blockDim.x) cost[x] = a.cost[x];
  __syncthreads();
  // The wave index as a scalar, so the read's anchor count and the loops over it are uniform.
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const uint32_t li = blockIdx.x * 4 + wv;
  if (li >= a.n_reads) return;
  const uint32_t r = read_id(a.list, li);
  const int n = (int)a.anchor_count[li];
  const uint64_t off = a.anchor_off[li];
  WaveChainMem s;
  if (n <= kWaveChainLds && a.min_anchors >= 2) {
    uint8_t* base = reinterpret_cast<uint8_t*>(cost + ((a.band + 4) & ~3)) + wv * kWaveChainBytes;
    s.keys = reinterpret_cast<uint64_t*>(base);
    s.order = s.keys + kWaveChainLds;
    s.f = reinterpret_cast<int32_t*>(s.order + kWaveChainLds);
    s.prev = reinterpret_cast<uint32_t*>(s.f + kWaveChainLds);
    s.chains = s.f + 2 * kWaveChainLds;
    for (int x = lane; x < n; x += 64) s.keys[x] = a.keys[off + x];
    wave_sync();
    chain_wave_read(a, r, n, s, cost, lane);
  } else {
    s.keys = a.keys + off;
    s.order = a.order + off;
    s.f = a.f + off;
    s.prev = reinterpret_cast<uint32_t*>(a.prev + off);
    s.chains = reinterpret_cast<int32_t*>(a.chains + off);
    chain_wave_read(a, r, n, s, cost, lane);
  }
}


// One read per thread; the read's anchor words sit in LDS, `cap` u64 slots per thread,
// interleaved by thread. The minimizer list goes to the block's part of `mins` (maxmin u64 per
// thread, interleaved: written by the sketch, read back by the probes). The fallback sketch's
// ring (w u64 hashes + w u32 positions) uses the anchor slots before any anchor is written.
template <int WT, bool P8, int CAP>
__global__ __launch_bounds__(kFusedBlock) DN_WAVES_PER_EU(CAP > 32 ? 2 : 3) void k_seed_chain_fused(SeedArgs s,
                                                                                                  ChainArgs c, int maxmin,
                                                                  uint64_t* mins, uint32_t* ovf, uint32_t* novf) {
  DN_DYN_LDS(uint64_t, lds);
  const int t = threadIdx.x;
  const uint32_t i = blockIdx.x * kFusedBlock + t;
  if (i >= s.n_reads) return;
  uint64_t* keys = lds;
  Ring ring;
  ring.h = keys + t;
  ring.ps = reinterpret_cast<uint32_t*>(keys + (size_t)s.w * kFusedBlock) + t;
  ring.stride = kFusedBlock;
  LdsView v;
  v.keys = keys + t;
  v.stride = kFusedBlock;
  const uint32_t r = read_id(s.list, i);
  uint64_t* my_mins = mins + (size_t)blockIdx.x * maxmin * kFusedBlock + t;
  const int nc = seed_chain_lds<8, 10, P8, CAP>(s, c, r, v, my_mins, kFusedBlock, maxmin, ring);
  if (nc < 0) {
    ovf[atomicAdd(novf, 1u)] = r;
    c.n_cand[r] = 0;
  } else {
    c.n_cand[r] = nc;
  }
}

struct ReadView {
  const uint8_t* codes;
  const uint64_t* offs;
  const int32_t* lens;
  const uint32_t* fnv;
};

// Rank tasks: every candidate of reads with at least two (src/mapper.cpp:124-130).
__global__ void k_rank_count(const int32_t* n_cand, uint32_t* cnt, uint32_t n) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) cnt[r] = n_cand[r] >= 2 ? (uint32_t)n_cand[r] : 0u;
}

__global__ void k_rank_tasks(ReadView rv, MapGeom g, const Cand* cands, const int32_t* n_cand, int max_cand,
                             const uint64_t* rank_off, SwTask* tasks, uint32_t n) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int nc = n_cand[r];
  if (nc < 2) return;
  const uint64_t q = (uint64_t)(uintptr_t)(rv.codes + rv.offs[r]);
  for (int c = 0; c < nc; ++c)
    tasks[rank_off[r] + c] = make_task(g, q, rv.lens[r], cands[(uint64_t)r * max_cand + c]);
}

__global__ void k_select(ReadView rv, const int32_t* n_cand, const uint64_t* rank_off, const SwResult* rank_res,
                         int32_t* chosen, int32_t* second, uint32_t* tbflag, uint32_t n) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int nc = n_cand[r];
  const SwResult* res = rank_res + (nc >= 2 ? rank_off[r] : 0);
  int32_t ch, s2;
  select_read(nc, [res](int c) { return res[c].score; }, rv.fnv[r], ch, s2);
  chosen[r] = ch;
  second[r] = s2;
  tbflag[r] = ch >= 0 ? 1u : 0u;
}

__global__ void k_tb_tasks(ReadView rv, MapGeom g, const Cand* cands, int max_cand, const int32_t* chosen,
                           const uint64_t* tb_off, SwTask* tasks, uint32_t n) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n || chosen[r] < 0) return;
  const uint64_t q = (uint64_t)(uintptr_t)(rv.codes + rv.offs[r]);
  tasks[tb_off[r]] = make_task(g, q, rv.lens[r], cands[(uint64_t)r * max_cand + chosen[r]]);
}

__global__ void k_records(ReadView rv, MapGeom g, const Cand* cands, const int32_t* n_cand, int max_cand,
                          const int32_t* chosen, const int32_t* second, const uint64_t* tb_off,
                          const SwResult* tb_res, int min_aln_score, const uint8_t* lut, int lut_len,
                          dnaln_record* rec, uint32_t* cig_len, int* err, uint32_t n) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int ch = chosen[r];
  const SwResult* res = ch >= 0 ? &tb_res[tb_off[r]] : nullptr;
  if (res && (res->n_cigar < 0 || res->status != 0)) atomicOr(err, res->n_cigar < 0 ? 1 : 2);
  dnaln_record m;
  record_read(g, rv.lens[r], n_cand[r], cands + (uint64_t)r * max_cand, ch, second[r], res, min_aln_score, lut,
              lut_len, m);
  if (res && res->n_cigar < 0) m.cigar_len = 0;
  rec[r] = m;
  cig_len[r] = m.cigar_len;
}

__global__ void k_cigar_gather(ReadView rv, const int32_t* chosen, const uint64_t* tb_off, const SwResult* tb_res,
                               const uint32_t* slots, int cap, const uint64_t* cig_off, uint64_t pool_base,
                               dnaln_record* rec, uint32_t* pool, uint32_t n) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  rec[r].cigar_off = pool_base + cig_off[r];
  if (rec[r].cigar_len == 0) return;
  const uint64_t t = tb_off[r];
  gather_cigar(rv.lens[r], tb_res[t], slots + t * (uint64_t)cap, pool + cig_off[r]);
}

// Waves per SIMD the fill must keep: two for the traceback fills through W = 65. Up to W = 49
// they fit in 256 registers, and the bound keeps scheduling changes from silently halving their
// occupancy. W = 65 spills 44 registers to scratch in the once-per-four-rows block section and
// still gains 4% (25.6 against 26.7 ms per launch on the 4M x 250 bp repeat workload).
constexpr int sw_min_waves(int W, bool TB) { return TB && W <= 65 ? 2 : 1; }

template <int W, bool TB>
__global__ __launch_bounds__(kSwBlock) DN_WAVES_PER_EU(sw_min_waves(W, TB)) void k_sw_packed(
    SwArgs a) {
  const int lane = threadIdx.x & 63;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t base = (uint64_t)wave * 128;
  if (base >= a.n) return;
  const uint32_t ka = (uint32_t)base + lane, kb = (uint32_t)base + 64 + lane;
  const bool va = ka < a.n, vb = kb < a.n;
  const uint32_t ida = va ? (a.order ? a.order[ka] : ka) : 0u;
  const uint32_t idb = vb ? (a.order ? a.order[kb] : kb) : 0u;
  int ql = max(va ? a.tasks[ida].qlen : 0, vb ? a.tasks[idb].qlen : 0);
  ql = wave_max(ql);
  uint64_t* slot = TB ? dir_slot(a, wave) : nullptr;
  if (TB && W > kBestRowInRegs) {
    DN_DYN_LDS(uint16_t, bh_rows);  // per wave of the block: W x 64 lanes x {A, B}
    uint16_t* bh = bh_rows + (size_t)(threadIdx.x >> 6) * best_row_lds_words(W) * 2 + 2 * lane;
    sw_pair<W, TB>(a, ida, idb, va, vb, ql, slot, lane, bh, 64);
  } else if (TB && flag_granules<W, 1>()) {
    DN_DYN_LDS(uint32_t, stage);  // per wave of the block: one step's flag words, lane-interleaved
    sw_pair<W, TB>(a, ida, idb, va, vb, ql, slot, lane, nullptr, 0,
                   stage + (size_t)(threadIdx.x >> 6) * fill_lds_words(W) + lane, 64);
  } else {
    sw_pair<W, TB>(a, ida, idb, va, vb, ql, slot, lane);
  }
}

// The packed fill split over P lanes per pair (sw_split.hpp): fill group g (tasks 128g ..
// 128g + 127, one direction slot) runs as waves gP .. gP + P-1; wave j holds pairs
// j*64/P .. (j+1)*64/P - 1, lane l segment l % P of pair j*64/P + l / P. Each wave keeps its
// lanes' best rows (S columns x 64 lanes x {A, B}) and one step's flag words (split_stage_words)
// in LDS. Two waves per SIMD: at S = 33 the traceback fill takes about 224 registers; the
// score-only fill held to three waves spills 17.
template <int W, int P, bool TB>
__global__ __launch_bounds__(kSwBlock) DN_WAVES_PER_EU(2) void k_sw_split(SwArgs a) {
  constexpr int S = seg_width(W, P);
  const int lane = threadIdx.x & 63;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t grp = wave / P, j = wave % P;
  const uint64_t base = (uint64_t)grp * 128;
  if (base >= a.n) return;
  const uint32_t ka = (uint32_t)base + j * (64 / P) + lane / P, kb = ka + 64;
  const bool va = ka < a.n, vb = kb < a.n;
  const uint32_t ida = va ? (a.order ? a.order[ka] : ka) : 0u;
  const uint32_t idb = vb ? (a.order ? a.order[kb] : kb) : 0u;
  int ql = max(va ? a.tasks[ida].qlen : 0, vb ? a.tasks[idb].qlen : 0);
  ql = wave_max(ql);
  uint64_t* slot = TB ? dir_slot(a, grp) : nullptr;
  uint16_t* bh = nullptr;
  uint32_t* stage = nullptr;
  if (TB) {
    DN_DYN_LDS(uint32_t, split_lds);  // best rows of the block's waves, then their flag staging
    bh = reinterpret_cast<uint16_t*>(split_lds + (size_t)(threadIdx.x >> 6) * S * 64) + 2 * lane;
    stage = split_lds + (size_t)(kSwBlock / 64) * S * 64 + (size_t)(threadIdx.x >> 6) * split_stage_words(W, P) + lane;
  }
  sw_split<W, P, TB>(a, ida, idb, va, vb, ql, slot, (int)j, lane, bh, 64, stage, 64);
}

// Traceback with one thread per task (walk_task), fills with P lanes per pair.
template <int W, int P = 1>
__global__ __launch_bounds__(256) void k_walk(SwArgs a) {
  __shared__ uint32_t seqbuf[kWalkSeqDwords * 256];
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  WalkSeq seq(seqbuf + threadIdx.x, 256);
  walk_task<W, P>(a, k, seq);
}

// Traceback of the 128 tasks of one fill wave by one block, through an LDS window of `rows`
// rows of the wave's flag words: the block loads rows [top - rows + 1, top] (top = the highest
// row any unfinished walk is on) with coalesced 16-byte loads, every walk steps while its row
// is inside, and the window moves down. Each slot is read once, instead of a cache line per
// step per task. Thread t walks task wave * 128 + t, the task the fill ran in lane t % 64, half
// t / 64.
template <int W>
__global__ __launch_bounds__(128) void k_walk_win(SwArgs a, int rows) {
  constexpr int G = (W + 7) / 8;
  constexpr bool kGran = flag_granules<W, 1>();
  DN_DYN_LDS(uint64_t, win);
  __shared__ int s_top;
  __shared__ uint32_t seqbuf[kWalkSeqDwords * 128];
  const int t = threadIdx.x, lane = t & 63, half = t >> 6;
  const uint32_t wave = blockIdx.x;
  const uint32_t k = wave * 128 + t;
  bool active = false;
  uint32_t id = 0;
  int score = 0, i_end = 0, d_end = 0;
  SwTask tk{};
  Walk w{};
  WalkSeq seq(seqbuf + t, 128);
  uint32_t* out = nullptr;
  if (k < a.n) {
    id = a.order ? a.order[k] : k;
    SwResult& r = a.res[id];
    score = r.score;
    i_end = r.qe;
    d_end = r.te;
    if (score <= 0) {
      r.qb = r.qe = r.tb = r.te = r.nm = r.n_cigar = r.status = 0;
    } else {
      tk = a.tasks[id];
      seq.begin(tk);
      out = a.cigar + (uint64_t)id * (uint64_t)a.cigar_cap;
      w = walk_init(i_end, d_end, score);
      active = true;
    }
  }
  const uint64_t* slot = dir_slot(a, wave);
  for (;;) {
    if (t == 0) s_top = -1;
    __syncthreads();
    if (active && !w.done) atomicMax(&s_top, w.i);
    __syncthreads();
    const int top = s_top;
    if (top < 0) break;
    // Flag granules: the window covers whole granules, [lo, top | 3] with lo a multiple of 4
    // (rows is a multiple of 4, and slots hold whole granules).
    const int lo = kGran ? max(0, (top - rows + 4) & ~3) : max(0, top - rows + 1);
    const int hi = kGran ? (top | 3) : top;
    const uint64_t n_words = (uint64_t)(hi - lo + 1) * G * 64;
    const uint64_t* src = slot + (uint64_t)lo * G * 64;
    for (uint64_t x = 2 * (uint64_t)t; x < n_words; x += 256) {
      const uint4 v = *reinterpret_cast<const uint4*>(src + x);
      *reinterpret_cast<uint4*>(win + x) = v;
    }
    __syncthreads();
    while (active && !w.done && (w.i >= lo || w.i < 0))
      walk_step<W>(w, tk, reinterpret_cast<const uint32_t*>(win) + (kGran ? (half * 64 + lane) * 4 : half * 64 + lane),
                   a.sp, out, a.cigar_cap, seq, lo);
    __syncthreads();
  }
  if (active) walk_finish(w, i_end, d_end, score, a.cigar_cap, a.res[id]);
}

__global__ void k_sw_generic(SwGenericArgs a) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= a.n) return;
  sw_generic_one(a, a.order ? a.order[tid] : tid, tid);
}

// The device copy of the index: a bucket whose minimizer occurs once holds that occurrence
// itself in `offset`, so a probe of a unique minimizer is one random access instead of two.
__global__ void k_inline_unique(Bucket* table, uint64_t n, const uint32_t* pos) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Bucket b = table[i];
  if (b.key != ~0ull && b.count == 1) table[i].offset = pos[b.offset];
}

// The 8-byte probe table (probe_table.hpp) from the inlined 16-byte table: words of pos2 per
// bucket (entry n is 0, so the exclusive scan ends with the total), then the entries, the
// mirrored tail and the blocks of pos2.
__global__ void k_ct_need(const Bucket* table, uint64_t n, uint32_t max_occ, uint32_t* need) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  need[i] = i < n ? c_need(table[i], max_occ) : 0u;
}

__global__ void k_ct_fill(const Bucket* table, uint64_t n, uint32_t max_occ, const uint32_t* pos, const uint64_t* at,
                          uint64_t* ctab, uint32_t* pos2) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n + kCTail) return;
  const uint64_t slot = i < n ? i : (i - n) & (n - 1);
  const Bucket b = table[slot];
  ctab[i] = c_entry(b, max_occ, b.offset, (uint32_t)at[slot]);  // count 1: offset is the occurrence
  if (i >= n || !c_need(b, max_occ)) return;
  uint32_t* out = pos2 + at[i];
  out[0] = b.count;
  for (uint32_t j = 0; j < b.count; ++j) out[1 + j] = pos[b.offset + j];
}

// ---- index build on the device (src/index.cpp:10-45): segments sketched one per thread, the
// occurrences radix-sorted by (hash, pos), each key placed by one thread with atomic linear
// probing. Probe chains only ever grow, so every key stays reachable from its home slot.
__global__ void k_ix_count(const uint8_t* ref, const IndexSeg* segs, uint32_t nseg, int k, int w, uint64_t* rh,
                           uint32_t* rps, uint64_t* counts) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nseg) return;
  uint64_t n = 0;
  if (i < nseg)
    sketch_segment(ref, segs[i], k, w, Ring{rh + i, rps + i, (int)nseg}, [&](uint64_t, uint32_t, uint32_t) { ++n; });
  counts[i] = n;  // entry nseg is 0, so the exclusive scan ends with the total
}

__global__ void k_ix_emit(const uint8_t* ref, const IndexSeg* segs, uint32_t nseg, int k, int w, uint64_t* rh,
                          uint32_t* rps, const uint64_t* offs, uint64_t* keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nseg) return;
  const IndexSeg g = segs[i];
  uint64_t* out = keys + offs[i];
  sketch_segment(ref, g, k, w, Ring{rh + i, rps + i, (int)nseg},
                 [&](uint64_t h, uint32_t p, uint32_t st) { *out++ = occ_key(h, g.base, p, st); });
}

// flags[i] = 1 where a new hash starts (entry n is 0); pos[i] = the occurrence's position word.
__global__ void k_ix_flags(const uint64_t* keys, uint64_t n, uint64_t* flags, uint32_t* pos) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  if (i == n) {
    flags[i] = 0;
    return;
  }
  const uint64_t k = keys[i];
  flags[i] = i == 0 || (keys[i - 1] >> 32) != (k >> 32);
  pos[i] = (uint32_t)k;
}

// run_start[r] = first occurrence of the r-th distinct hash; run_start[n_keys] = n.
__global__ void k_ix_starts(const uint64_t* flags, const uint64_t* ridx, uint64_t n, uint32_t* run_start) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  if (i == n) run_start[ridx[n]] = (uint32_t)n;
  else if (flags[i]) run_start[ridx[i]] = (uint32_t)i;
}

__global__ void k_ix_clear(Bucket* table, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) table[i] = Bucket{~0ull, 0, 0};
}

// One key per thread; the load factor stays <= 0.5, so every probe ends.
__global__ void k_ix_insert(const uint64_t* keys, const uint32_t* run_start, uint64_t n_keys, Bucket* table,
                            uint64_t mask, uint32_t* max_count) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_keys) return;
  const uint32_t off = run_start[r], cnt = run_start[r + 1] - off;
  const unsigned long long h = keys[off] >> 32;
  for (uint64_t s = h & mask;; s = (s + 1) & mask) {
    if (atomicCAS(reinterpret_cast<unsigned long long*>(&table[s].key), ~0ull, h) == ~0ull) {
      table[s].offset = off;
      table[s].count = cnt;
      break;
    }
  }
  atomicMax(max_count, cnt);
}

__global__ void k_pool_tasks(const dnaln_task* in, uint64_t pool, int bw, SwTask* out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const dnaln_task t = in[i];
  SwTask s;
  s.q = pool + t.q_off;
  s.t = pool + t.t_off;
  s.qlen = t.qlen;
  s.tlo = 0;
  const int wl = t.qlen + 2 * bw;
  s.thi = t.tlen < wl ? t.tlen : wl;
  if (s.thi < 0) s.thi = 0;
  s.flags = 0;
  out[i] = s;
}

// Threads per block of the packed fill: four waves, or one where each wave's best row sits in
// LDS (so a block stays within 64 KiB of LDS).
constexpr int sw_block(int W, bool TB) { return TB && W > kBestRowInRegs ? 64 : kSwBlock; }

// grid: blocks of kSwBlock threads.
template <bool TB>
bool launch_packed_bw(int bw, const SwArgs& a, unsigned grid, hipStream_t st) {
  switch (bw) {
#define DN_CASE(B)                                                                    \
  case B:                                                                             \
    hipLaunchKernelGGL((k_sw_packed<2 * B + 1, TB>), dim3(grid * (kSwBlock / sw_block(2 * B + 1, TB))),  \
                       dim3(sw_block(2 * B + 1, TB)),                                                     \
                       TB ? (size_t)fill_lds_words(2 * B + 1) * 4 * (sw_block(2 * B + 1, TB) / 64) : 0,     \
                       st, a);                                                                            \
    return true;
    DN_PACKED_BW_LIST(DN_CASE)
#undef DN_CASE
    default:
      return false;
  }
}

// True when band half-width bw has a fill split over `lanes` lanes per pair.
constexpr bool split_available(int bw, int lanes) {
#define DN_CASE(B, P) (bw == B && lanes == P) ||
  return DN_SPLIT_LIST(DN_CASE) false;
#undef DN_CASE
}

// The split fill: groups = 128-task fill groups of the launch.
template <bool TB>
bool launch_split_bw(int bw, int lanes, const SwArgs& a, unsigned groups, hipStream_t st) {
#define DN_CASE(B, P)                                                                                       \
  if (bw == B && lanes == P) {                                                                              \
    constexpr int kW = 2 * B + 1;                                                                           \
    hipLaunchKernelGGL((k_sw_split<kW, P, TB>), dim3((groups * P + kSwBlock / 64 - 1) / (kSwBlock / 64)),   \
                       dim3(kSwBlock),                                                                      \
                       TB ? (size_t)(seg_width(kW, P) * 64 + split_stage_words(kW, P)) * 4 * (kSwBlock / 64) : 0, \
                       st, a);                                                                              \
    return true;                                                                                            \
  }
  DN_SPLIT_LIST(DN_CASE)
#undef DN_CASE
  return false;
}

// Lanes per pair of the packed fill when the batch config leaves it open: the split where the
// band width has one, for the batched form.
int default_sw_lanes(int bw, int mode) {
  if (mode != DNALN_MODE_BATCHED) return 1;
  for (const int lanes : {2, 4})
    if (split_available(bw, lanes)) return lanes;
  return 1;
}

// The walk after a split fill (one thread per task).
bool launch_walk_split(int bw, int lanes, const SwArgs& a, hipStream_t st) {
#define DN_CASE(B, P)                                                                                  \
  if (bw == B && lanes == P) {                                                                         \
    hipLaunchKernelGGL((k_walk<2 * B + 1, P>), dim3((a.n + 255) / 256), dim3(256), 0, st, a);          \
    return true;                                                                                       \
  }
  DN_SPLIT_LIST(DN_CASE)
#undef DN_CASE
  return false;
}

// Rows per LDS window of the walk: 16 KiB of flag words, a multiple of 4 rows (whole flag
// granules), at least 4.
int walk_rows(int W) { return std::max(4, (16384 / (((W + 7) / 8) * 512)) & ~3); }

// walk_kind 0: one thread per task (k_walk); 1: one block per fill wave through LDS windows
// (k_walk_win).
bool launch_walk_bw(int bw, const SwArgs& a, int walk_kind, hipStream_t st) {
  const int W = 2 * bw + 1;
  const int rows = walk_rows(W);
  const size_t lds = (size_t)rows * ((W + 7) / 8) * 512;
  switch (bw) {
#define DN_CASE(B)                                                                                       \
  case B:                                                                                                \
    if (walk_kind == 1)                                                                                  \
      hipLaunchKernelGGL((k_walk_win<2 * B + 1>), dim3((a.n + 127) / 128), dim3(128), lds, st, a, rows); \
    else                                                                                                 \
      hipLaunchKernelGGL((k_walk<2 * B + 1>), dim3((a.n + 255) / 256), dim3(256), 0, st, a);             \
    return true;
    DN_PACKED_BW_LIST(DN_CASE)
#undef DN_CASE
    default:
      return false;
  }
}

bool packed_bw_available(int bw) {
  switch (bw) {
#define DN_CASE(B) case B:
    DN_PACKED_BW_LIST(DN_CASE)
#undef DN_CASE
    return true;
    default:
      return false;
  }
}

}  // namespace

#include "batched_graph.hpp"

// ------------------------------------------------------------------------------------ context
// Sizes of a device index, and the wall time of its GPU build (0 when uploaded from the host).
struct IndexInfo {
  uint64_t n_buckets = 0, n_pos = 0, n_keys = 0;
  uint32_t max_count = 0;
  double ms = 0;
};

struct dnaln_hip_ctx {
  dnaln_map_params p{};
  IndexInfo ix;
  int device = 0;
  hipStream_t stream = nullptr;
  DevBuf ref, coff, clen, table, pos, cost, lut;
  DevBuf ctab, pos2;            // the 8-byte probe table and its occurrence blocks (probe_table.hpp)
  bool probe8 = false;          // the seed probes use ctab / pos2
  uint8_t* ref_base = nullptr;  // reference codes, kSeqGuard bytes into `ref`
  std::vector<uint64_t> h_coff;
  std::vector<int32_t> h_cost;  // chain gap cost by drift, 0..band
  std::vector<uint32_t> h_clen;
  uint64_t ref_len = 0, mask = 0, n_pos = 0;
  int nc = 0, k = 0, w = 0, lut_len = 0;
  // workspace of dnaln_hip_align
  DevBuf a_pool, a_in, a_tasks, a_res, a_cig;
  SwWork a_sw;
  dnaln_hip_batch* batch = nullptr;  // used by dnaln_hip_map_batch
  uint8_t* scratch = nullptr;         // device address of this GPU's static scratch array
  ~dnaln_hip_ctx() {
    if (stream) (void)hipStreamDestroy(stream);  // also when create fails part-way
  }
};

namespace {

int validate(const dnaln_map_params& p) {
  if (p.k < 4 || p.k > 28) return fail(DNALN_E_INVALID, "k must be in [4, 28]");
  if (p.w < 1 || p.w > 255) return fail(DNALN_E_INVALID, "w must be in [1, 255]");
  if (p.sc.bw < 1 || p.sc.bw > 1000) return fail(DNALN_E_INVALID, "--bw must be in [1, 1000]");
  if (p.sc.match < 1 || p.sc.mismatch < 0 || p.sc.gap_open < 0 || p.sc.gap_ext < 1)
    return fail(DNALN_E_INVALID, "scores must satisfy A >= 1, B >= 0, O >= 0, E >= 1");
  if (p.sc.ambig < 0) return fail(DNALN_E_INVALID, "ambig must be >= 0");
  if (p.max_cand < 0 || p.ch.lookback < 0 || p.ch.band < 0) return fail(DNALN_E_INVALID, "negative parameter");
  return DNALN_OK;
}


// Runs a score-only or traceback fill over tasks[0, n) (device arrays), results in res[n],
// reversed CIGARs in cig[n * cap]. Chunks the traceback launches so the direction scratch stays
// under `dir_budget` bytes.

__global__ void k_iota(uint32_t* out, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (uint32_t)i;
}

__global__ void k_len_keys(const SwTask* tasks, uint32_t n, uint32_t* keys, uint32_t* ids) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = (uint32_t)tasks[i].qlen;
  ids[i] = i;
}

// Slot words of every fill group, from its longest query (the first of its 128 tasks, which are
// sorted longest first) and the fill's lanes per pair; entry `waves` is 0 so that the exclusive
// scan ends with the total.
__global__ void k_wave_words(const uint32_t* sorted_len, uint32_t waves, uint64_t row_words, int lanes,
                             uint64_t* words) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w > waves) return;
  words[w] = w < waves ? slot_rows((int)sorted_len[(uint64_t)w * 128], lanes) * row_words : 0;
}

int scan_u64(DevBuf& tmp, const uint64_t* in, uint64_t* out, uint32_t n, hipStream_t st) {
  size_t bytes = 0;
  DN_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, st));
  DN_TRY(tmp.reserve(bytes, st));
  DN_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, n, st));
  return DNALN_OK;
}

// Fills (and with traceback walks) n tasks. Query lengths lie in [qmin, qmax]. When they differ,
// the tasks are sorted longest first on the device, so each wave of 128 runs to nearly its own
// tasks' length, and each wave's direction slot is sized by its longest query. Launches are cut
// at wave boundaries so that one launch's direction words fit wk.dir_budget. With wk.lanes > 1
// (where the band width has that split) each group of 128 tasks runs as wk.lanes waves of the
// split fill, walked one thread per task.
int run_sw(const ScoreP& sp, const SwTask* tasks, uint32_t n, int qmin, int qmax, bool traceback, SwResult* res,
           uint32_t* cig, int cap, SwWork& wk, hipStream_t st, uint64_t* launches, SpanSink* spans = nullptr,
           int fill_family = 0) {
  if (n == 0) return DNALN_OK;
  const int W = 2 * sp.bw + 1;
  const bool packed = packed_bw_available(sp.bw) && packed_ok(sp, qmax);
  const uint64_t budget = wk.dir_budget;
  if (packed) {
    SwArgs a{};
    a.sp = sp;
    score_table(sp, a.tab0, a.tab1);
    a.cigar_cap = cap;
    // The score-only fill stays unsplit up to W = 65, where it keeps two waves per SIMD without
    // spills and runs faster than split (4.66 against 4.83 ms per step on the 4M x 250 bp repeat
    // workload); wider, unsplit it holds one wave.
    const int lanes = split_available(sp.bw, wk.lanes) && (traceback || W > 65) ? wk.lanes : 1;
    const uint64_t row_words = lanes > 1 ? slot_row_words(W, lanes) : (uint64_t)((W + 7) / 8) * 64;
    const uint64_t slot_words = slot_rows(qmax, lanes) * row_words;
    const uint32_t waves = (n + 127) / 128;
    const uint32_t* order = nullptr;
    const uint64_t* woff = nullptr;
    if (qmax > qmin && n > 128) {
      DN_TRY(wk.keys.reserve((uint64_t)n * 4, st));
      DN_TRY(wk.keys2.reserve((uint64_t)n * 4, st));
      DN_TRY(wk.ids.reserve((uint64_t)n * 4, st));
      DN_TRY(wk.ids2.reserve((uint64_t)n * 4, st));
      hipLaunchKernelGGL(k_len_keys, dim3((n + 255) / 256), dim3(256), 0, st, tasks, n, wk.keys.as<uint32_t>(),
                         wk.ids.as<uint32_t>());
      DN_HIP(hipGetLastError());
      const int bits = 32 - __builtin_clz((unsigned)qmax);
      size_t bytes = 0;
      DN_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, bytes, wk.keys.as<uint32_t>(), wk.keys2.as<uint32_t>(),
                                                          wk.ids.as<uint32_t>(), wk.ids2.as<uint32_t>(), n, 0, bits, st));
      DN_TRY(wk.tmp.reserve(bytes, st));
      DN_HIP(hipcub::DeviceRadixSort::SortPairsDescending(wk.tmp.p, bytes, wk.keys.as<uint32_t>(), wk.keys2.as<uint32_t>(),
                                                          wk.ids.as<uint32_t>(), wk.ids2.as<uint32_t>(), n, 0, bits, st));
      order = wk.ids2.as<uint32_t>();
      if (traceback) {
        DN_TRY(wk.woff.reserve((uint64_t)(waves + 1) * 16, st));
        uint64_t* words = wk.woff.as<uint64_t>();
        uint64_t* offs = words + (waves + 1);
        hipLaunchKernelGGL(k_wave_words, dim3((waves + 256) / 256), dim3(256), 0, st, wk.keys2.as<uint32_t>(), waves,
                           row_words, lanes, words);
        DN_HIP(hipGetLastError());
        DN_TRY(scan_u64(wk.tmp, words, offs, waves + 1, st));
        wk.woff_h.resize(waves + 1);
        DN_HIP(hipMemcpyAsync(wk.woff_h.data(), offs, (waves + 1) * 8, hipMemcpyDeviceToHost, st));
        DN_HIP(hipStreamSynchronize(st));
        woff = offs;
      }
    }
    // Launch cuts [w0, w1) in waves.
    const uint64_t budget_words = std::max<uint64_t>(budget / 8, 1);
    std::vector<std::pair<uint32_t, uint32_t>> cuts;
    uint64_t dir_words = 0;
    if (!traceback) {
      cuts.emplace_back(0, waves);
    } else if (woff) {
      for (uint32_t w0 = 0; w0 < waves;) {
        uint32_t w1 = w0 + 1;
        while (w1 < waves && wk.woff_h[w1 + 1] - wk.woff_h[w0] <= budget_words) ++w1;
        cuts.emplace_back(w0, w1);
        dir_words = std::max(dir_words, wk.woff_h[w1] - wk.woff_h[w0]);
        w0 = w1;
      }
    } else {
      const uint32_t per = (uint32_t)std::max<uint64_t>(1, budget_words / slot_words);
      for (uint32_t w0 = 0; w0 < waves; w0 += per) cuts.emplace_back(w0, std::min(waves, w0 + per));
      dir_words = (uint64_t)std::min(waves, per) * slot_words;
    }
    if (traceback) DN_TRY(wk.dir.reserve(dir_words * 8, st));
    for (const auto& c : cuts) {
      const uint32_t t0 = c.first * 128, m = std::min<uint64_t>(n, (uint64_t)c.second * 128) - t0;
      if (order) {
        a.tasks = tasks;
        a.order = order + t0;
        a.res = res;
        a.cigar = cig;
      } else {
        a.tasks = tasks + t0;
        a.order = nullptr;
        a.res = res + t0;
        a.cigar = cig + (uint64_t)t0 * cap;
      }
      a.n = m;
      a.dir = traceback ? wk.dir.as<uint64_t>() : nullptr;
      a.dir_slot_words = slot_words;
      a.woff = woff ? woff + c.first : nullptr;
      a.woff_base = woff ? wk.woff_h[c.first] : 0;
      const unsigned lw = (m + 127) / 128;
      const unsigned grid = (lw + kSwBlock / 64 - 1) / (kSwBlock / 64);
      size_t si = 0;
      if (spans) DN_TRY(spans->begin(fill_family, si));
      bool ok;
      if (lanes > 1)
        ok = traceback ? launch_split_bw<true>(sp.bw, lanes, a, lw, st) : launch_split_bw<false>(sp.bw, lanes, a, lw, st);
      else
        ok = traceback ? launch_packed_bw<true>(sp.bw, a, grid, st) : launch_packed_bw<false>(sp.bw, a, grid, st);
      if (!ok) return fail(DNALN_E_INTERNAL, "no packed kernel for this band width");
      DN_HIP(hipGetLastError());
      if (spans) DN_TRY(spans->end(si));
      if (launches) ++launches[fill_family];
      if (traceback) {
        if (spans) DN_TRY(spans->begin(4, si));
        const bool wok = lanes > 1 ? launch_walk_split(sp.bw, lanes, a, st) : launch_walk_bw(sp.bw, a, wk.walk_kind, st);
        if (!wok) return fail(DNALN_E_INTERNAL, "no walk kernel for this band width");
        DN_HIP(hipGetLastError());
        if (spans) DN_TRY(spans->end(si));
        if (launches) ++launches[4];
      }
    }
  } else {
    SwGenericArgs g{};
    g.sp = sp;
    g.cigar_cap = cap;
    g.traceback = traceback;
    const uint64_t stride = traceback ? (uint64_t)std::max(qmax, 1) * W : 0;
    uint32_t per_launch = n;
    if (traceback) per_launch = (uint32_t)std::min<uint64_t>(n, std::max<uint64_t>(256, budget / std::max<uint64_t>(stride, 1)));
    DN_TRY(wk.rows.reserve((uint64_t)per_launch * 2 * (W + 1) * sizeof(int32_t), st));
    if (traceback) DN_TRY(wk.dir.reserve((uint64_t)per_launch * stride, st));
    for (uint32_t t0 = 0; t0 < n; t0 += per_launch) {
      const uint32_t m = std::min(per_launch, n - t0);
      g.tasks = tasks + t0;
      g.order = nullptr;
      g.n = m;
      g.res = res + t0;
      g.cigar = cig + (uint64_t)t0 * cap;
      g.rows = wk.rows.as<int32_t>();
      g.dir = traceback ? wk.dir.as<uint8_t>() : nullptr;
      g.dir_stride = stride;
      size_t si = 0;
      if (spans) DN_TRY(spans->begin(fill_family, si));
      hipLaunchKernelGGL(k_sw_generic, dim3((m + 127) / 128), dim3(128), 0, st, g);
      DN_HIP(hipGetLastError());
      if (spans) DN_TRY(spans->end(si));
      if (launches) ++launches[fill_family];
    }
  }
  return DNALN_OK;
}

template <class T>
int exclusive_scan(DevBuf& tmp, const T* in, uint64_t* out, uint32_t n, hipStream_t st) {
  size_t bytes = 0;
  DN_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, st));
  DN_TRY(tmp.reserve(bytes, st));
  DN_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, n, st));
  return DNALN_OK;
}

}  // namespace

// ------------------------------------------------------------------------------------ batches
// What a captured graph of the batched form bakes in: the stages it runs, their shapes and
// grids, the traceback launch cuts, and its buffers (`gen` changes whenever one is reallocated).
// A run whose plan differs from the captured one captures a new graph.
struct GraphPlan {
  int32_t sorted, bucket, split0, split1;
  int32_t fused_on[2], fs_cap[2], fs_maxmin[2];
  int32_t retry_on, retry_cap, retry_maxmin;
  uint32_t bins;
  int32_t maxlen;          // CIGAR slot words per task: 2 * maxlen + 4
  int32_t packed, walk_kind, wave_chain, seed_lds;
  uint32_t n_cuts, cut_waves, generic_threads;
  uint64_t slot_words;     // direction words per fill wave when slots are of one size (packed),
                           // direction bytes per thread (int32 fill)
  uint64_t dir_words;      // direction buffer of one cut
  uint64_t anchor_cap;
  uint64_t n_grid;         // reads the grids cover (a power of two >= the batch size)
  uint64_t gen;
};

// A captured graph and the plan it was captured under.
struct GraphEntry {
  GraphPlan plan;
  hipGraphExec_t exec;
  uint64_t launches[6];  // SW launches it holds, per timing family
};

struct dnaln_hip_batch : SpanSink {
  dnaln_hip_ctx* ctx = nullptr;
  dnaln_batch_config cfg{};
  hipStream_t st = nullptr;
  // inputs
  DevBuf codes, offs, lens, fnv;
  uint64_t n = 0, bases = 0, sum_len = 0;
  int minlen = 0, maxlen = 0;
  int split[2] = {0, 0};             // resolved len_split
  uint64_t n_short = 0, n_mid = 0;   // reads with length <= split[0], <= split[1]
  std::vector<int32_t> lens_h;       // host copy of the read lengths (bucket counts)
  bool tuned = false;
  DevBuf fmins;                      // minimizer lists of the fused kernel
  // The host-synchronized batched form (run_chunk): length-sorted read order, second-pass list.
  DevBuf lids, lkeys, lorder, ltmp, ovf2;
  bool device_form = false;          // the last run took the device-count sequence
  // workspace
  DevBuf acnt, aoff, keys, f, prev, order, chains, ring_h, ring_ps;
  DevBuf hits, nhits;                // probe results of the global path's count pass
  DevBuf cands, ncand, rcnt, roff, rtasks, rres, chosen, second, tbflag, tboff, ttasks, tres, tcig;
  SwWork sw;
  DevBuf scan_tmp, ciglen, cigoff, err, scratch8, ovf;
  // outputs
  DevBuf rec, pool;
  uint64_t pool_used = 0;
  dnaln_map_stats stats{};
  // The batched form as a graph (batched_graph.hpp): control block, length-sort histogram and
  // order, global list, scan partials, per-position and per-read task offsets, slot offsets.
  DevBuf gctl, ghist, gorder, glist, gpart, groffp, grank_off, gtoffp, gtb_idx, gwoff;
  std::vector<GraphEntry> graphs;  // captured graphs, most recently used first (a few plans)
  uint64_t enqueued = 0;    // kernels the last enqueue_batched launched
  GraphPlan gplan{};        // plan of the current run
  hipGraphExec_t gexec = nullptr;  // graph of the current run
  uint64_t gen = 0;         // bumped when a buffer of the graph is reallocated
  uint64_t anchor_cap = 0;  // anchors the global path may hold (grown when a run overflows)
  BatchCtl ctl_h{};         // the control block of the last finished run
  std::vector<uint32_t> len_hist;  // host histogram of read lengths (direction-slot bound)
  // Timing: event pairs recorded on the batch stream around each launch group, summed per
  // family (kFamilies) at poll time.
  struct Span {
    hipEvent_t a = nullptr, b = nullptr;
    int family = 0;
  };
  std::vector<Span> spans;
  size_t n_spans = 0;
  double ms[kFamilies] = {};
  uint64_t launches[kFamilies] = {};
  hipEvent_t done = nullptr;
  bool pending = false;
  int last_rc = DNALN_OK;  // status of the last run; poll and download report it

  // Spans may nest (the whole-run span holds the stage spans), so each caller keeps its index.
  int span_begin(int family, size_t& idx) {
    if (n_spans == spans.size()) {
      Span s;
      DN_HIP(hipEventCreate(&s.a));
      DN_HIP(hipEventCreate(&s.b));
      spans.push_back(s);
    }
    idx = n_spans++;
    spans[idx].family = family;
    DN_HIP(hipEventRecord(spans[idx].a, st));
    return DNALN_OK;
  }
  int span_end(size_t idx) {
    if (idx >= n_spans) return fail(DNALN_E_INTERNAL, "timing span out of range");
    DN_HIP(hipEventRecord(spans[idx].b, st));
    return DNALN_OK;
  }
  int begin(int family, size_t& idx) override { return span_begin(family, idx); }
  int end(size_t idx) override { return span_end(idx); }
};

namespace {

// Total of an exclusive scan: off[n-1] + cnt[n-1].
template <class T>
int scan_total(hipStream_t st, const uint64_t* off, const T* cnt, uint32_t n, uint64_t& total) {
  if (n == 0) {
    total = 0;
    return DNALN_OK;
  }
  uint64_t o = 0;
  T c = 0;
  DN_HIP(hipMemcpyAsync(&o, off + n - 1, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
  DN_HIP(hipMemcpyAsync(&c, cnt + n - 1, sizeof(T), hipMemcpyDeviceToHost, st));
  DN_HIP(hipStreamSynchronize(st));
  total = o + (uint64_t)c;
  return DNALN_OK;
}

// Global-memory seed + chain path for reads list[0, nl) (list null: reads 0 .. nl-1): count,
// scan, emit anchors, chain. No limit on anchors per read.
// Seeds and chains the listed reads (all reads when list is null) through global memory: count,
// scan, emit, chain. wave_chain: one wave per read in the chain step (batched form), else one
// thread per read.
int seed_chain_global(dnaln_hip_batch* b, SeedArgs sa, ChainArgs ca, const uint32_t* list, uint32_t nl,
                      size_t& sp_i, bool wave_chain = false) {
  dnaln_hip_ctx* c = b->ctx;
  hipStream_t st = b->st;
  DN_TRY(b->acnt.reserve((uint64_t)nl * 4 + 4, st));
  DN_TRY(b->aoff.reserve((uint64_t)nl * 8 + 8, st));
  // The count pass keeps each read's probe results for the emit pass (12 bytes per base of the
  // batch at most, the region of read r at 3 * offs[r] words).
  DN_TRY(b->hits.reserve(b->bases * 12 + 16, st));
  DN_TRY(b->nhits.reserve((uint64_t)nl * 4 + 4, st));
  sa.n_reads = nl;
  sa.list = list;
  sa.anchor_count = b->acnt.as<uint32_t>();
  sa.anchor_off = b->aoff.as<uint64_t>();
  sa.hits = b->hits.as<uint32_t>();
  sa.n_hits = b->nhits.as<uint32_t>();
  const size_t lds = (size_t)c->w * kSeedBlock * 12;
  const int use_lds = lds <= 48 * 1024;
  if (!use_lds) {
    DN_TRY(b->ring_h.reserve((uint64_t)nl * c->w * 8, st));
    DN_TRY(b->ring_ps.reserve((uint64_t)nl * c->w * 4, st));
  }
  const unsigned sgrid = (nl + kSeedBlock - 1) / kSeedBlock;
  DN_TRY(b->span_begin(0, sp_i));
  if (nl) {
    hipLaunchKernelGGL(k_seed_count, dim3(sgrid), dim3(kSeedBlock), use_lds ? lds : 0, st, sa,
                       b->ring_h.as<uint64_t>(), b->ring_ps.as<uint32_t>(), use_lds);
    DN_HIP(hipGetLastError());
  }
  DN_TRY(exclusive_scan(b->scan_tmp, b->acnt.as<uint32_t>(), b->aoff.as<uint64_t>(), nl, st));
  DN_TRY(b->span_end(sp_i));
  uint64_t na = 0;
  DN_TRY(scan_total(st, b->aoff.as<uint64_t>(), b->acnt.as<uint32_t>(), nl, na));
  DN_TRY(b->keys.reserve(na * 8 + 8, st));
  DN_TRY(b->f.reserve(na * 4 + 4, st));
  DN_TRY(b->prev.reserve(na * 4 + 4, st));
  DN_TRY(b->order.reserve(na * 8 + 8, st));
  DN_TRY(b->chains.reserve(na * sizeof(ChainRec) + sizeof(ChainRec), st));
  sa.keys = b->keys.as<uint64_t>();
  DN_TRY(b->span_begin(0, sp_i));
  if (nl) {
    hipLaunchKernelGGL(k_seed_emit, dim3(sgrid), dim3(kSeedBlock), use_lds ? lds : 0, st, sa,
                       b->ring_h.as<uint64_t>(), b->ring_ps.as<uint32_t>(), use_lds);
    DN_HIP(hipGetLastError());
  }
  DN_TRY(b->span_end(sp_i));
  b->stats.anchors += na;
  ca.n_reads = nl;
  ca.list = list;
  ca.anchor_count = b->acnt.as<uint32_t>();
  ca.anchor_off = b->aoff.as<uint64_t>();
  ca.keys = b->keys.as<uint64_t>();
  ca.f = b->f.as<int32_t>();
  ca.prev = b->prev.as<int32_t>();
  ca.order = b->order.as<uint64_t>();
  ca.chains = b->chains.as<ChainRec>();
  DN_TRY(b->span_begin(1, sp_i));
  if (nl) {
    if (wave_chain && ca.lookback <= 64 && ca.max_cand <= 64) {
      hipLaunchKernelGGL(k_chain_wave, dim3((nl + 3) / 4), dim3(256), wave_chain_lds_bytes(ca.band), st, ca);
    } else {
      hipLaunchKernelGGL(k_chain, dim3(sgrid), dim3(kSeedBlock), 0, st, ca);
    }
    DN_HIP(hipGetLastError());
  }
  DN_TRY(b->span_end(sp_i));
  return DNALN_OK;
}

// Shape of the fused seed + chain kernel: anchor slots (LDS) and minimizer-list entries (global
// scratch) per read.
struct FusedShape {
  int cap = 0, maxmin = 0;
  size_t lds = 0;
};

// Minimizer-list entries for reads of up to maxlen bases: 0.22 per k-mer plus slack (a random
// read has 0.18 on average), in steps of 8.
int fused_list_need(int k, int maxlen) {
  const int nk = std::max(maxlen - k + 1, 1);
  return ((nk * 22 / 100 + 2) + 7) / 8 * 8;
}

// Sized for the longest read of a bucket: 32 anchors up to 200 bp, else 64 (the chain
// extraction's used-mask is 64 bits); the list holds fused_list_need entries, at most 64.
// cap == 0: the fused path does not apply.
FusedShape fused_shape(int k, int w, int maxlen) {
  FusedShape f;
  f.cap = maxlen <= 200 ? 32 : 64;
  f.maxmin = std::min(64, std::max(16, fused_list_need(k, maxlen)));
  f.lds = (size_t)kFusedBlock * 8 * f.cap;
  if (12 * w > 8 * f.cap || f.lds > 64 * 1024 || maxlen >= kFusedMaxLen) f.cap = 0;
  return f;
}

// The shape of the second fused pass over the reads a 32-anchor shape declined: 64 anchors and
// 64 list entries (cap == 0 where it cannot be used).
FusedShape fused_retry_shape(int k, int w, int maxlen) {
  FusedShape f;
  f.cap = 64;
  f.maxmin = 64;
  f.lds = (size_t)kFusedBlock * 8 * f.cap;
  if (12 * w > 8 * f.cap || f.lds > 64 * 1024 || maxlen >= kFusedMaxLen) f.cap = 0;
  return f;
}


// Minimizer scratch of a fused launch over n reads (u64 entries).
uint64_t fused_mins_words(const FusedShape& fs, uint64_t n) {
  return (n + kFusedBlock - 1) / kFusedBlock * kFusedBlock * (uint64_t)fs.maxmin;
}

// The fused kernel over sa.list[0, sa.n_reads) (or reads 0 .. n_reads-1); reads it declines go
// to ovf, counted in *novf.
int launch_fused(const FusedShape& fs, const SeedArgs& sa, const ChainArgs& ca, uint64_t* mins, uint32_t* ovf,
                 uint32_t* novf, hipStream_t st) {
  const unsigned grid = (sa.n_reads + kFusedBlock - 1) / kFusedBlock;
  if (fs.cap != 32 && fs.cap != 64) return fail(DNALN_E_INTERNAL, "fused seed + chain shape without a kernel");
#define DN_FUSED(P8, CAP)                                                                                            \
  hipLaunchKernelGGL((k_seed_chain_fused<10, P8, CAP>), dim3(grid), dim3(kFusedBlock), fs.lds, st, sa, ca, fs.maxmin, \
                     mins, ovf, novf)
  if (sa.ctab) {
    if (fs.cap == 32) DN_FUSED(true, 32);
    else DN_FUSED(true, 64);
  } else {
    if (fs.cap == 32) DN_FUSED(false, 32);
    else DN_FUSED(false, 64);
  }
#undef DN_FUSED
  DN_HIP(hipGetLastError());
  return DNALN_OK;
}

// Default length buckets: the longest reads whose expected minimizer list fits the 32-entry
// shape (200 bp at most) and the 64-entry shape; 0 where the shape cannot be used.
void default_splits(int k, int w, int out[2]) {
  out[0] = out[1] = 0;
  for (int L = 200; L >= k; --L)
    if (fused_list_need(k, L) <= 32) {
      out[0] = fused_shape(k, w, L).cap ? L : 0;
      break;
    }
  for (int L = 4000; L >= k; --L)
    if (fused_list_need(k, L) <= 64) {
      out[1] = fused_shape(k, w, L).cap ? L : 0;
      break;
    }
}


// The context's chain and candidate parameters (the chain band already clamped to bw).
ChainSetup chain_setup(const dnaln_hip_ctx& c) {
  ChainSetup s;
  s.k = c.k;
  s.max_gap = c.p.ch.max_gap;
  s.band = c.p.ch.band;
  s.lookback = c.p.ch.lookback;
  s.min_score = c.p.ch.min_score;
  s.min_anchors = c.p.ch.min_anchors;
  s.max_cand = std::max(c.p.max_cand, 0);
  s.bw = c.p.sc.bw;
  s.pri_ratio = c.p.pri_ratio;
  return s;
}

// ------------------------------------------------------------------ batched form as a graph
// Compute units of a device (persistent grids are sized from it).
int device_cus(int device) {
  static std::mutex mu;
  static int cache[256] = {};
  std::lock_guard<std::mutex> lk(mu);
  if (device < 0 || device >= 256) return 256;
  if (!cache[device]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || v <= 0) {
      (void)hipGetLastError();
      v = 256;
    }
    cache[device] = v;
  }
  return cache[device];
}

// Threads of the per-read kernels of the graph (grid-stride loops), from the CU count.
unsigned read_blocks(int device) { return (unsigned)device_cus(device) * 8; }

// Grows buffer d of batch b for the graph path; a reallocation changes the buffer generation,
// so the next run captures a new graph. Only the batch's own stream can be using d.
int greserve(dnaln_hip_batch* b, DevBuf& d, uint64_t bytes) {
  if (bytes <= d.cap) return DNALN_OK;
  DN_TRY(d.reserve(bytes, b->st));
  ++b->gen;
  return DNALN_OK;
}

// Smallest planned batch size: grids of batches up to this many reads are the same, so their
// plans (and graphs) are shared. The CPU emulation, where every block of a grid costs fiber
// set-up, plans smaller grids.
#if defined(DN_EMU) && !defined(DN_EMU_GPU_SHAPES)
constexpr uint64_t kGridFloor = 128;
#else
constexpr uint64_t kGridFloor = 4096;
#endif

uint64_t next_pow2(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

// Bound on the traceback direction words of a run, from the batch's read lengths. The device
// sorts reads by length bin, bin = min(len, bins - 1), longest first; a fill wave's slot holds
// the rows of its longest task. Count a read of the top bin as maxlen (its bin is not sorted
// inside): then the traceback tasks, a subset of the reads in the same order, put no task longer
// than the (128 w)-th read of this order into wave w. Walks the histogram from the top.
uint64_t dir_bound_words(const std::vector<uint32_t>& hist, uint64_t n, uint64_t row_words, uint32_t bins) {
  uint64_t total = 0, seen = 0, next = 0;  // next: position of the next wave's first task
  const int64_t maxlen = (int64_t)hist.size() - 1;
  for (int64_t L = maxlen; L >= 0 && next < n; --L) {
    const uint64_t c = hist[(size_t)L];
    const int eff = bins && L >= (int64_t)bins - 1 ? (int)maxlen : (int)L;
    while (next < seen + c && next < n) {
      total += fill_rows(eff) * row_words;
      next += 128;
    }
    seen += c;
  }
  return total;
}

// Sizes and reserves every buffer of the batched graph for the resident batch, and fills the
// plan. Reallocations bump b->gen before pl.gen is taken.
int plan_batched(dnaln_hip_batch* b, GraphPlan& pl) {
  dnaln_hip_ctx* c = b->ctx;
  const dnaln_map_params& p = c->p;
  const ScoreP sp{p.sc.match, p.sc.mismatch, p.sc.gap_open, p.sc.gap_ext, p.sc.ambig, p.sc.bw};
  std::memset(&pl, 0, sizeof pl);
  const uint64_t n = b->n, mc = (uint64_t)std::max(p.max_cand, 1);
  const int W = 2 * p.sc.bw + 1;
  const uint64_t row_words = (uint64_t)((W + 7) / 8) * 64;
  pl.n_grid = std::max<uint64_t>(kGridFloor, next_pow2(n));
  pl.sorted = b->minlen < b->maxlen;
  // Buckets: reads up to split[0] / split[1] bases take the fused kernel in the shape for that
  // length; a bucket whose shape cannot hold the sketch ring (cap 0) joins the next one, so its
  // reads take the global path.
  int cut[2] = {b->split[0], b->split[1]};
  for (int x = 0; x < 2; ++x) {
    const FusedShape fs = fused_shape(c->k, c->w, std::min(b->split[x], b->maxlen));
    pl.fs_cap[x] = fs.cap;
    pl.fs_maxmin[x] = fs.maxmin;
    if (!fs.cap) cut[x] = x == 0 ? -1 : cut[0];
  }
  pl.split0 = cut[0];
  pl.split1 = std::max(cut[1], cut[0]);
  pl.bucket = pl.sorted ? -1 : (b->maxlen <= pl.split0 ? 0 : b->maxlen <= pl.split1 ? 1 : 2);
  bool small = false;
  for (int x = 0; x < 2; ++x) {
    pl.fused_on[x] = pl.fs_cap[x] && (pl.sorted ? (x == 0 ? pl.split0 >= 0 : pl.split1 > pl.split0) : pl.bucket == x);
    if (pl.fused_on[x] && pl.fs_cap[x] < 64) small = true;
  }
  const FusedShape retry = fused_retry_shape(c->k, c->w, b->maxlen);
  pl.retry_on = small && retry.cap;
  pl.retry_cap = retry.cap;
  pl.retry_maxmin = retry.maxmin;
  pl.bins = pl.sorted ? (uint32_t)std::min<uint64_t>(kMaxLenBins, std::max<uint64_t>(256, next_pow2(b->maxlen + 1))) : 0;
  pl.maxlen = (int)std::max<uint64_t>(64, next_pow2((uint64_t)std::max(b->maxlen, 1)));
  pl.packed = packed_bw_available(p.sc.bw) && packed_ok(sp, b->maxlen);
  pl.walk_kind = b->sw.walk_kind;
  pl.wave_chain = p.ch.lookback <= 64 && p.max_cand <= 64;
  pl.seed_lds = (size_t)c->w * kSeedBlock * 12 <= 48 * 1024;
  // Traceback directions: one cut when the bound fits the budget, else cuts of cut_waves waves,
  // each no larger than the budget at the longest read (at most kMaxCuts cuts).
  const uint64_t waves = (n + 127) / 128, budget_words = std::max<uint64_t>(b->sw.dir_budget / 8, 1);
  uint64_t dir_bytes = 0, rows_bytes = 0;
  if (pl.packed) {
    const uint64_t at_max = fill_rows(b->maxlen) * row_words;
    const uint64_t bound = pl.sorted ? dir_bound_words(b->len_hist, n, row_words, pl.bins) : waves * at_max;
    pl.slot_words = pl.sorted ? 0 : at_max;
    if (bound <= budget_words) {
      pl.n_cuts = 1;
      pl.cut_waves = 0;  // all
      dir_bytes = bound * 8;
    } else {
      uint64_t per = std::max<uint64_t>(1, budget_words / std::max<uint64_t>(at_max, 1));
      if ((waves + per - 1) / per > (uint64_t)kMaxCuts) per = (waves + kMaxCuts - 1) / kMaxCuts;
      pl.cut_waves = (uint32_t)per;
      pl.n_cuts = (uint32_t)((waves + per - 1) / per);
      dir_bytes = per * at_max * 8;
    }
  } else {
    // Per thread of a fixed grid: the direction bytes of the longest read the plan covers.
    const uint64_t per_task = (uint64_t)pl.maxlen * W;
    uint64_t t = std::min<uint64_t>(std::max<uint64_t>(n, 1), std::max<uint64_t>(256, b->sw.dir_budget / per_task));
    t = (t + 127) / 128 * 128;
    pl.generic_threads = (uint32_t)t;
    pl.slot_words = per_task;  // bytes of directions per thread
    dir_bytes = t * per_task;
    rows_bytes = t * 2 * (uint64_t)(W + 1) * 4;
    pl.n_cuts = 1;
  }
  // Anchors of the global path: anchors_per_read per read when configured, else an estimate
  // (8 per read for declined reads, one per two bases of the long reads); a run that needs more
  // grows the buffer and runs again.
  {
    uint64_t want = 0;
    if (b->cfg.anchors_per_read) {
      want = n * b->cfg.anchors_per_read + 1;
    } else {
      uint64_t long_bases = 0;
      for (size_t L = (size_t)std::max(b->split[1], 0) + 1; L < b->len_hist.size(); ++L)
        long_bases += L * b->len_hist[L];
      want = n * 8 + long_bases / 2 + 65536;
    }
    b->anchor_cap = std::max(b->anchor_cap, want);
  }
  const uint64_t na = b->anchor_cap, cap = 2 * (uint64_t)pl.maxlen + 4;
  // Reserve (grow-only) every buffer the graph touches.
  DN_TRY(greserve(b, b->gctl, sizeof(BatchCtl)));
  // The control block holds no stale sizes between runs: zeroed here, outside the graph, so no
  // kernel of a run can see another run's counts before k_begin and k_reset set them.
  DN_HIP(hipMemsetAsync(b->gctl.p, 0, sizeof(BatchCtl), b->st));
  DN_TRY(greserve(b, b->ghist, std::max<uint64_t>(pl.bins, 1) * 4));
  DN_TRY(greserve(b, b->gorder, n * 4 + 4));
  DN_TRY(greserve(b, b->cands, n * mc * sizeof(Cand) + sizeof(Cand)));
  DN_TRY(greserve(b, b->ncand, n * 4 + 4));
  DN_TRY(greserve(b, b->ovf, n * 4 + 4));
  DN_TRY(greserve(b, b->glist, n * 4 + 4));
  DN_TRY(greserve(b, b->fmins, (n + kFusedBlock) / kFusedBlock * kFusedBlock * 64 * 8));
  DN_TRY(greserve(b, b->acnt, n * 4 + 4));
  DN_TRY(greserve(b, b->aoff, n * 8 + 8));
  DN_TRY(greserve(b, b->keys, na * 8 + 8));
  DN_TRY(greserve(b, b->f, na * 4 + 4));
  DN_TRY(greserve(b, b->prev, na * 4 + 4));
  DN_TRY(greserve(b, b->order, na * 8 + 8));
  DN_TRY(greserve(b, b->chains, na * sizeof(ChainRec) + sizeof(ChainRec)));
  if (!pl.seed_lds) {
    const uint64_t threads = (pl.n_grid + kSeedBlock - 1) / kSeedBlock * kSeedBlock;
    DN_TRY(greserve(b, b->ring_h, threads * c->w * 8));
    DN_TRY(greserve(b, b->ring_ps, threads * c->w * 4));
  }
  DN_TRY(greserve(b, b->gpart, kScanBlocks * 8));
  DN_TRY(greserve(b, b->groffp, n * 8 + 8));
  DN_TRY(greserve(b, b->grank_off, n * 8 + 8));
  DN_TRY(greserve(b, b->rtasks, n * mc * sizeof(SwTask) + sizeof(SwTask)));
  DN_TRY(greserve(b, b->rres, n * mc * sizeof(SwResult) + sizeof(SwResult)));
  DN_TRY(greserve(b, b->chosen, n * 4 + 4));
  DN_TRY(greserve(b, b->second, n * 4 + 4));
  DN_TRY(greserve(b, b->gtoffp, n * 8 + 8));
  DN_TRY(greserve(b, b->gtb_idx, n * 8 + 8));
  DN_TRY(greserve(b, b->ttasks, n * sizeof(SwTask) + sizeof(SwTask)));
  DN_TRY(greserve(b, b->tres, n * sizeof(SwResult) + sizeof(SwResult)));
  DN_TRY(greserve(b, b->tcig, n * cap * 4 + 4));
  DN_TRY(greserve(b, b->gwoff, (waves + 2) * 8));
  DN_TRY(greserve(b, b->sw.dir, std::max<uint64_t>(dir_bytes, 8)));
  if (rows_bytes) DN_TRY(greserve(b, b->sw.rows, rows_bytes));
  DN_TRY(greserve(b, b->ciglen, n * 4 + 4));
  DN_TRY(greserve(b, b->cigoff, n * 8 + 8));
  DN_TRY(greserve(b, b->pool, (2 * b->sum_len + 4 * n + 4) * 4));
  DN_TRY(greserve(b, b->rec, n * sizeof(dnaln_record) + sizeof(dnaln_record)));
  pl.anchor_cap = na;
  pl.dir_words = b->sw.dir.cap / 8;
  pl.gen = b->gen;
  return DNALN_OK;
}

// Every stage of the batched form on the batch stream (captured into a graph, or launched
// directly). Nothing here reads a count on the host or touches memory outside a kernel.
int enqueue_batched(dnaln_hip_batch* b, const GraphPlan& pl, bool spans) {
  dnaln_hip_ctx* c = b->ctx;
  const dnaln_map_params& p = c->p;
  hipStream_t st = b->st;
  const int max_cand = std::max(p.max_cand, 0);
  const ScoreP sp{p.sc.match, p.sc.mismatch, p.sc.gap_open, p.sc.gap_ext, p.sc.ambig, p.sc.bw};
  const int W = 2 * sp.bw + 1, cap = 2 * pl.maxlen + 4;
  const uint64_t row_words = (uint64_t)((W + 7) / 8) * 64;
  const uint64_t ng = pl.n_grid, mc = (uint64_t)std::max(max_cand, 1);
  const unsigned rb = read_blocks(c->device);
  BatchCtl* ctl = b->gctl.as<BatchCtl>();
  uint32_t* order = pl.sorted ? b->gorder.as<uint32_t>() : nullptr;
  size_t si = 0;
  auto open = [&](int family) -> int { return spans ? b->span_begin(family, si) : DNALN_OK; };
  auto close = [&]() -> int { return spans ? b->span_end(si) : DNALN_OK; };

  ReadView rv;
  rv.codes = b->codes.as<uint8_t>() + kReadGuard;
  rv.offs = b->offs.as<uint64_t>();
  rv.lens = b->lens.as<int32_t>();
  rv.fnv = b->fnv.as<uint32_t>();
  MapGeom g;
  g.ref = (uint64_t)(uintptr_t)c->ref_base;
  g.coff = c->coff.as<uint64_t>();
  g.clen = c->clen.as<uint32_t>();
  g.bw = p.sc.bw;
  SeedArgs sa{};
  sa.codes = rv.codes;
  sa.offs = rv.offs;
  sa.lens = rv.lens;
  sa.table = c->table.as<Bucket>();
  sa.mask = c->mask;
  sa.pos = c->probe8 ? c->pos2.as<uint32_t>() : c->pos.as<uint32_t>();
  sa.ctab = c->probe8 ? c->ctab.as<uint64_t>() : nullptr;
  sa.pos_inline = 1;
  sa.k = c->k;
  sa.w = c->w;
  sa.max_occ = p.max_occ;
  ChainArgs ca{};
  init_chain_args(ca, chain_setup(*c), c->cost.as<int32_t>(), c->h_cost.data(), c->coff.as<uint64_t>(), c->nc);
  ca.cands = b->cands.as<Cand>();
  ca.n_cand = b->ncand.as<int32_t>();

  // 1. reset, length order, fused seed + chain
  DN_TRY(open(0));
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(256), 0, st, ctl, pl.sorted ? -1 : pl.bucket, b->ghist.as<uint32_t>(),
                     pl.bins);
  b->enqueued += 1;
  DN_HIP(hipGetLastError());
  if (pl.sorted) {
    hipLaunchKernelGGL(k_len_hist, dim3(rb / 4), dim3(kHistThreads), (size_t)pl.bins * 4, st, rv.lens, ctl,
                       pl.bins, b->ghist.as<uint32_t>(), pl.split0, pl.split1);
    b->enqueued += 1;
    hipLaunchKernelGGL(k_len_starts, dim3(1), dim3(kHistThreads), 0, st, ctl, pl.bins, b->ghist.as<uint32_t>());
    b->enqueued += 1;
    hipLaunchKernelGGL(k_len_scatter, dim3(rb), dim3(256), 0, st, rv.lens, ctl, pl.bins,
                       b->ghist.as<uint32_t>(), order);
    b->enqueued += 1;
    DN_HIP(hipGetLastError());
  }
  uint32_t* ovf = b->ovf.as<uint32_t>();
  uint32_t* glist = b->glist.as<uint32_t>();
  uint64_t* fmins = b->fmins.as<uint64_t>();
  for (int x = 0; x < 2; ++x) {
    if (!pl.fused_on[x]) continue;
    hipLaunchKernelGGL((k_fused_g<10>), dim3((unsigned)(ng / kFusedBlock)), dim3(kFusedBlock),
                       (size_t)kFusedBlock * 8 * pl.fs_cap[x], st,
                       sa, ca, pl.fs_cap[x], pl.fs_maxmin[x], fmins, ctl, x, order, (const uint32_t*)nullptr,
                       pl.retry_on ? ovf : glist, !pl.retry_on);
    b->enqueued += 1;
    DN_HIP(hipGetLastError());
  }
  if (pl.retry_on) {
    hipLaunchKernelGGL((k_fused_g<10>), dim3((unsigned)(ng / kFusedBlock)), dim3(kFusedBlock),
                       (size_t)kFusedBlock * 8 * pl.retry_cap, st,
                       sa, ca, pl.retry_cap, pl.retry_maxmin, fmins, ctl, 2, order, (const uint32_t*)ovf, glist, true);
    b->enqueued += 1;
    DN_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_long_to_global, dim3(1), dim3(1024), 0, st, ctl, order, glist, &ctl->nglobal);
  b->enqueued += 1;
  DN_HIP(hipGetLastError());
  DN_TRY(close());

  // 2. global path: long reads and declined reads
  DN_TRY(open(0));
  sa.list = glist;
  sa.n_reads = 0;
  sa.anchor_count = b->acnt.as<uint32_t>();
  sa.anchor_off = b->aoff.as<uint64_t>();
  sa.keys = b->keys.as<uint64_t>();
  const size_t seed_lds = pl.seed_lds ? (size_t)c->w * kSeedBlock * 12 : 0;
  uint64_t* ring_h = pl.seed_lds ? nullptr : b->ring_h.as<uint64_t>();
  uint32_t* ring_ps = pl.seed_lds ? nullptr : b->ring_ps.as<uint32_t>();
  const unsigned sgrid = (unsigned)((ng + kSeedBlock - 1) / kSeedBlock);
  hipLaunchKernelGGL(k_gseed_count, dim3(sgrid), dim3(kSeedBlock), seed_lds, st, sa, ctl, ring_h, ring_ps);
  b->enqueued += 1;
  DN_HIP(hipGetLastError());
  enqueue_scan(ScanAnchors{b->acnt.as<uint32_t>()}, &ctl->nglobal, b->gpart.as<uint64_t>(), b->aoff.as<uint64_t>(),
               (uint32_t*)nullptr, &ctl->anchors, st);
  b->enqueued += 2;
  hipLaunchKernelGGL(k_gseed_emit, dim3(sgrid), dim3(kSeedBlock), seed_lds, st, sa, ctl, pl.anchor_cap, ring_h,
                     ring_ps);
  b->enqueued += 1;
  DN_HIP(hipGetLastError());
  DN_TRY(close());
  DN_TRY(open(1));
  ChainArgs cg = ca;
  cg.list = glist;
  cg.n_reads = 0;
  cg.anchor_count = b->acnt.as<uint32_t>();
  cg.anchor_off = b->aoff.as<uint64_t>();
  cg.keys = b->keys.as<uint64_t>();
  cg.f = b->f.as<int32_t>();
  cg.prev = b->prev.as<int32_t>();
  cg.order = b->order.as<uint64_t>();
  cg.chains = b->chains.as<ChainRec>();
  if (pl.wave_chain)
    hipLaunchKernelGGL(k_gchain, dim3((unsigned)(ng / 4)), dim3(256), wave_chain_lds_bytes(cg.band), st, cg, ctl,
                       pl.anchor_cap);
  else
    hipLaunchKernelGGL(k_gchain_thread, dim3(sgrid), dim3(kSeedBlock), 0, st, cg, ctl, pl.anchor_cap);
  b->enqueued += 1;
  DN_HIP(hipGetLastError());
  DN_TRY(close());

  // 3. rank pass
  enqueue_scan(ScanRank{order, b->ncand.as<int32_t>()}, &ctl->n, b->gpart.as<uint64_t>(), b->groffp.as<uint64_t>(),
               &ctl->nr, (uint64_t*)nullptr, st);
  b->enqueued += 2;
  hipLaunchKernelGGL(k_grank_tasks, dim3(rb), dim3(256), 0, st, rv, g, b->cands.as<Cand>(),
                     b->ncand.as<int32_t>(), max_cand, order, b->groffp.as<uint64_t>(), b->grank_off.as<uint64_t>(),
                     b->rtasks.as<SwTask>(), ctl);
  b->enqueued += 1;
  DN_HIP(hipGetLastError());
  DN_TRY(open(2));
  SwArgs a{};
  a.sp = sp;
  score_table(sp, a.tab0, a.tab1);
  SwGenericArgs ga{};
  ga.sp = sp;
  ga.cigar_cap = cap;
  ga.order = nullptr;
  ga.rows = b->sw.rows.as<int32_t>();
  ga.dir = b->sw.dir.as<uint8_t>();
  ga.dir_stride = pl.packed ? 0 : pl.slot_words;
  const unsigned gen_grid = (pl.generic_threads + 127) / 128;
  if (pl.packed) {
    a.tasks = b->rtasks.as<SwTask>();
    a.order = nullptr;
    a.res = b->rres.as<SwResult>();
    a.n = 0xffffffffu;
    if (!launch_packed_g<false>(sp.bw, a, (ng * mc + 127) / 128, &ctl->nr, 0, st))
      return fail(DNALN_E_INTERNAL, "no packed kernel for this band width");
    b->enqueued += 1;
  } else {
    ga.tasks = b->rtasks.as<SwTask>();
    ga.res = b->rres.as<SwResult>();
    ga.cigar = nullptr;
    ga.traceback = 0;
    hipLaunchKernelGGL(k_sw_generic_g, dim3(gen_grid), dim3(128), 0, st, ga, &ctl->nr);
    b->enqueued += 1;
  }
  DN_HIP(hipGetLastError());
  ++b->launches[2];
  DN_TRY(close());
  hipLaunchKernelGGL(k_gselect, dim3(rb), dim3(256), 0, st, rv, b->ncand.as<int32_t>(),
                     b->grank_off.as<uint64_t>(), b->rres.as<SwResult>(), b->chosen.as<int32_t>(),
                     b->second.as<int32_t>(), ctl);
  b->enqueued += 1;
  DN_HIP(hipGetLastError());

  // 4. traceback pass
  enqueue_scan(ScanChosen{order, b->chosen.as<int32_t>()}, &ctl->n, b->gpart.as<uint64_t>(), b->gtoffp.as<uint64_t>(),
               &ctl->nt, (uint64_t*)nullptr, st);
  b->enqueued += 2;
  hipLaunchKernelGGL(k_gtb_tasks, dim3(rb), dim3(256), 0, st, rv, g, b->cands.as<Cand>(), max_cand,
                     b->chosen.as<int32_t>(), order, b->gtoffp.as<uint64_t>(), b->gtb_idx.as<int64_t>(),
                     b->ttasks.as<SwTask>(), ctl);
  b->enqueued += 1;
  DN_HIP(hipGetLastError());
  SwTask* tt = b->ttasks.as<SwTask>();
  SwResult* tr = b->tres.as<SwResult>();
  uint32_t* tc = b->tcig.as<uint32_t>();
  if (pl.packed) {
    if (pl.sorted) {
      enqueue_scan(ScanWaveWords{tt, &ctl->nt, row_words}, &ctl->tb_waves, b->gpart.as<uint64_t>(), b->gwoff.as<uint64_t>(),
                   (uint32_t*)nullptr, (uint64_t*)nullptr, st);
      b->enqueued += 2;
    }
    a.cigar_cap = cap;
    a.dir = b->sw.dir.as<uint64_t>();
    a.dir_slot_words = pl.slot_words;
    for (uint32_t j = 0; j < pl.n_cuts; ++j) {
      const uint64_t w0 = (uint64_t)j * pl.cut_waves, first = w0 * 128;
      a.tasks = tt + first;
      a.res = tr + first;
      a.cigar = tc + first * cap;
      a.n = pl.cut_waves ? pl.cut_waves * 128u : 0xffffffffu;
      a.woff = pl.sorted ? b->gwoff.as<uint64_t>() + w0 : nullptr;
      a.woff_base = 0;
      const uint64_t cut_waves = pl.cut_waves ? pl.cut_waves : (ng + 127) / 128;
      DN_TRY(open(3));
      if (!launch_packed_g<true>(sp.bw, a, cut_waves, &ctl->nt, (uint32_t)first, st))
        return fail(DNALN_E_INTERNAL, "no packed kernel for this band width");
      b->enqueued += 1;
      DN_HIP(hipGetLastError());
      DN_TRY(close());
      ++b->launches[3];
      DN_TRY(open(4));
      if (!launch_walk_g(sp.bw, a, pl.walk_kind, cut_waves * 128, &ctl->nt, (uint32_t)first, st))
        return fail(DNALN_E_INTERNAL, "no walk kernel for this band width");
      b->enqueued += 1;
      DN_HIP(hipGetLastError());
      DN_TRY(close());
      ++b->launches[4];
    }
  } else {
    ga.tasks = tt;
    ga.res = tr;
    ga.cigar = tc;
    ga.traceback = 1;
    DN_TRY(open(3));
    hipLaunchKernelGGL(k_sw_generic_g, dim3(gen_grid), dim3(128), 0, st, ga, &ctl->nt);
    b->enqueued += 1;
    DN_HIP(hipGetLastError());
    DN_TRY(close());
    ++b->launches[3];
  }

  // 5. records and CIGARs
  hipLaunchKernelGGL(k_grecords, dim3(rb), dim3(256), 0, st, rv, g, b->cands.as<Cand>(), b->ncand.as<int32_t>(),
                     max_cand, b->chosen.as<int32_t>(), b->second.as<int32_t>(), b->gtb_idx.as<int64_t>(), tr,
                     p.min_aln_score, c->lut.as<uint8_t>(), c->lut_len, b->rec.as<dnaln_record>(),
                     b->ciglen.as<uint32_t>(), ctl);
  b->enqueued += 1;
  DN_HIP(hipGetLastError());
  enqueue_scan(ScanCigar{b->ciglen.as<uint32_t>()}, &ctl->n, b->gpart.as<uint64_t>(), b->cigoff.as<uint64_t>(),
               (uint32_t*)nullptr, &ctl->ncig, st);
  b->enqueued += 2;
  hipLaunchKernelGGL(k_gcigar, dim3(rb), dim3(256), 0, st, rv, b->chosen.as<int32_t>(), b->gtb_idx.as<int64_t>(),
                     tr, tc, cap, b->cigoff.as<uint64_t>(), b->rec.as<dnaln_record>(), b->pool.as<uint32_t>(), ctl);
  b->enqueued += 1;
  DN_HIP(hipGetLastError());
  return DNALN_OK;
}

// Captures the batched form under plan pl into a new graph at the front of b->graphs. The
// capture holds kernel launches only: no event is recorded while the stream captures (an event
// recorded in a capturing stream is only a capture-time marker, not a graph node, and may not be
// queried afterwards), and the graph is checked to be exactly the chain of the launches made,
// each depending on the one before, before it is instantiated.
constexpr size_t kGraphCache = 6;  // graphs kept per batch (plans differing in grid size or shape)

int check_chain(hipGraph_t graph, uint64_t launches) {
  size_t n = 0;
  DN_HIP(hipGraphGetNodes(graph, nullptr, &n));
  if (n != launches)
    return fail(DNALN_E_INTERNAL, "captured graph holds " + std::to_string(n) + " nodes for " +
                                      std::to_string(launches) + " launches");
  std::vector<hipGraphNode_t> nodes(n);
  DN_HIP(hipGraphGetNodes(graph, nodes.data(), &n));
  size_t roots = 0;
  for (hipGraphNode_t nd : nodes) {
    hipGraphNodeType type;
    DN_HIP(hipGraphNodeGetType(nd, &type));
    if (type != hipGraphNodeTypeKernel) return fail(DNALN_E_INTERNAL, "captured graph holds a node that is no kernel");
    size_t deps = 0, users = 0;
    DN_HIP(hipGraphNodeGetDependencies(nd, nullptr, &deps));
    DN_HIP(hipGraphNodeGetDependentNodes(nd, nullptr, &users));
    if (deps > 1 || users > 1) return fail(DNALN_E_INTERNAL, "captured graph is not a chain of launches");
    roots += deps == 0;
  }
  if (n && roots != 1) return fail(DNALN_E_INTERNAL, "captured graph is not one chain of launches");
  return DNALN_OK;
}

int capture_batched(dnaln_hip_batch* b, const GraphPlan& pl) {
  for (auto& l : b->launches) l = 0;
  b->enqueued = 0;
  DN_HIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeRelaxed));
  const int rc = enqueue_batched(b, pl, false);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(b->st, &graph);
  if (rc != DNALN_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    return rc;
  }
  if (e != hipSuccess) return fail(DNALN_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  const int ok = check_chain(graph, b->enqueued);
  if (ok != DNALN_OK) {
    (void)hipGraphDestroy(graph);
    return ok;
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) return fail(DNALN_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  if (b->graphs.size() >= kGraphCache) {
    (void)hipGraphExecDestroy(b->graphs.back().exec);
    b->graphs.pop_back();
  }
  GraphEntry ge{pl, exec, {}};
  std::memcpy(ge.launches, b->launches, sizeof ge.launches);
  b->graphs.insert(b->graphs.begin(), ge);
  return DNALN_OK;
}

// Plans the next run of the batched form and reserves its buffers; in graph mode finds the
// graph captured under the same plan, or captures one.
int prepare_batched(dnaln_hip_batch* b) {
  GraphPlan pl;
  DN_TRY(plan_batched(b, pl));
  b->gplan = pl;
  if (b->cfg.graph != kFormGraph) return DNALN_OK;
  // Graphs from before a buffer was reallocated point at freed memory.
  for (size_t i = 0; i < b->graphs.size();) {
    if (b->graphs[i].plan.gen != pl.gen) {
      DN_HIP(hipGraphExecDestroy(b->graphs[i].exec));
      b->graphs.erase(b->graphs.begin() + (long)i);
    } else {
      ++i;
    }
  }
  size_t hit = b->graphs.size();
  for (size_t i = 0; i < b->graphs.size(); ++i)
    if (std::memcmp(&b->graphs[i].plan, &pl, sizeof pl) == 0) hit = i;
  if (hit < b->graphs.size()) {
    std::rotate(b->graphs.begin(), b->graphs.begin() + (long)hit, b->graphs.begin() + (long)hit + 1);
  } else {
    DN_TRY(capture_batched(b, pl));
  }
  b->gexec = b->graphs[0].exec;
  std::memcpy(b->launches, b->graphs[0].launches, sizeof b->launches);
  return DNALN_OK;
}

// Enqueues the prepared run: the batch size, then the stages launched one by one (or, with
// cfg.graph == 1, the captured graph). Returns without waiting for anything.
int launch_batched(dnaln_hip_batch* b) {
  hipLaunchKernelGGL(k_begin, dim3(1), dim3(1), 0, b->st, b->gctl.as<BatchCtl>(), (uint32_t)b->n);
  DN_HIP(hipGetLastError());
  if (b->cfg.graph != kFormGraph) return enqueue_batched(b, b->gplan, true);
  DN_HIP(hipGraphLaunch(b->gexec, b->st));
  return DNALN_OK;
}

// One chunk of reads [r0, r0 + n) through the whole pipeline, launch by launch, with host reads
// of the task counts: the plain form (chunks of chunk_reads), and the host-synchronized batched
// form (the whole batch as one chunk, length buckets, length-sorted SW tasks).
int run_chunk(dnaln_hip_batch* b, uint64_t r0, uint32_t n, bool first_chunk) {
  dnaln_hip_ctx* c = b->ctx;
  const dnaln_map_params& p = c->p;
  hipStream_t st = b->st;
  const int max_cand = std::max(p.max_cand, 0);
  const ScoreP sp{p.sc.match, p.sc.mismatch, p.sc.gap_open, p.sc.gap_ext, p.sc.ambig, p.sc.bw};
  const int cap = 2 * b->maxlen + 4;
  // The batched form sorts SW tasks by length; the plain form runs them in read order.
  const int qmin = b->cfg.mode == DNALN_MODE_BATCHED ? b->minlen : b->maxlen;
  size_t sp_i = 0;

  ReadView rv;
  rv.codes = b->codes.as<uint8_t>() + kReadGuard;
  rv.offs = b->offs.as<uint64_t>() + r0;
  rv.lens = b->lens.as<int32_t>() + r0;
  rv.fnv = b->fnv.as<uint32_t>() + r0;
  MapGeom g;
  g.ref = (uint64_t)(uintptr_t)c->ref_base;
  g.coff = c->coff.as<uint64_t>();
  g.clen = c->clen.as<uint32_t>();
  g.bw = p.sc.bw;

  // 1-2. seeds, chains and candidates
  DN_TRY(b->cands.reserve((uint64_t)n * std::max(max_cand, 1) * sizeof(Cand), st));
  DN_TRY(b->ncand.reserve((uint64_t)n * 4 + 4, st));
  SeedArgs sa{};
  sa.codes = rv.codes;  // kReadGuard bytes into the buffer
  sa.offs = rv.offs;
  sa.lens = rv.lens;
  sa.table = c->table.as<Bucket>();
  sa.mask = c->mask;
  sa.pos = c->probe8 ? c->pos2.as<uint32_t>() : c->pos.as<uint32_t>();
  sa.ctab = c->probe8 ? c->ctab.as<uint64_t>() : nullptr;
  sa.pos_inline = 1;
  sa.k = c->k;
  sa.w = c->w;
  sa.max_occ = p.max_occ;
  ChainArgs ca{};
  init_chain_args(ca, chain_setup(*c), c->cost.as<int32_t>(), c->h_cost.data(), c->coff.as<uint64_t>(), c->nc);
  ca.cands = b->cands.as<Cand>();
  ca.n_cand = b->ncand.as<int32_t>();
  const unsigned sgrid = (n + kSeedBlock - 1) / kSeedBlock;
  if (b->cfg.mode == DNALN_MODE_BATCHED && n) {
    // Batched form (one chunk: the whole batch). Length buckets: reads of up to split[0] and
    // split[1] bases run the fused LDS kernel in its two shapes, longer ones the global-memory
    // path, and reads the fused kernel declines are listed for the global path too. With more
    // than one bucket in use the reads are sorted by length on the device first.
    if (r0 != 0 || n != b->n) return fail(DNALN_E_INTERNAL, "batched form runs the whole batch as one chunk");
    const uint64_t na = b->n_short, nb = b->n_mid;
    const int used = (na > 0) + (nb > na) + (n > nb);
    const uint32_t* lorder = nullptr;
    DN_TRY(b->span_begin(0, sp_i));
    if (used > 1) {
      DN_TRY(b->lids.reserve((uint64_t)n * 4, st));
      DN_TRY(b->lkeys.reserve((uint64_t)n * 4, st));
      DN_TRY(b->lorder.reserve((uint64_t)n * 4, st));
      hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, st, b->lids.as<uint32_t>(), n);
      DN_HIP(hipGetLastError());
      const int bits = 32 - __builtin_clz((unsigned)std::max(b->maxlen, 1));
      const uint32_t* lens_u = reinterpret_cast<const uint32_t*>(b->lens.as<int32_t>());
      size_t bytes = 0;
      DN_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, lens_u, b->lkeys.as<uint32_t>(), b->lids.as<uint32_t>(),
                                                b->lorder.as<uint32_t>(), (uint32_t)n, 0, bits, st));
      DN_TRY(b->ltmp.reserve(bytes, st));
      DN_HIP(hipcub::DeviceRadixSort::SortPairs(b->ltmp.p, bytes, lens_u, b->lkeys.as<uint32_t>(), b->lids.as<uint32_t>(),
                                                b->lorder.as<uint32_t>(), (uint32_t)n, 0, bits, st));
      lorder = b->lorder.as<uint32_t>();
    }
    DN_TRY(b->ovf.reserve((uint64_t)n * 4 + 8, st));
    uint32_t* ovf = b->ovf.as<uint32_t>();
    uint32_t* novf_d = ovf + n;
    DN_HIP(hipMemsetAsync(novf_d, 0, 4, st));
    const uint64_t lo[2] = {0, na}, hi[2] = {na, nb};
    for (int x = 0; x < 2; ++x) {
      if (hi[x] <= lo[x]) continue;
      const FusedShape fs = fused_shape(c->k, c->w, std::min(b->split[x], b->maxlen));
      DN_TRY(b->fmins.reserve(fused_mins_words(fs, hi[x] - lo[x]) * 8, st));
    }
    for (int x = 0; x < 2; ++x) {
      if (hi[x] <= lo[x]) continue;
      const FusedShape fs = fused_shape(c->k, c->w, std::min(b->split[x], b->maxlen));
      sa.list = lorder ? lorder + lo[x] : nullptr;
      sa.n_reads = (uint32_t)(hi[x] - lo[x]);
      DN_TRY(launch_fused(fs, sa, ca, b->fmins.as<uint64_t>(), ovf, novf_d, st));
    }
    uint32_t novf = 0;
    DN_HIP(hipMemcpyAsync(&novf, novf_d, 4, hipMemcpyDeviceToHost, st));
    DN_HIP(hipStreamSynchronize(st));
    // Reads a 32-anchor shape declined (more minimizers or anchors than its slots) get a second
    // fused pass in the 64-anchor shape; only what that declines too takes the global path. When
    // long reads take the global path anyway, the declined reads join them instead: on 1M uneven
    // reads (50-1,000 bp, bw 16) that saved 0.15 ms of a 13.5 ms step, while on 2M x 150 bp (no
    // long reads) the global path for the ~2,800 declined reads alone cost 0.1 ms more than the
    // second pass (records equal either way).
    bool small = false;
    for (int x = 0; x < 2; ++x)
      if (hi[x] > lo[x]) small = small || fused_shape(c->k, c->w, std::min(b->split[x], b->maxlen)).cap < 64;
    const FusedShape retry = fused_retry_shape(c->k, c->w, b->maxlen);
    if (novf && small && retry.cap && n <= nb) {
      DN_TRY(b->ovf2.reserve((uint64_t)n * 4 + 8, st));
      uint32_t* ovf2 = b->ovf2.as<uint32_t>();
      uint32_t* novf2_d = ovf2 + n;
      DN_HIP(hipMemsetAsync(novf2_d, 0, 4, st));
      DN_TRY(b->fmins.reserve(fused_mins_words(retry, novf) * 8, st));
      sa.list = ovf;
      sa.n_reads = novf;
      DN_TRY(launch_fused(retry, sa, ca, b->fmins.as<uint64_t>(), ovf2, novf2_d, st));
      b->stats.fused_retry_reads += novf;
      DN_HIP(hipMemcpyAsync(&novf, novf2_d, 4, hipMemcpyDeviceToHost, st));
      DN_HIP(hipStreamSynchronize(st));
      ovf = ovf2;
    }
    DN_TRY(b->span_end(sp_i));
    const uint32_t* list = ovf;
    uint32_t nl = novf;
    if (n > nb) {
      if (lorder) {
        DN_HIP(hipMemcpyAsync(ovf + novf, lorder + nb, (n - nb) * 4, hipMemcpyDeviceToDevice, st));
        nl += (uint32_t)(n - nb);
      } else {  // every read is long: the global path in read order
        list = nullptr;
        nl = n;
      }
    }
    b->stats.global_path_reads += nl;
    if (nl) DN_TRY(seed_chain_global(b, sa, ca, list, nl, sp_i, true));
  } else {
    DN_TRY(seed_chain_global(b, sa, ca, nullptr, n, sp_i));
  }

  // 3. rank pass
  DN_TRY(b->rcnt.reserve((uint64_t)n * 4 + 4, st));
  DN_TRY(b->roff.reserve((uint64_t)n * 8 + 8, st));
  if (n) hipLaunchKernelGGL(k_rank_count, dim3(sgrid), dim3(kSeedBlock), 0, st, b->ncand.as<int32_t>(), b->rcnt.as<uint32_t>(), n);
  DN_TRY(exclusive_scan(b->scan_tmp, b->rcnt.as<uint32_t>(), b->roff.as<uint64_t>(), n, st));
  uint64_t nr = 0;
  DN_TRY(scan_total(st, b->roff.as<uint64_t>(), b->rcnt.as<uint32_t>(), n, nr));
  DN_TRY(b->rtasks.reserve(nr * sizeof(SwTask) + sizeof(SwTask), st));
  DN_TRY(b->rres.reserve(nr * sizeof(SwResult) + sizeof(SwResult), st));
  if (n)
    hipLaunchKernelGGL(k_rank_tasks, dim3(sgrid), dim3(kSeedBlock), 0, st, rv, g, b->cands.as<Cand>(),
                       b->ncand.as<int32_t>(), max_cand, b->roff.as<uint64_t>(), b->rtasks.as<SwTask>(), n);
  DN_TRY(run_sw(sp, b->rtasks.as<SwTask>(), (uint32_t)nr, qmin, b->maxlen, false, b->rres.as<SwResult>(),
                nullptr, 0, b->sw, st, b->launches, b, 2));
  b->stats.rank_tasks += nr;

  // 4. select + traceback pass
  DN_TRY(b->chosen.reserve((uint64_t)n * 4 + 4, st));
  DN_TRY(b->second.reserve((uint64_t)n * 4 + 4, st));
  DN_TRY(b->tbflag.reserve((uint64_t)n * 4 + 4, st));
  DN_TRY(b->tboff.reserve((uint64_t)n * 8 + 8, st));
  if (n)
    hipLaunchKernelGGL(k_select, dim3(sgrid), dim3(kSeedBlock), 0, st, rv, b->ncand.as<int32_t>(),
                       b->roff.as<uint64_t>(), b->rres.as<SwResult>(), b->chosen.as<int32_t>(),
                       b->second.as<int32_t>(), b->tbflag.as<uint32_t>(), n);
  DN_TRY(exclusive_scan(b->scan_tmp, b->tbflag.as<uint32_t>(), b->tboff.as<uint64_t>(), n, st));
  uint64_t nt = 0;
  DN_TRY(scan_total(st, b->tboff.as<uint64_t>(), b->tbflag.as<uint32_t>(), n, nt));
  DN_TRY(b->ttasks.reserve(nt * sizeof(SwTask) + sizeof(SwTask), st));
  DN_TRY(b->tres.reserve(nt * sizeof(SwResult) + sizeof(SwResult), st));
  DN_TRY(b->tcig.reserve(nt * (uint64_t)cap * 4 + 4, st));
  if (n)
    hipLaunchKernelGGL(k_tb_tasks, dim3(sgrid), dim3(kSeedBlock), 0, st, rv, g, b->cands.as<Cand>(), max_cand,
                       b->chosen.as<int32_t>(), b->tboff.as<uint64_t>(), b->ttasks.as<SwTask>(), n);
  DN_TRY(run_sw(sp, b->ttasks.as<SwTask>(), (uint32_t)nt, qmin, b->maxlen, true, b->tres.as<SwResult>(),
                b->tcig.as<uint32_t>(), cap, b->sw, st, b->launches, b, 3));
  b->stats.tb_tasks += nt;

  // 5. records and CIGARs
  DN_TRY(b->ciglen.reserve((uint64_t)n * 4 + 4, st));
  DN_TRY(b->cigoff.reserve((uint64_t)n * 8 + 8, st));
  dnaln_record* rec = b->rec.as<dnaln_record>() + r0;
  if (n)
    hipLaunchKernelGGL(k_records, dim3(sgrid), dim3(kSeedBlock), 0, st, rv, g, b->cands.as<Cand>(),
                       b->ncand.as<int32_t>(), max_cand, b->chosen.as<int32_t>(), b->second.as<int32_t>(),
                       b->tboff.as<uint64_t>(), b->tres.as<SwResult>(), p.min_aln_score, c->lut.as<uint8_t>(),
                       c->lut_len, rec, b->ciglen.as<uint32_t>(), b->err.as<int>(), n);
  DN_TRY(exclusive_scan(b->scan_tmp, b->ciglen.as<uint32_t>(), b->cigoff.as<uint64_t>(), n, st));
  uint64_t ncig = 0;
  DN_TRY(scan_total(st, b->cigoff.as<uint64_t>(), b->ciglen.as<uint32_t>(), n, ncig));
  // The pool grows across chunks; keep what earlier chunks wrote.
  const uint64_t need = (b->pool_used + ncig) * 4 + 4;
  if (need > b->pool.cap) {
    DevBuf bigger;
    DN_TRY(bigger.reserve(need * 2));
    if (b->pool_used) DN_HIP(hipMemcpyAsync(bigger.p, b->pool.p, b->pool_used * 4, hipMemcpyDeviceToDevice, st));
    DN_HIP(hipStreamSynchronize(st));
    std::swap(b->pool.p, bigger.p);
    std::swap(b->pool.cap, bigger.cap);
  }
  if (n)
    hipLaunchKernelGGL(k_cigar_gather, dim3(sgrid), dim3(kSeedBlock), 0, st, rv, b->chosen.as<int32_t>(),
                       b->tboff.as<uint64_t>(), b->tres.as<SwResult>(), b->tcig.as<uint32_t>(), cap,
                       b->cigoff.as<uint64_t>(), b->pool_used, rec, b->pool.as<uint32_t>() + b->pool_used, n);
  DN_HIP(hipGetLastError());
  b->pool_used += ncig;
  return DNALN_OK;
}

// k-mer starts per index-build thread; the overlap each thread re-reads is 2(w - 1) + k - 1 bases.
constexpr uint32_t kIndexSegment = 1024;

// Builds MinimizerIndex (src/index.cpp:10-45) for the reference codes at `ref` (device memory
// with guard bytes on both sides) into `table` (the host build's form: no inlined occurrences)
// and `pos`, on stream st. The bucket count and pos[] equal the host build's; a key may sit in
// another slot of its probe chain, so lookups return the same slices.
int build_index_dev(const uint8_t* ref, const uint64_t* coff, const uint32_t* clen, int nc, int k, int w,
                    hipStream_t st, DevBuf& table, DevBuf& pos, IndexInfo& info) {
  if (k < 4 || k > 16)
    return fail(DNALN_E_INVALID, "the GPU index build needs k in [4, 16] (the hash fills the upper word of a "
                                 "64-bit sort key); use a host-built index");
  if (w < 1 || w > 255) return fail(DNALN_E_INVALID, "w must be in [1, 255]");
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<IndexSeg> segs;
  for (int c = 0; c < nc; ++c) {
    if (clen[c] < (uint32_t)k) continue;
    const uint32_t nk = clen[c] - (uint32_t)k + 1;
    for (uint32_t s = 0; s < nk; s += kIndexSegment)
      segs.push_back(IndexSeg{coff[c], clen[c], s, std::min(nk, s + kIndexSegment), 0});
  }
  const uint32_t nseg = (uint32_t)segs.size();
  DevBuf dsegs, ring_h, ring_p, counts, offs, keys, keys2, flags, ridx, starts, tmp, mx;
  uint64_t n_occ = 0, n_keys = 0;
  uint32_t max_count = 0;
  if (nseg) {
    DN_TRY(dsegs.reserve(nseg * sizeof(IndexSeg)));
    DN_HIP(hipMemcpyAsync(dsegs.p, segs.data(), nseg * sizeof(IndexSeg), hipMemcpyHostToDevice, st));
    DN_TRY(ring_h.reserve((uint64_t)nseg * w * 8));
    DN_TRY(ring_p.reserve((uint64_t)nseg * w * 4));
    DN_TRY(counts.reserve((nseg + 1ull) * 8));
    DN_TRY(offs.reserve((nseg + 1ull) * 8));
    hipLaunchKernelGGL(k_ix_count, dim3((nseg + 1 + 255) / 256), dim3(256), 0, st, ref, dsegs.as<IndexSeg>(), nseg, k,
                       w, ring_h.as<uint64_t>(), ring_p.as<uint32_t>(), counts.as<uint64_t>());
    DN_HIP(hipGetLastError());
    DN_TRY(scan_u64(tmp, counts.as<uint64_t>(), offs.as<uint64_t>(), nseg + 1, st));
    DN_HIP(hipMemcpyAsync(&n_occ, offs.as<uint64_t>() + nseg, 8, hipMemcpyDeviceToHost, st));
    DN_HIP(hipStreamSynchronize(st));
  }
  if (n_occ >= (1ull << 31)) return fail(DNALN_E_INVALID, "too many minimizer occurrences for the GPU index build");
  DN_TRY(pos.reserve(n_occ * 4 + 4));
  if (n_occ) {
    DN_TRY(keys.reserve(n_occ * 8));
    DN_TRY(keys2.reserve(n_occ * 8));
    hipLaunchKernelGGL(k_ix_emit, dim3((nseg + 255) / 256), dim3(256), 0, st, ref, dsegs.as<IndexSeg>(), nseg, k, w,
                       ring_h.as<uint64_t>(), ring_p.as<uint32_t>(), offs.as<uint64_t>(), keys.as<uint64_t>());
    DN_HIP(hipGetLastError());
    size_t bytes = 0;
    DN_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(), (int)n_occ, 0,
                                             32 + 2 * k, st));
    DN_TRY(tmp.reserve(bytes));
    DN_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(), (int)n_occ, 0,
                                             32 + 2 * k, st));
    DN_TRY(flags.reserve((n_occ + 1) * 8));
    DN_TRY(ridx.reserve((n_occ + 1) * 8));
    const unsigned g = (unsigned)((n_occ + 1 + 255) / 256);
    hipLaunchKernelGGL(k_ix_flags, dim3(g), dim3(256), 0, st, keys2.as<uint64_t>(), n_occ, flags.as<uint64_t>(),
                       pos.as<uint32_t>());
    DN_HIP(hipGetLastError());
    DN_TRY(scan_u64(tmp, flags.as<uint64_t>(), ridx.as<uint64_t>(), (uint32_t)(n_occ + 1), st));
    DN_HIP(hipMemcpyAsync(&n_keys, ridx.as<uint64_t>() + n_occ, 8, hipMemcpyDeviceToHost, st));
    DN_HIP(hipStreamSynchronize(st));
    DN_TRY(starts.reserve((n_keys + 1) * 4));
    hipLaunchKernelGGL(k_ix_starts, dim3(g), dim3(256), 0, st, flags.as<uint64_t>(), ridx.as<uint64_t>(), n_occ,
                       starts.as<uint32_t>());
    DN_HIP(hipGetLastError());
  }
  uint64_t cap = 16;  // src/index.cpp:28-30
  while (cap < n_keys * 2) cap <<= 1;
  DN_TRY(table.reserve(cap * sizeof(Bucket)));
  hipLaunchKernelGGL(k_ix_clear, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, st, table.as<Bucket>(), cap);
  DN_HIP(hipGetLastError());
  if (n_keys) {
    DN_TRY(mx.reserve(4));
    DN_HIP(hipMemsetAsync(mx.p, 0, 4, st));
    hipLaunchKernelGGL(k_ix_insert, dim3((unsigned)((n_keys + 255) / 256)), dim3(256), 0, st, keys2.as<uint64_t>(),
                       starts.as<uint32_t>(), n_keys, table.as<Bucket>(), cap - 1, mx.as<uint32_t>());
    DN_HIP(hipGetLastError());
    DN_HIP(hipMemcpyAsync(&max_count, mx.p, 4, hipMemcpyDeviceToHost, st));
  }
  DN_HIP(hipStreamSynchronize(st));
  info.n_buckets = cap;
  info.n_pos = n_occ;
  info.n_keys = n_keys;
  info.max_count = max_count;
  info.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return DNALN_OK;
}

// The 8-byte probe table and pos2 (probe_table.hpp) from the context's inlined 16-byte table,
// where hashes fit 30 bits (k <= 15); *built = false leaves the probes on the 16-byte table.
int build_probe_table(const DevBuf& table, uint64_t n, const DevBuf& pos, uint32_t max_occ, int k, hipStream_t st,
                      DevBuf& ctab, DevBuf& pos2, bool* built) {
  *built = false;
  if (2 * k > 30 || n == 0 || n + 1 > 0xffffffffull) return DNALN_OK;
  DevBuf need, at, tmp;
  DN_TRY(need.reserve((n + 1) * 4));
  DN_TRY(at.reserve((n + 1) * 8));
  const unsigned blocks = (unsigned)((n + 1 + 255) / 256);
  hipLaunchKernelGGL(k_ct_need, dim3(blocks), dim3(256), 0, st, table.as<Bucket>(), n, max_occ, need.as<uint32_t>());
  DN_HIP(hipGetLastError());
  DN_TRY(exclusive_scan(tmp, need.as<uint32_t>(), at.as<uint64_t>(), (uint32_t)(n + 1), st));
  uint64_t total = 0;
  DN_HIP(hipMemcpyAsync(&total, at.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, st));
  DN_HIP(hipStreamSynchronize(st));
  if (total >= 0xffffffffull) return DNALN_OK;
  DN_TRY(ctab.reserve((n + kCTail) * 8));
  DN_TRY(pos2.reserve(total * 4 + 4));
  hipLaunchKernelGGL(k_ct_fill, dim3((unsigned)((n + kCTail + 255) / 256)), dim3(256), 0, st, table.as<Bucket>(), n,
                     max_occ, pos.as<uint32_t>(), at.as<uint64_t>(), ctab.as<uint64_t>(), pos2.as<uint32_t>());
  DN_HIP(hipGetLastError());
  DN_HIP(hipStreamSynchronize(st));  // the temporaries are freed on return
  *built = true;
  return DNALN_OK;
}

// Reference codes in device memory between two guard regions of N (kSeqGuard bytes each): the
// chunked readers may touch bytes just outside a contig.
int upload_reference(DevBuf& buf, const uint8_t* codes, uint64_t len, hipStream_t st, uint8_t** base) {
  DN_TRY(buf.reserve(len + 2 * kSeqGuard));
  DN_HIP(hipMemsetAsync(buf.p, 4, len + 2 * kSeqGuard, st));
  *base = buf.as<uint8_t>() + kSeqGuard;
  if (len) DN_HIP(hipMemcpyAsync(*base, codes, len, hipMemcpyHostToDevice, st));
  DN_HIP(hipStreamSynchronize(st));
  return DNALN_OK;
}

}  // namespace

// ------------------------------------------------------------------ alignment without allocation
// dnaln_hip_align_static works entirely inside this array: one copy per GPU, held for the life of
// the process. In host code the name `scratch` is a host-side stand-in, not a device address, so
// host code never hands it to a kernel or a copy: it uses the address hipGetSymbolAddress gave
// for the device (scratch_base, looked up once per device with that device current).
static __device__ __align__(256) uint8_t scratch[DNALN_STATIC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
uint8_t* g_scratch_at[256] = {};  // by device ordinal

int scratch_base(int device, uint8_t** out) {
  if (device < 0 || device >= 256) return fail(DNALN_E_INVALID, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_at[device]) {
    void* p = nullptr;
    DN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_at[device] = static_cast<uint8_t*>(p);
  }
  *out = g_scratch_at[device];
  return DNALN_OK;
}

// Byte offsets of one dnaln_hip_align_static call's buffers inside the array (256-aligned).
struct StaticLayout {
  uint64_t ctl, pool, tin, tasks, res, cig, dir, rows, off, out, outcig, end;
};

// Status words of a static call: the CIGAR total and error bits (1: a CIGAR overflowed its slot
// or the walk met an inconsistent cell).
struct StaticCtl {
  uint64_t total;
  uint32_t err;
  uint32_t reserved;
};

StaticLayout static_layout(uint64_t n, uint64_t pool_len, int qmax, int W, bool packed, bool tb) {
  auto up = [](uint64_t x) { return (x + 255) & ~(uint64_t)255; };
  const uint64_t cap = 2 * (uint64_t)qmax + 4;
  StaticLayout L{};
  uint64_t o = 0;
  L.ctl = o;
  o += 256;
  L.pool = o;
  o = up(o + pool_len + 2 * kReadGuard);
  L.tin = o;
  o = up(o + n * sizeof(dnaln_task));
  L.tasks = o;
  o = up(o + n * sizeof(SwTask));
  L.res = o;
  o = up(o + n * sizeof(SwResult));
  L.cig = o;
  o = up(o + (tb ? n * cap * 4 : 0));
  L.dir = o;
  if (packed) {
    const uint64_t waves = (n + 127) / 128, row_words = (uint64_t)((W + 7) / 8) * 64;
    o = up(o + (tb ? waves * fill_rows(qmax) * row_words * 8 : 0));
    L.rows = o;
  } else {
    o = up(o + (tb ? n * (uint64_t)std::max(qmax, 1) * W : 0));
    L.rows = o;
    o = up(o + n * 2 * (uint64_t)(W + 1) * 4);
  }
  L.off = o;
  o = up(o + n * 8);
  L.out = o;
  o = up(o + n * sizeof(dnaln_result));
  L.outcig = o;
  o = up(o + (tb ? n * cap * 4 : 0));
  L.end = o;
  return L;
}

// Exclusive scan of the CIGAR lengths of n results, by one block: off[i], and the total with
// the error bit of any failed walk in ctl.
__global__ __launch_bounds__(1024) void k_static_scan(const SwResult* res, uint32_t n, uint64_t* off, StaticCtl* ctl) {
  __shared__ uint64_t part[1024];
  __shared__ uint64_t carry;
  __shared__ uint32_t err;
  const uint32_t t = threadIdx.x;
  if (t == 0) carry = 0, err = 0;
  __syncthreads();
  for (uint32_t base = 0; base < n; base += 1024) {
    const uint32_t i = base + t;
    uint64_t v = 0;
    if (i < n) {
      const SwResult r = res[i];
      if (r.n_cigar < 0 || r.status != 0) atomicOr(&err, 1u);
      else v = (uint64_t)r.n_cigar;
    }
    part[t] = v;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {  // inclusive Hillis-Steele scan
      const uint64_t x = t >= d ? part[t - d] : 0;
      __syncthreads();
      part[t] += x;
      __syncthreads();
    }
    if (i < n) off[i] = carry + part[t] - v;
    __syncthreads();
    if (t == 1023) carry += part[1023];
    __syncthreads();
  }
  if (t == 0) {
    ctl->total = carry;
    ctl->err = err;
  }
}

// AlignResult of task i, and its CIGAR in forward order at outcig[off[i]] (the walk stored it
// reversed in the task's slot of cap words).
__global__ void k_static_out(const SwResult* res, const uint32_t* slots, int cap, const uint64_t* off, uint32_t n,
                             dnaln_result* out, uint32_t* outcig) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SwResult r = res[i];
  dnaln_result o;
  o.score = r.score;
  o.qb = r.qb;
  o.qe = r.qe;
  o.tb = r.tb;
  o.te = r.te;
  o.nm = r.nm;
  o.cigar_off = off[i];
  o.cigar_len = r.n_cigar > 0 && r.status == 0 ? (uint32_t)r.n_cigar : 0u;
  o.reserved = 0;
  out[i] = o;
  const uint32_t* s = slots + (uint64_t)i * (uint64_t)cap;
  for (uint32_t k = 0; k < o.cigar_len; ++k) outcig[o.cigar_off + k] = s[o.cigar_len - 1 - k];
}

}  // namespace

// A GPU-built index held on its device until downloaded (dnaln_hip_index_build).
struct dnaln_hip_index {
  int device = 0;
  DevBuf table, pos;
  IndexInfo info;
};
static_assert(sizeof(Bucket) == sizeof(dnaln_bucket), "device bucket layout");

extern "C" {

void dnaln_hip_default_params(dnaln_map_params* p) {
  std::memset(p, 0, sizeof *p);
  p->k = 15;
  p->w = 10;
  p->max_occ = 500;
  p->sc = {2, 4, 4, 2, 1, 16};
  p->ch = {5000, 16, 50, 20, 2};
  p->max_cand = 10;
  p->pri_ratio = 0.5;
  p->min_aln_score = 40;
}

const char* dnaln_hip_last_error(void) { return g_err.c_str(); }

int dnaln_hip_create(const dnaln_map_params* params, const uint8_t* ref_codes, uint64_t ref_len,
                     const dnaln_contig* contigs, int32_t n_contigs, const dnaln_index_view* index, int32_t device,
                     dnaln_hip_ctx** out) {
  return guarded([&]() -> int {
    if (!params || !out || (!ref_codes && ref_len) || n_contigs < 1 || !contigs)
      return fail(DNALN_E_INVALID, "null argument");
    DN_TRY(validate(*params));
    if (index && (index->k != params->k || index->w != params->w))
      return fail(DNALN_E_INVALID, "index k/w differ from params");
    if (index && (index->n_buckets == 0 || (index->n_buckets & (index->n_buckets - 1))))
      return fail(DNALN_E_INVALID, "index bucket count must be a power of two");
    if (ref_len >= (1ull << 31)) return fail(DNALN_E_INVALID, "reference is too large (limit: 2^31 bp)");
    std::unique_ptr<dnaln_hip_ctx> c(new dnaln_hip_ctx);
    c->p = *params;
    c->p.ch.band = std::min(c->p.ch.band, c->p.sc.bw);  // src/mapper.cpp:9
    c->device = device;
    DN_HIP(hipSetDevice(device));
    DN_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    DN_TRY(scratch_base(device, &c->scratch));
    c->ref_len = ref_len;
    c->nc = n_contigs;
    c->k = params->k;
    c->w = params->w;
    // The packed kernel streams aligned dwords and may touch bytes just outside a window: the
    // reference sits between two guard regions of kSeqGuard bytes (never used as sequence).
    DN_TRY(upload_reference(c->ref, ref_codes, ref_len, c->stream, &c->ref_base));
    c->h_coff.resize(n_contigs);
    c->h_clen.resize(n_contigs);
    for (int i = 0; i < n_contigs; ++i) {
      c->h_coff[i] = contigs[i].offset;
      c->h_clen[i] = contigs[i].len;
      if (contigs[i].offset + contigs[i].len > ref_len) return fail(DNALN_E_INVALID, "contig outside the reference");
    }
    DN_TRY(c->coff.reserve(n_contigs * 8));
    DN_TRY(c->clen.reserve(n_contigs * 4));
    DN_HIP(hipMemcpy(c->coff.p, c->h_coff.data(), n_contigs * 8, hipMemcpyHostToDevice));
    DN_HIP(hipMemcpy(c->clen.p, c->h_clen.data(), n_contigs * 4, hipMemcpyHostToDevice));
    if (index) {
      c->ix.n_buckets = index->n_buckets;
      c->ix.n_pos = index->n_pos;
      DN_TRY(c->table.reserve(index->n_buckets * sizeof(dnaln_bucket)));
      DN_HIP(hipMemcpy(c->table.p, index->table, index->n_buckets * sizeof(dnaln_bucket), hipMemcpyHostToDevice));
      DN_TRY(c->pos.reserve(index->n_pos * 4 + 4));
      if (index->n_pos) DN_HIP(hipMemcpy(c->pos.p, index->pos, index->n_pos * 4, hipMemcpyHostToDevice));
    } else {
      DN_TRY(build_index_dev(c->ref_base, c->h_coff.data(), c->h_clen.data(), n_contigs, c->k, c->w, c->stream, c->table,
                             c->pos, c->ix));
    }
    c->mask = c->ix.n_buckets - 1;
    c->n_pos = c->ix.n_pos;
    hipLaunchKernelGGL(k_inline_unique, dim3((unsigned)((c->ix.n_buckets + 255) / 256)), dim3(256), 0, c->stream,
                       c->table.as<Bucket>(), c->ix.n_buckets, c->pos.as<uint32_t>());
    DN_HIP(hipGetLastError());
    DN_TRY(build_probe_table(c->table, c->ix.n_buckets, c->pos, (uint32_t)c->p.max_occ, c->k, c->stream, c->ctab,
                             c->pos2, &c->probe8));
    DN_HIP(hipStreamSynchronize(c->stream));
    // Chain gap cost by drift dd (src/chain.cpp:36), evaluated in double exactly as the reference.
    std::vector<int32_t> cost(c->p.ch.band + 1, 0);
    for (int dd = 1; dd <= c->p.ch.band; ++dd)
      cost[dd] = (int)(0.01 * c->k * (double)dd + 0.5 * (63 - __builtin_clzll((uint64_t)dd)));
    c->h_cost = cost;
    DN_TRY(c->cost.reserve(cost.size() * 4));
    DN_HIP(hipMemcpy(c->cost.p, cost.data(), cost.size() * 4, hipMemcpyHostToDevice));
    // MAPQ by score difference (src/mapper.cpp:212-214).
    std::vector<uint8_t> lut;
    for (int d = 0;; ++d) {
      const double q = 6.02 * d / c->p.sc.match + 0.499;
      const uint8_t v = (uint8_t)std::min(60.0, q);
      if (v >= 60) break;
      lut.push_back(v);
    }
    c->lut_len = (int)lut.size();
    DN_TRY(c->lut.reserve(lut.size() + 1));
    if (!lut.empty()) DN_HIP(hipMemcpy(c->lut.p, lut.data(), lut.size(), hipMemcpyHostToDevice));
    *out = c.release();
    return DNALN_OK;
  });
}

void dnaln_hip_destroy(dnaln_hip_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  if (c->batch) dnaln_hip_batch_destroy(c->batch);
  delete c;
}

int dnaln_hip_align(dnaln_hip_ctx* c, const dnaln_score_params* spp, const dnaln_task* tasks, uint64_t n,
                    const uint8_t* seq_pool, uint64_t pool_len, int traceback, dnaln_result* out,
                    uint32_t* cigar_pool, uint64_t cigar_cap) {
  return guarded([&]() -> int {
    if (!c || (n && (!tasks || !out || !seq_pool))) return fail(DNALN_E_INVALID, "null argument");
    if (n >= (1ull << 31)) return fail(DNALN_E_INVALID, "too many tasks");
    const dnaln_score_params s = spp ? *spp : c->p.sc;
    if (s.bw < 1 || s.bw > 1000 || s.match < 1 || s.mismatch < 0 || s.gap_open < 0 || s.gap_ext < 1 || s.ambig < 0)
      return fail(DNALN_E_INVALID, "score parameters outside the validated range");
    const ScoreP sp{s.match, s.mismatch, s.gap_open, s.gap_ext, s.ambig, s.bw};
    DN_HIP(hipSetDevice(c->device));
    int qmax = 0, qmin = n ? tasks[0].qlen : 0;
    for (uint64_t i = 0; i < n; ++i) {
      if (tasks[i].qlen < 0 || tasks[i].tlen < 0 || tasks[i].q_off + (uint64_t)tasks[i].qlen > pool_len ||
          tasks[i].t_off + (uint64_t)tasks[i].tlen > pool_len)
        return fail(DNALN_E_INVALID, "task outside the sequence pool");
      qmax = std::max(qmax, tasks[i].qlen);
      qmin = std::min(qmin, tasks[i].qlen);
    }
    if (n == 0) return DNALN_OK;
    const int cap = 2 * qmax + 4;
    hipStream_t st = c->stream;
    DN_TRY(c->a_pool.reserve(pool_len + 2 * kReadGuard));
    DN_TRY(c->a_in.reserve(n * sizeof(dnaln_task)));
    DN_TRY(c->a_tasks.reserve(n * sizeof(SwTask)));
    DN_TRY(c->a_res.reserve(n * sizeof(SwResult)));
    if (traceback) DN_TRY(c->a_cig.reserve(n * (uint64_t)cap * 4));
    DN_HIP(hipMemcpyAsync(c->a_pool.as<uint8_t>() + kReadGuard, seq_pool, pool_len, hipMemcpyHostToDevice, st));
    DN_HIP(hipMemcpyAsync(c->a_in.p, tasks, n * sizeof(dnaln_task), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_pool_tasks, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, c->a_in.as<dnaln_task>(),
                       (uint64_t)(uintptr_t)(c->a_pool.as<uint8_t>() + kReadGuard), sp.bw, c->a_tasks.as<SwTask>(),
                       (uint32_t)n);
    DN_HIP(hipGetLastError());
    DN_TRY(run_sw(sp, c->a_tasks.as<SwTask>(), (uint32_t)n, qmin, qmax, traceback != 0, c->a_res.as<SwResult>(),
                  c->a_cig.as<uint32_t>(), cap, c->a_sw, st, nullptr));
    std::vector<SwResult> res(n);
    DN_HIP(hipMemcpyAsync(res.data(), c->a_res.p, n * sizeof(SwResult), hipMemcpyDeviceToHost, st));
    std::vector<uint32_t> cig;
    if (traceback) {
      cig.resize(n * (uint64_t)cap);
      DN_HIP(hipMemcpyAsync(cig.data(), c->a_cig.p, n * (uint64_t)cap * 4, hipMemcpyDeviceToHost, st));
    }
    DN_HIP(hipStreamSynchronize(st));
    uint64_t used = 0;
    for (uint64_t i = 0; i < n; ++i) {
      const SwResult& r = res[i];
      if (r.status != 0 || r.n_cigar < 0) return fail(DNALN_E_INTERNAL, "traceback failed on task " + std::to_string(i));
      dnaln_result& o = out[i];
      o.score = r.score;
      o.qb = r.qb;
      o.qe = r.qe;
      o.tb = r.tb;
      o.te = r.te;
      o.nm = r.nm;
      o.cigar_off = used;
      o.cigar_len = (uint32_t)r.n_cigar;
      o.reserved = 0;
      if (r.n_cigar > 0) {
        if (!cigar_pool || used + (uint64_t)r.n_cigar > cigar_cap)
          return fail(DNALN_E_CAPACITY, "cigar_pool too small");
        for (int k = 0; k < r.n_cigar; ++k) cigar_pool[used + k] = cig[i * (uint64_t)cap + (r.n_cigar - 1 - k)];
        used += (uint64_t)r.n_cigar;
      }
    }
    return DNALN_OK;
  });
}

int dnaln_hip_align_static(dnaln_hip_ctx* c, const dnaln_score_params* spp, const dnaln_task* tasks, uint64_t n,
                           const uint8_t* seq_pool, uint64_t pool_len, int traceback, dnaln_result* out,
                           uint32_t* cigar_pool, uint64_t cigar_cap) {
  return guarded([&]() -> int {
    if (!c || (n && (!tasks || !out || !seq_pool))) return fail(DNALN_E_INVALID, "null argument");
    if (n >= (1ull << 31)) return fail(DNALN_E_INVALID, "too many tasks");
    const dnaln_score_params s = spp ? *spp : c->p.sc;
    if (s.bw < 1 || s.bw > 1000 || s.match < 1 || s.mismatch < 0 || s.gap_open < 0 || s.gap_ext < 1 || s.ambig < 0)
      return fail(DNALN_E_INVALID, "score parameters outside the validated range");
    const ScoreP sp{s.match, s.mismatch, s.gap_open, s.gap_ext, s.ambig, s.bw};
    int qmax = 0, qmin = n ? tasks[0].qlen : 0;
    for (uint64_t i = 0; i < n; ++i) {
      if (tasks[i].qlen < 0 || tasks[i].tlen < 0 || tasks[i].q_off + (uint64_t)tasks[i].qlen > pool_len ||
          tasks[i].t_off + (uint64_t)tasks[i].tlen > pool_len)
        return fail(DNALN_E_INVALID, "task outside the sequence pool");
      qmax = std::max(qmax, tasks[i].qlen);
      qmin = std::min(qmin, tasks[i].qlen);
    }
    if (n == 0) return DNALN_OK;
    const int W = 2 * sp.bw + 1, cap = 2 * qmax + 4;
    const bool tb = traceback != 0;
    const bool packed = packed_bw_available(sp.bw) && packed_ok(sp, qmax);
    const StaticLayout L = static_layout(n, pool_len, qmax, W, packed, tb);
    if (L.end > DNALN_STATIC_SCRATCH_BYTES)
      return fail(DNALN_E_CAPACITY, "the call needs " + std::to_string(L.end) + " bytes of device scratch, more than " +
                                        std::to_string((unsigned long long)DNALN_STATIC_SCRATCH_BYTES));
    DN_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    uint8_t* base = c->scratch;
    auto at = [base](uint64_t o) { return base + o; };
    uint8_t* pool = at(L.pool) + kReadGuard;
    DN_HIP(hipMemcpyAsync(pool, seq_pool, pool_len, hipMemcpyHostToDevice, st));
    DN_HIP(hipMemcpyAsync(at(L.tin), tasks, n * sizeof(dnaln_task), hipMemcpyHostToDevice, st));
    SwTask* tk = reinterpret_cast<SwTask*>(at(L.tasks));
    SwResult* res = reinterpret_cast<SwResult*>(at(L.res));
    uint32_t* cig = reinterpret_cast<uint32_t*>(at(L.cig));
    hipLaunchKernelGGL(k_pool_tasks, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       reinterpret_cast<const dnaln_task*>(at(L.tin)), (uint64_t)(uintptr_t)pool, sp.bw, tk, (uint32_t)n);
    DN_HIP(hipGetLastError());
    if (packed) {
      // The fill of dnaln_hip_align in one launch, tasks in call order, fixed-size slots.
      SwArgs a{};
      a.sp = sp;
      score_table(sp, a.tab0, a.tab1);
      a.cigar_cap = cap;
      a.tasks = tk;
      a.order = nullptr;
      a.res = res;
      a.cigar = cig;
      a.n = (uint32_t)n;
      a.dir = tb ? reinterpret_cast<uint64_t*>(at(L.dir)) : nullptr;
      a.dir_slot_words = fill_rows(qmax) * (uint64_t)((W + 7) / 8) * 64;
      a.woff = nullptr;
      a.woff_base = 0;
      const unsigned lw = (unsigned)((n + 127) / 128);
      const unsigned grid = (lw + kSwBlock / 64 - 1) / (kSwBlock / 64);
      const bool ok = tb ? launch_packed_bw<true>(sp.bw, a, grid, st) : launch_packed_bw<false>(sp.bw, a, grid, st);
      if (!ok) return fail(DNALN_E_INTERNAL, "no packed kernel for this band width");
      DN_HIP(hipGetLastError());
      if (tb) {
        if (!launch_walk_bw(sp.bw, a, 0, st)) return fail(DNALN_E_INTERNAL, "no walk kernel for this band width");
        DN_HIP(hipGetLastError());
      }
    } else {
      SwGenericArgs g{};
      g.sp = sp;
      g.cigar_cap = cap;
      g.traceback = tb;
      g.tasks = tk;
      g.order = nullptr;
      g.n = (uint32_t)n;
      g.res = res;
      g.cigar = cig;
      g.rows = reinterpret_cast<int32_t*>(at(L.rows));
      g.dir = tb ? at(L.dir) : nullptr;
      g.dir_stride = tb ? (uint64_t)std::max(qmax, 1) * W : 0;
      hipLaunchKernelGGL(k_sw_generic, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, g);
      DN_HIP(hipGetLastError());
    }
    StaticCtl* ctl = reinterpret_cast<StaticCtl*>(at(L.ctl));
    uint64_t* off = reinterpret_cast<uint64_t*>(at(L.off));
    dnaln_result* dout = reinterpret_cast<dnaln_result*>(at(L.out));
    uint32_t* outcig = reinterpret_cast<uint32_t*>(at(L.outcig));
    hipLaunchKernelGGL(k_static_scan, dim3(1), dim3(1024), 0, st, res, (uint32_t)n, off, ctl);
    DN_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_static_out, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, res, cig, tb ? cap : 0, off,
                       (uint32_t)n, dout, outcig);
    DN_HIP(hipGetLastError());
    StaticCtl h{};
    DN_HIP(hipMemcpyAsync(&h, ctl, sizeof h, hipMemcpyDeviceToHost, st));
    DN_HIP(hipStreamSynchronize(st));
    if (h.err) return fail(DNALN_E_INTERNAL, "traceback failed on a task");
    if (h.total && (!cigar_pool || h.total > cigar_cap)) return fail(DNALN_E_CAPACITY, "cigar_pool too small");
    DN_HIP(hipMemcpyAsync(out, dout, n * sizeof(dnaln_result), hipMemcpyDeviceToHost, st));
    if (h.total) DN_HIP(hipMemcpyAsync(cigar_pool, outcig, h.total * 4, hipMemcpyDeviceToHost, st));
    DN_HIP(hipStreamSynchronize(st));
    return DNALN_OK;
  });
}

void dnaln_hip_default_batch_config(dnaln_batch_config* cfg) {
  std::memset(cfg, 0, sizeof *cfg);
  cfg->mode = DNALN_MODE_PLAIN;
  cfg->chunk_reads = 65536;
  cfg->max_reads = 0;
  cfg->max_bases = 0;
  cfg->anchors_per_read = 0;
  cfg->tune = 0;
  cfg->len_split[0] = cfg->len_split[1] = 0;
  cfg->dir_budget = 0;
}

int dnaln_hip_batch_create(dnaln_hip_ctx* c, const dnaln_batch_config* cfg, dnaln_hip_batch** out) {
  return guarded([&]() -> int {
    if (!c || !out) return fail(DNALN_E_INVALID, "null argument");
    std::unique_ptr<dnaln_hip_batch> b(new dnaln_hip_batch);
    b->ctx = c;
    if (cfg) b->cfg = *cfg;
    else dnaln_hip_default_batch_config(&b->cfg);
    if (b->cfg.chunk_reads == 0) b->cfg.chunk_reads = 65536;
    DN_HIP(hipSetDevice(c->device));
    b->sw.dir_budget = b->cfg.dir_budget ? std::max<uint64_t>(b->cfg.dir_budget, 1 << 20) : default_dir_budget();
    if (b->cfg.walk_kind < 0 || b->cfg.walk_kind > 1) return fail(DNALN_E_INVALID, "walk_kind must be 0 or 1");
    if (b->cfg.sw_lanes < 0 || (b->cfg.sw_lanes > 1 && !split_available(c->p.sc.bw, b->cfg.sw_lanes)))
      return fail(DNALN_E_INVALID, "sw_lanes = " + std::to_string(b->cfg.sw_lanes) + ": no such split for bw " +
                                       std::to_string(c->p.sc.bw));
    if (b->cfg.graph < 0 || b->cfg.graph > 2) return fail(DNALN_E_INVALID, "graph must be 0, 1 or 2");
#if !defined(DN_GRAPH_CAPTURE)
    // The device-count sequence (graph = 2) and its captured form (graph = 1) faulted the GPU on
    // their first run, and the cause has not been found: on the CPU emulation every case passes,
    // bounds-checked, also on stale device memory and with the device's grid sizes. They stay
    // switched off on the device until it is. Builds that define DN_GRAPH_CAPTURE (the CPU
    // emulation) keep them.
    if (b->cfg.graph != 0)
      return fail(DNALN_E_INVALID, "graph = " + std::to_string(b->cfg.graph) +
                                       ": the device-count forms of the batched pipeline are switched off");
#endif
    b->sw.walk_kind = b->cfg.walk_kind;
    b->sw.lanes = b->cfg.sw_lanes ? b->cfg.sw_lanes : default_sw_lanes(c->p.sc.bw, b->cfg.mode);
    DN_HIP(hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking));
    DN_HIP(hipEventCreate(&b->done));
    DN_TRY(b->err.reserve(64));
    if (b->cfg.max_reads) {
      DN_TRY(b->offs.reserve(b->cfg.max_reads * 8));
      DN_TRY(b->lens.reserve(b->cfg.max_reads * 4));
      DN_TRY(b->fnv.reserve(b->cfg.max_reads * 4));
      DN_TRY(b->rec.reserve(b->cfg.max_reads * sizeof(dnaln_record)));
    }
    if (b->cfg.max_bases) DN_TRY(b->codes.reserve(b->cfg.max_bases));
    *out = b.release();
    return DNALN_OK;
  });
}

void dnaln_hip_batch_destroy(dnaln_hip_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  if (b->st) (void)hipStreamSynchronize(b->st);
  for (auto& sp : b->spans) {
    (void)hipEventDestroy(sp.a);
    (void)hipEventDestroy(sp.b);
  }
  if (b->done) (void)hipEventDestroy(b->done);
  for (auto& ge : b->graphs) (void)hipGraphExecDestroy(ge.exec);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
}

// Resolves requested bucket limits (0: default, -1: bucket unused) and counts the reads of
// each bucket.
static void set_splits(dnaln_hip_batch* b, int want0, int want1) {
  int def[2];
  default_splits(b->ctx->k, b->ctx->w, def);
  const int want[2] = {want0, want1};
  for (int x = 0; x < 2; ++x)
    b->split[x] = want[x] < 0 ? 0 : (want[x] > 0 && fused_shape(b->ctx->k, b->ctx->w, want[x]).cap ? want[x] : def[x]);
  b->split[1] = std::max(b->split[1], b->split[0]);
  b->n_short = b->n_mid = 0;
  for (const int32_t l : b->lens_h) {
    b->n_short += l <= b->split[0];
    b->n_mid += l <= b->split[1];
  }
}

int dnaln_hip_batch_upload(dnaln_hip_batch* b, const uint8_t* codes, const uint64_t* offs, const int32_t* lens,
                           const uint32_t* fnv, uint64_t n) {
  return guarded([&]() -> int {
    if (!b || (n && (!codes || !offs || !lens || !fnv))) return fail(DNALN_E_INVALID, "null argument");
    if (n >= (1ull << 31)) return fail(DNALN_E_INVALID, "too many reads in one batch");
    DN_HIP(hipSetDevice(b->ctx->device));
    b->last_rc = DNALN_OK;
    uint64_t bases = 0;
    int minlen = n ? lens[0] : 0, maxlen = 0;
    for (uint64_t i = 0; i < n; ++i) {
      if (lens[i] < 0) return fail(DNALN_E_INVALID, "negative read length");
      bases = std::max(bases, offs[i] + (uint64_t)lens[i]);
      minlen = std::min(minlen, lens[i]);
      maxlen = std::max(maxlen, lens[i]);
    }
    DN_HIP(hipStreamSynchronize(b->st));
    // (greserve: a reallocated input buffer also retires the graphs that point at it)
    DN_TRY(greserve(b, b->codes, bases + 2 * kReadGuard));
    DN_TRY(greserve(b, b->offs, n * 8 + 8));
    DN_TRY(greserve(b, b->lens, n * 4 + 4));
    DN_TRY(greserve(b, b->fnv, n * 4 + 4));
    DN_TRY(greserve(b, b->rec, n * sizeof(dnaln_record) + sizeof(dnaln_record)));
    if (bases) DN_HIP(hipMemcpyAsync(b->codes.as<uint8_t>() + kReadGuard, codes, bases, hipMemcpyHostToDevice, b->st));
    if (n) {
      DN_HIP(hipMemcpyAsync(b->offs.p, offs, n * 8, hipMemcpyHostToDevice, b->st));
      DN_HIP(hipMemcpyAsync(b->lens.p, lens, n * 4, hipMemcpyHostToDevice, b->st));
      DN_HIP(hipMemcpyAsync(b->fnv.p, fnv, n * 4, hipMemcpyHostToDevice, b->st));
    }
    b->n = n;
    b->bases = bases;
    b->minlen = minlen;
    b->maxlen = maxlen;
    b->lens_h.assign(lens, lens + n);
    b->sum_len = 0;
    for (uint64_t i = 0; i < n; ++i) b->sum_len += (uint64_t)lens[i];
    if (b->cfg.mode == DNALN_MODE_BATCHED) {
      b->len_hist.assign((size_t)maxlen + 1, 0);
      for (uint64_t i = 0; i < n; ++i) ++b->len_hist[(size_t)lens[i]];
    }
    set_splits(b, b->cfg.len_split[0], b->cfg.len_split[1]);
    return DNALN_OK;
  });
}

int dnaln_hip_batch_run(dnaln_hip_batch* b) {
  return guarded([&]() -> int {
    if (!b) return fail(DNALN_E_INVALID, "null batch");
    if (b->cfg.tune && !b->tuned && b->cfg.mode == DNALN_MODE_BATCHED && b->n) DN_TRY(dnaln_hip_batch_tune(b, nullptr));
    dnaln_hip_ctx* c = b->ctx;
    DN_HIP(hipSetDevice(c->device));
    std::memset(&b->stats, 0, sizeof b->stats);
    b->pool_used = 0;
    b->n_spans = 0;
    for (auto& l : b->launches) l = 0;
    b->last_rc = DNALN_OK;
    size_t run_span = 0;
    int rc = DNALN_OK;
    b->device_form = b->cfg.mode == DNALN_MODE_BATCHED && b->cfg.graph != 0;
    if (b->device_form) {
      // The whole batch as launches whose sizes live on the device (batched_graph.hpp): nothing
      // waits on the device.
      rc = prepare_batched(b);
      if (rc == DNALN_OK) rc = b->span_begin(5, run_span);
      if (rc == DNALN_OK) rc = launch_batched(b);
    } else {
      // Plain form: chunks of chunk_reads. Batched form: the whole batch as one chunk.
      DN_HIP(hipMemsetAsync(b->err.p, 0, 4, b->st));
      rc = b->span_begin(5, run_span);
      const uint64_t chunk = b->cfg.mode == DNALN_MODE_PLAIN ? b->cfg.chunk_reads : std::max<uint64_t>(b->n, 1);
      for (uint64_t r0 = 0; rc == DNALN_OK && (r0 < b->n || (r0 == 0 && b->n == 0)); r0 += chunk) {
        const uint32_t m = (uint32_t)std::min<uint64_t>(chunk, b->n - r0);
        rc = run_chunk(b, r0, m, r0 == 0);
        if (b->n == 0) break;
      }
    }
    if (rc != DNALN_OK) {
      (void)hipStreamSynchronize(b->st);
      b->last_rc = rc;
      b->pending = false;
      return rc;
    }
    DN_TRY(b->span_end(run_span));
    DN_HIP(hipEventRecord(b->done, b->st));
    b->stats.reads = b->n;
    b->pending = true;
    return DNALN_OK;
  });
}

int dnaln_hip_batch_tune(dnaln_hip_batch* b, dnaln_tune_report* rep) {
  return guarded([&]() -> int {
    if (!b) return fail(DNALN_E_INVALID, "null batch");
    if (b->cfg.mode != DNALN_MODE_BATCHED) return fail(DNALN_E_INVALID, "the autotuner applies to the batched form");
    b->tuned = true;
    dnaln_tune_report r{};
    // One candidate: apply, run twice (the first run grows the workspace), keep the second time.
    auto trial = [&](int s0, int s1, uint64_t budget, int walk, int lanes, bool& skipped) -> int {
      set_splits(b, s0, s1);
      b->sw.dir_budget = budget;
      b->sw.walk_kind = walk;
      b->sw.lanes = lanes;
      const int e0 = b->n_short > 0 ? b->split[0] : 0, e1 = b->n_mid > b->n_short ? b->split[1] : 0;
      skipped = false;
      for (int x = 0; x < r.n_tried; ++x)
        if (r.cand_split[x][0] == e0 && r.cand_split[x][1] == e1 && r.cand_budget[x] == budget &&
            r.cand_walk[x] == walk && r.cand_lanes[x] == lanes)
          skipped = true;
      if (skipped || r.n_tried >= 16) {
        skipped = true;
        return DNALN_OK;
      }
      double ms = 0;
      for (int rep_i = 0; rep_i < 2; ++rep_i) {
        DN_TRY(dnaln_hip_batch_run(b));
        DN_TRY(dnaln_hip_batch_poll(b, 1));
        ms = b->ms[5];
      }
      const int t = r.n_tried++;
      r.ms[t] = ms;
      r.cand_split[t][0] = e0;
      r.cand_split[t][1] = e1;
      r.cand_budget[t] = budget;
      r.cand_walk[t] = walk;
      r.cand_lanes[t] = lanes;
      return DNALN_OK;
    };
    const int bw = b->ctx->p.sc.bw;
    const int lanes0 = b->sw.lanes;
    int split = 1;  // the band width's split fill, if any
    for (const int l : {2, 4})
      if (split_available(bw, l)) split = l;
    const uint64_t big = default_dir_budget();
    std::vector<uint64_t> budgets{big};
    if (big != (1ull << 30)) budgets.push_back(1ull << 30);
    bool skipped = false;
    // Stage 1: read-length bucketing x direction budget, with the per-thread walk and the
    // configured lanes per pair.
    for (const auto& sp : {std::make_pair(0, 0), std::make_pair(0, -1), std::make_pair(-1, 0), std::make_pair(-1, -1)})
      for (const uint64_t bud : budgets) DN_TRY(trial(sp.first, sp.second, bud, 0, lanes0, skipped));
    auto best_of = [&]() {
      int bi = 0;
      for (int x = 1; x < r.n_tried; ++x)
        if (r.ms[x] < r.ms[bi]) bi = x;
      return bi;
    };
    // Stage 2: the SW fill's other lanes per pair (where the band width has a split), then the
    // other walk (which applies to the unsplit fill) on the best setting.
    int bi = best_of();
    auto s0_of = [&](int x) { return r.cand_split[x][0] > 0 ? r.cand_split[x][0] : -1; };
    auto s1_of = [&](int x) { return r.cand_split[x][1] > 0 ? r.cand_split[x][1] : -1; };
    if (split > 1) {
      DN_TRY(trial(s0_of(bi), s1_of(bi), r.cand_budget[bi], 0, lanes0 > 1 ? 1 : split, skipped));
      bi = best_of();
    }
    if (r.cand_lanes[bi] == 1) {
      DN_TRY(trial(s0_of(bi), s1_of(bi), r.cand_budget[bi], 1, 1, skipped));
      bi = best_of();
    }
    r.chosen = bi;
    r.len_split[0] = r.cand_split[bi][0];
    r.len_split[1] = r.cand_split[bi][1];
    r.dir_budget = r.cand_budget[bi];
    r.walk_kind = r.cand_walk[bi];
    r.sw_lanes = r.cand_lanes[bi];
    b->cfg.len_split[0] = r.len_split[0] > 0 ? r.len_split[0] : -1;
    b->cfg.len_split[1] = r.len_split[1] > 0 ? r.len_split[1] : -1;
    b->cfg.dir_budget = r.dir_budget;
    b->cfg.walk_kind = r.walk_kind;
    b->cfg.sw_lanes = r.sw_lanes;
    set_splits(b, b->cfg.len_split[0], b->cfg.len_split[1]);
    b->sw.dir_budget = r.dir_budget;
    b->sw.walk_kind = r.walk_kind;
    b->sw.lanes = r.sw_lanes;
    if (rep) *rep = r;
    return DNALN_OK;
  });
}

int dnaln_hip_batch_poll(dnaln_hip_batch* b, int wait) {
  return guarded([&]() -> int {
    if (!b) return fail(DNALN_E_INVALID, "null batch");
    if (b->last_rc != DNALN_OK) return b->last_rc;
    if (!b->pending) return DNALN_OK;
    uint32_t err = 0;
    for (;;) {
      if (!wait) {
        const hipError_t q = hipEventQuery(b->done);
        if (q == hipErrorNotReady) return DNALN_E_PENDING;
        if (q != hipSuccess) return fail(DNALN_E_DEVICE, hipGetErrorString(q));
      }
      DN_HIP(hipEventSynchronize(b->done));
      b->pending = false;
      for (auto& m : b->ms) m = 0;
      for (size_t i = 0; i < b->n_spans; ++i) {
        float t = 0;
        DN_HIP(hipEventElapsedTime(&t, b->spans[i].a, b->spans[i].b));
        b->ms[b->spans[i].family] += t;
      }
      if (!b->device_form) {
        int e = 0;
        DN_HIP(hipMemcpy(&e, b->err.p, 4, hipMemcpyDeviceToHost));
        err = (uint32_t)e;
        break;
      }
      DN_HIP(hipMemcpy(&b->ctl_h, b->gctl.p, sizeof(BatchCtl), hipMemcpyDeviceToHost));
      const BatchCtl& k = b->ctl_h;
      if (k.anchor_overflow) {
        // The global path needed more anchors than planned: grow the buffer and run again.
        b->anchor_cap = k.anchors + k.anchors / 4 + 65536;
        const int rc = dnaln_hip_batch_run(b);
        if (rc != DNALN_OK) return rc;
        wait = 1;
        continue;
      }
      err = k.err;
      b->stats.anchors = k.anchors;
      b->stats.global_path_reads = k.nglobal;
      b->stats.fused_retry_reads = b->gplan.retry_on ? k.novf : 0;
      b->stats.rank_tasks = k.nr;
      b->stats.tb_tasks = k.nt;
      b->pool_used = k.ncig;
      break;
    }
    if (err) {
      b->last_rc = DNALN_E_INTERNAL;
      return fail(DNALN_E_INTERNAL, err & 1 ? "CIGAR slot overflow" : "inconsistent traceback");
    }
    return DNALN_OK;
  });
}

int dnaln_hip_batch_download(dnaln_hip_batch* b, dnaln_record* out, uint32_t* cigar_pool, uint64_t cigar_cap,
                             dnaln_map_stats* stats) {
  return guarded([&]() -> int {
    if (!b) return fail(DNALN_E_INVALID, "null batch");
    DN_TRY(dnaln_hip_batch_poll(b, 1));
    if (b->pool_used > cigar_cap) return fail(DNALN_E_CAPACITY, "cigar_pool too small");
    if (b->n && !out) return fail(DNALN_E_INVALID, "null output");
    if (b->n) DN_HIP(hipMemcpy(out, b->rec.p, b->n * sizeof(dnaln_record), hipMemcpyDeviceToHost));
    if (b->pool_used) DN_HIP(hipMemcpy(cigar_pool, b->pool.p, b->pool_used * 4, hipMemcpyDeviceToHost));
    if (stats) {
      // Band cells as the reference counts them (src/mapper.cpp:139, 186): every candidate of a
      // read with two or more in the rank pass, and the chosen one in the traceback pass.
      dnaln_map_stats s = b->stats;
      const uint64_t W = 2 * (uint64_t)b->ctx->p.sc.bw + 1;
      uint64_t mapped = 0, cand = 0, cells = 0;
      for (uint64_t i = 0; i < b->n; ++i) {
        const uint64_t nc = (uint64_t)std::max(out[i].n_cand, 0), len = (uint64_t)b->lens_h[i];
        mapped += out[i].contig >= 0;
        cand += nc;
        cells += ((nc >= 2 ? nc : 0) + (nc >= 1 ? 1 : 0)) * len * W;
      }
      s.candidates = cand;
      s.mapped = mapped;
      s.cells = cells;
      s.ms_seed = b->ms[0];
      s.ms_chain = b->ms[1];
      s.ms_rank = b->ms[2];
      s.ms_traceback = b->ms[3] + b->ms[4];
      s.ms_total = b->ms[5];
      *stats = s;
    }
    return DNALN_OK;
  });
}

int dnaln_hip_batch_candidates(dnaln_hip_batch* b, int32_t* n_cand, dnaln_candidate* cands, uint64_t per_read) {
  return guarded([&]() -> int {
    static_assert(sizeof(Cand) == sizeof(dnaln_candidate), "candidate layout");
    if (!b || (b->n && (!n_cand || !cands))) return fail(DNALN_E_INVALID, "null argument");
    const uint64_t mc = (uint64_t)std::max(b->ctx->p.max_cand, 0);
    if (per_read < mc) return fail(DNALN_E_CAPACITY, "per_read is below max_cand");
    if (b->cfg.mode == DNALN_MODE_PLAIN && b->n > b->cfg.chunk_reads)
      return fail(DNALN_E_INVALID, "the plain form keeps the candidates of one chunk only");
    DN_TRY(dnaln_hip_batch_poll(b, 1));
    if (b->n == 0) return DNALN_OK;
    DN_HIP(hipSetDevice(b->ctx->device));
    DN_HIP(hipMemcpy(n_cand, b->ncand.p, b->n * 4, hipMemcpyDeviceToHost));
    std::vector<Cand> h(b->n * std::max<uint64_t>(mc, 1));
    if (mc) DN_HIP(hipMemcpy(h.data(), b->cands.p, b->n * mc * sizeof(Cand), hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < b->n; ++i)
      for (uint64_t c = 0; c < per_read; ++c) {
        dnaln_candidate& o = cands[i * per_read + c];
        if (c < mc && (int64_t)c < n_cand[i]) std::memcpy(&o, &h[i * mc + c], sizeof o);
        else std::memset(&o, 0, sizeof o);
      }
    return DNALN_OK;
  });
}

uint64_t dnaln_hip_batch_device_bytes(dnaln_hip_ctx* c, const dnaln_batch_config* cfg) {
  if (!c || !cfg) return 0;
  const uint64_t n = cfg->max_reads, bases = cfg->max_bases;
  const uint64_t per_read = 8 + 4 + 4 + sizeof(dnaln_record) + 4 * 12 + (uint64_t)std::max(c->p.max_cand, 1) *
                            (sizeof(Cand) + sizeof(SwTask) + sizeof(SwResult));
  return bases + n * per_read + n * (uint64_t)cfg->anchors_per_read * 48;
}

int dnaln_hip_batch_timings(dnaln_hip_batch* b, double* ms, uint64_t* launches) {
  return guarded([&]() -> int {
    static_assert(kFamilies == DNALN_TIMING_FAMILIES, "timing families");
    if (!b || !ms) return fail(DNALN_E_INVALID, "null argument");
    for (int i = 0; i < kFamilies; ++i) ms[i] = b->ms[i];
    if (launches)
      for (int i = 0; i < kFamilies; ++i) launches[i] = b->launches[i];
    return DNALN_OK;
  });
}

int dnaln_hip_batch_reserve(dnaln_hip_batch* b, uint64_t n, uint64_t bases, int32_t maxlen,
                            uint32_t anchors_per_read) {
  return guarded([&]() -> int {
    if (!b || maxlen < 0) return fail(DNALN_E_INVALID, "bad argument");
    dnaln_hip_ctx* c = b->ctx;
    DN_HIP(hipSetDevice(c->device));
    DN_HIP(hipStreamSynchronize(b->st));
    const uint64_t mc = (uint64_t)std::max(c->p.max_cand, 1);
    const uint64_t cap = 2 * (uint64_t)maxlen + 4;
    const uint64_t na = n * anchors_per_read;
    const uint64_t tasks = n * mc;
    // inputs
    DN_TRY(b->codes.reserve(bases + 2 * kReadGuard, b->st));
    DN_TRY(b->offs.reserve(n * 8 + 8, b->st));
    DN_TRY(b->lens.reserve(n * 4 + 4, b->st));
    DN_TRY(b->fnv.reserve(n * 4 + 4, b->st));
    if (b->cfg.mode == DNALN_MODE_BATCHED && b->cfg.graph != kFormHostSync) {
      // Plan the device-count sequence for the largest such batch: n reads of every length up to
      // maxlen, all (for the direction bound) of length maxlen.
      const uint64_t n0 = b->n, bases0 = b->bases, sum0 = b->sum_len;
      const int min0 = b->minlen, max0 = b->maxlen;
      std::vector<uint32_t> hist0;
      hist0.swap(b->len_hist);
      b->n = n;
      b->bases = bases;
      b->sum_len = std::max<uint64_t>(bases, n * (uint64_t)maxlen);
      b->minlen = 0;
      b->maxlen = maxlen;
      b->len_hist.assign((size_t)maxlen + 1, 0);
      b->len_hist[(size_t)maxlen] = (uint32_t)n;
      b->anchor_cap = std::max<uint64_t>(b->anchor_cap, na + 65536);
      GraphPlan pl;
      const int rc = plan_batched(b, pl);
      b->n = n0;
      b->bases = bases0;
      b->sum_len = sum0;
      b->minlen = min0;
      b->maxlen = max0;
      b->len_hist.swap(hist0);
      DN_TRY(rc);
      DN_HIP(hipStreamSynchronize(b->st));
      return DNALN_OK;
    }
    DN_TRY(b->rec.reserve(n * sizeof(dnaln_record) + sizeof(dnaln_record)));
    DN_TRY(b->pool.reserve(n * (cap + 2) * 4 + 4));
    // seeds and chains
    DN_TRY(b->cands.reserve(n * mc * sizeof(Cand)));
    DN_TRY(b->ncand.reserve(n * 4 + 4));
    DN_TRY(b->acnt.reserve(n * 4 + 4));
    DN_TRY(b->aoff.reserve(n * 8 + 8));
    DN_TRY(b->keys.reserve(na * 8 + 8));
    DN_TRY(b->f.reserve(na * 4 + 4));
    DN_TRY(b->prev.reserve(na * 4 + 4));
    DN_TRY(b->order.reserve(na * 8 + 8));
    DN_TRY(b->chains.reserve(na * sizeof(ChainRec) + sizeof(ChainRec)));
    DN_TRY(b->hits.reserve(bases * 12 + 16));
    DN_TRY(b->nhits.reserve(n * 4 + 4));
    // rank and traceback passes
    DN_TRY(b->rcnt.reserve(n * 4 + 4));
    DN_TRY(b->roff.reserve(n * 8 + 8));
    DN_TRY(b->rtasks.reserve(tasks * sizeof(SwTask) + sizeof(SwTask)));
    DN_TRY(b->rres.reserve(tasks * sizeof(SwResult) + sizeof(SwResult)));
    DN_TRY(b->chosen.reserve(n * 4 + 4));
    DN_TRY(b->second.reserve(n * 4 + 4));
    DN_TRY(b->tbflag.reserve(n * 4 + 4));
    DN_TRY(b->tboff.reserve(n * 8 + 8));
    DN_TRY(b->ttasks.reserve(n * sizeof(SwTask) + sizeof(SwTask)));
    DN_TRY(b->tres.reserve(n * sizeof(SwResult) + sizeof(SwResult)));
    DN_TRY(b->tcig.reserve(n * cap * 4 + 4));
    DN_TRY(b->ciglen.reserve(n * 4 + 4));
    DN_TRY(b->cigoff.reserve(n * 8 + 8));
    // directions of one launch (the plain form runs its tasks in read order, fixed slots)
    const int W = 2 * c->p.sc.bw + 1;
    uint64_t slot = fill_rows(std::max(maxlen, 1)) * ((W + 7) / 8) * 512;
    for (const int lanes : {2, 4})  // the split fill's slots, which the autotuner may choose
      if (split_available(c->p.sc.bw, lanes))
        slot = std::max<uint64_t>(slot, slot_rows(std::max(maxlen, 1), lanes) * slot_row_words(W, lanes) * 8);
    const uint64_t dir = std::min<uint64_t>(b->sw.dir_budget, ((n + 127) / 128 + 1) * slot);
    DN_TRY(b->sw.dir.reserve(dir));
    // scan temporaries for the largest counts
    size_t t1 = 0;
    DN_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, t1, (const uint32_t*)nullptr, (uint64_t*)nullptr, (uint32_t)tasks,
                                            b->st));
    DN_TRY(b->scan_tmp.reserve(t1));
    if (b->cfg.mode == DNALN_MODE_BATCHED) {
      // The host-synchronized batched form: read length order, the fused kernel's lists, and the
      // SW tasks' length sort and per-wave slot offsets.
      DN_TRY(b->lids.reserve(n * 4));
      DN_TRY(b->lkeys.reserve(n * 4));
      DN_TRY(b->lorder.reserve(n * 4));
      DN_TRY(b->ovf.reserve(n * 4 + 8));
      DN_TRY(b->ovf2.reserve(n * 4 + 8));
      FusedShape fs;
      fs.maxmin = 64;
      DN_TRY(b->fmins.reserve(fused_mins_words(fs, n) * 8));
      DN_TRY(b->sw.keys.reserve(tasks * 4));
      DN_TRY(b->sw.keys2.reserve(tasks * 4));
      DN_TRY(b->sw.ids.reserve(tasks * 4));
      DN_TRY(b->sw.ids2.reserve(tasks * 4));
      const uint64_t waves = (tasks + 127) / 128;
      DN_TRY(b->sw.woff.reserve((waves + 2) * 16));
      size_t t2 = 0, t3 = 0;
      DN_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, t2, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                                          (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                                          (uint32_t)tasks, 0, 16, b->st));
      DN_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, t3, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                              (uint32_t)(waves + 1), b->st));
      DN_TRY(b->sw.tmp.reserve(std::max(t2, t3)));
      DN_TRY(b->ltmp.reserve(t2));
    }
    DN_HIP(hipStreamSynchronize(b->st));
    return DNALN_OK;
  });
}

int dnaln_hip_device_memory(dnaln_hip_ctx* c, uint64_t* free_bytes, uint64_t* total_bytes) {
  return guarded([&]() -> int {
    if (!c || !free_bytes || !total_bytes) return fail(DNALN_E_INVALID, "null argument");
    DN_HIP(hipSetDevice(c->device));
    size_t f = 0, t = 0;
    DN_HIP(hipMemGetInfo(&f, &t));
    *free_bytes = f;
    *total_bytes = t;
    return DNALN_OK;
  });
}

int dnaln_hip_host_alloc(uint64_t bytes, void** out) {
  return guarded([&]() -> int {
    if (!out) return fail(DNALN_E_INVALID, "null argument");
    *out = nullptr;
    if (hipHostMalloc(out, std::max<uint64_t>(bytes, 1), hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      *out = nullptr;
      return fail(DNALN_E_NOMEM, "hipHostMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    return DNALN_OK;
  });
}

void dnaln_hip_host_free(void* p) {
  if (p) (void)hipHostFree(p);
}

int dnaln_hip_context_index(dnaln_hip_ctx* c, uint64_t* n_buckets, uint64_t* n_pos, double* build_ms) {
  return guarded([&]() -> int {
    if (!c) return fail(DNALN_E_INVALID, "null argument");
    if (n_buckets) *n_buckets = c->ix.n_buckets;
    if (n_pos) *n_pos = c->ix.n_pos;
    if (build_ms) *build_ms = c->ix.ms;
    return DNALN_OK;
  });
}

int dnaln_hip_index_build(const uint8_t* ref_codes, uint64_t ref_len, const dnaln_contig* contigs, int32_t n_contigs,
                          int32_t k, int32_t w, int32_t device, dnaln_hip_index** out) {
  return guarded([&]() -> int {
    if (!out || !contigs || n_contigs < 1 || (!ref_codes && ref_len)) return fail(DNALN_E_INVALID, "null argument");
    *out = nullptr;
    if (ref_len >= (1ull << 31)) return fail(DNALN_E_INVALID, "reference is too large (limit: 2^31 bp)");
    std::vector<uint64_t> coff(n_contigs);
    std::vector<uint32_t> clen(n_contigs);
    for (int i = 0; i < n_contigs; ++i) {
      if (contigs[i].offset + contigs[i].len > ref_len) return fail(DNALN_E_INVALID, "contig outside the reference");
      coff[i] = contigs[i].offset;
      clen[i] = (uint32_t)contigs[i].len;
    }
    std::unique_ptr<dnaln_hip_index> ix(new dnaln_hip_index);
    ix->device = device;
    DN_HIP(hipSetDevice(device));
    hipStream_t st = nullptr;
    DN_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    int rc;
    {
      DevBuf ref;
      uint8_t* base = nullptr;
      rc = upload_reference(ref, ref_codes, ref_len, st, &base);
      if (rc == DNALN_OK)
        rc = build_index_dev(base, coff.data(), clen.data(), n_contigs, k, w, st, ix->table, ix->pos, ix->info);
      (void)hipStreamSynchronize(st);
    }
    (void)hipStreamDestroy(st);
    if (rc != DNALN_OK) return rc;
    *out = ix.release();
    return DNALN_OK;
  });
}

int dnaln_hip_index_info(const dnaln_hip_index* ix, uint64_t* n_buckets, uint64_t* n_pos, uint64_t* n_keys,
                         uint32_t* max_count, double* build_ms) {
  return guarded([&]() -> int {
    if (!ix) return fail(DNALN_E_INVALID, "null argument");
    if (n_buckets) *n_buckets = ix->info.n_buckets;
    if (n_pos) *n_pos = ix->info.n_pos;
    if (n_keys) *n_keys = ix->info.n_keys;
    if (max_count) *max_count = ix->info.max_count;
    if (build_ms) *build_ms = ix->info.ms;
    return DNALN_OK;
  });
}

int dnaln_hip_index_download(const dnaln_hip_index* ix, dnaln_bucket* table, uint32_t* pos) {
  return guarded([&]() -> int {
    if (!ix || !table || (!pos && ix->info.n_pos)) return fail(DNALN_E_INVALID, "null argument");
    DN_HIP(hipSetDevice(ix->device));
    DN_HIP(hipMemcpy(table, ix->table.p, ix->info.n_buckets * sizeof(dnaln_bucket), hipMemcpyDeviceToHost));
    if (ix->info.n_pos) DN_HIP(hipMemcpy(pos, ix->pos.p, ix->info.n_pos * 4, hipMemcpyDeviceToHost));
    return DNALN_OK;
  });
}

void dnaln_hip_index_free(dnaln_hip_index* ix) {
  if (!ix) return;
  (void)hipSetDevice(ix->device);
  delete ix;
}

int dnaln_hip_context_params(dnaln_hip_ctx* c, dnaln_map_params* out) {
  return guarded([&]() -> int {
    if (!c || !out) return fail(DNALN_E_INVALID, "null argument");
    *out = c->p;
    return DNALN_OK;
  });
}

void dnaln_hip_alloc_counters(uint64_t* events, uint64_t* bytes) {
  if (events) *events = g_alloc_events.load(std::memory_order_relaxed);
  if (bytes) *bytes = g_alloc_bytes.load(std::memory_order_relaxed);
}

// ------------------------------------------------------------------ packed int16 VALU peak
// Sixteen independent chains of v_pk_add_u16 + v_pk_max_i16 per lane: 64 int16 operations per
// lane per iteration and no memory access in the loop. The values depend on the thread and the
// seed, and a thread stores only when its result equals a value the host chose, so every
// operation stays live while the kernel (almost surely) writes nothing.
namespace {
constexpr int kPeakChains = 16, kPeakThreads = 256, kPeakWavesPerSimd = 8;

__global__ __launch_bounds__(kPeakThreads) void k_peak_pk16(int iters, uint32_t seed, uint32_t magic,
                                                             uint32_t* sink) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  pk16 a[kPeakChains];
#pragma unroll
  for (int j = 0; j < kPeakChains; ++j) a[j] = pk_from(t * 2654435761u + (uint32_t)j * 40503u + seed);
  const pk16 inc = pk_from(seed | 0x00010001u), lim = pk_from(seed >> 3);
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int j = 0; j < kPeakChains; ++j) a[j] = pk_max(pk_add(a[j], inc), lim);
  }
  uint32_t x = 0;
#pragma unroll
  for (int j = 0; j < kPeakChains; ++j) x ^= pk_bits(a[j]);
  if (x == magic) sink[t & 255] = x;
}
}  // namespace

int dnaln_hip_peak_int16(int32_t device, int32_t iters, double* tops, double* ms) {
  return guarded([&]() -> int {
    if (!tops || iters <= 0) return fail(DNALN_E_INVALID, "bad argument");
    DN_HIP(hipSetDevice(device));
    const unsigned blocks = (unsigned)device_cus(device) * 4 * kPeakWavesPerSimd / (kPeakThreads / 64);
    DevBuf sink;
    DN_TRY(sink.reserve(256 * 4));
    hipStream_t st = nullptr;
    DN_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = DNALN_OK;
    float t = 0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = fail(DNALN_E_DEVICE, "hipEventCreate");
    if (rc == DNALN_OK) {
      // One short launch to load the code object and raise the clocks, then the timed one.
      hipLaunchKernelGGL(k_peak_pk16, dim3(blocks), dim3(kPeakThreads), 0, st, std::max(iters / 16, 1), 12345u,
                         0xffffffffu, sink.as<uint32_t>());
      (void)hipEventRecord(e0, st);
      hipLaunchKernelGGL(k_peak_pk16, dim3(blocks), dim3(kPeakThreads), 0, st, iters, 0x2468ace1u, 0xffffffffu,
                         sink.as<uint32_t>());
      (void)hipEventRecord(e1, st);
      const hipError_t e = hipEventSynchronize(e1);
      if (e != hipSuccess) rc = fail(DNALN_E_DEVICE, std::string("k_peak_pk16: ") + hipGetErrorString(e));
      else if (hipEventElapsedTime(&t, e0, e1) != hipSuccess || t <= 0) rc = fail(DNALN_E_DEVICE, "event timing");
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    DN_TRY(rc);
    const double ops = (double)blocks * kPeakThreads * (double)iters * kPeakChains * 2 * 2;
    *tops = ops / (t * 1e-3) / 1e12;
    if (ms) *ms = t;
    return DNALN_OK;
  });
}

int dnaln_hip_map_batch(dnaln_hip_ctx* c, const uint8_t* read_codes, const uint64_t* offs, const int32_t* lens,
                        const uint32_t* name_fnv1a, uint64_t n, dnaln_record* out, uint32_t* cigar_pool,
                        uint64_t cigar_cap, dnaln_map_stats* stats) {
  return guarded([&]() -> int {
    if (!c) return fail(DNALN_E_INVALID, "null context");
    if (!c->batch) {
      dnaln_batch_config cfg;
      dnaln_hip_default_batch_config(&cfg);
      DN_TRY(dnaln_hip_batch_create(c, &cfg, &c->batch));
    }
    DN_TRY(dnaln_hip_batch_upload(c->batch, read_codes, offs, lens, name_fnv1a, n));
    DN_TRY(dnaln_hip_batch_run(c->batch));
    return dnaln_hip_batch_download(c->batch, out, cigar_pool, cigar_cap, stats);
  });
}

}  // extern "C"
