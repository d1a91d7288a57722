// batched_graph.hpp — the batched form of Mapper::map (src/mapper.cpp:75-223) as one HIP graph.
//
// Included once, by dnaln_hip.hip. The whole batch runs as a fixed sequence of kernels that never
// hands a count back to the host: every kernel reads its sizes (reads per bucket, declined reads,
// anchors, rank and traceback tasks, fill waves) from a control block in device memory that
// earlier kernels of the same run wrote. Grids are sized on the host for the worst case of the
// planned batch size (every read on the global path, every candidate in the rank pass), and the
// waves past the device-side count exit at once. So the sequence is captured once per plan
// (buffer addresses, shapes, grids and cuts fixed on the host) and each batch is one
// hipGraphLaunch; host and device meet only when the batch is polled. (A persistent loop drawing
// work from a device counter was tried first: it raised the fills' and the walk's register
// counts, e.g. the score-only W = 33 fill from 155 to 175 VGPRs, 3 to 2 waves per SIMD.)
//
// Stages, in stream order (every stage is a kernel; the graph holds no memset or copy node):
//   reset        zero the control block's counters (the graph's first node)
//   length sort  reads by length, longest first: histogram, scan, scatter (when lengths differ)
//   fused        seed + chain + candidates, one read per lane, anchors in LDS: short and mid
//                buckets, then a 64-anchor pass over what a 32-anchor shape declined
//   global path  long and declined reads: count anchors, scan, emit, one wave per read chains
//   rank pass    tasks of reads with >= 2 candidates (in length order, so every fill wave holds
//                nearly equal lengths), score-only fill
//   select       src/mapper.cpp:142-174
//   traceback    tasks of the chosen candidates in length order, per-wave direction slots sized
//                by each wave's longest query (a scan on the device), fill + walk per cut
//   records      records, MAPQ, CIGAR offsets (scan), CIGAR gather
// Anchors of the global path live in a buffer of planned capacity; a run that needs more sets
// a flag, and the host grows the buffer and runs the batch again when it polls.
#pragma once

namespace {

// ------------------------------------------------------------------------------ control block
constexpr int kMaxCuts = 256;  // traceback fill launches per run (direction budget)
// dnaln_batch_config::graph of the batched form: 0 the host-synchronized form (run_chunk), 1 this
// sequence captured as a graph, 2 this sequence launched one by one.
constexpr int kFormHostSync = 0, kFormGraph = 1, kFormDirect = 2;

// One run's sizes and counters in device memory. `n` is written before the graph is launched
// (k_begin); every other field starts at zero (k_reset) or is written by an earlier stage.
struct BatchCtl {
  uint32_t n;            // reads in the batch
  uint32_t seg_mid;      // length order: [0, seg_mid) long, [seg_mid, seg_short) mid,
  uint32_t seg_short;    //   [seg_short, n) short
  uint32_t n_short, n_mid;  // reads with length <= split[0] and <= split[1]
  uint32_t novf;         // reads the first fused pass declined
  uint32_t nglobal;      // reads on the global path
  uint32_t retry;        // reads the second fused pass took
  uint32_t nr, nt;       // rank and traceback tasks
  uint32_t tb_waves;     // fill waves of the traceback pass
  uint32_t err;          // bit 0: CIGAR slot overflow, bit 1: inconsistent traceback
  uint32_t anchor_overflow;
  uint32_t reserved;
  uint64_t anchors;      // anchors of the global path
  uint64_t ncig;         // CIGAR words of the batch
};

__global__ void k_begin(BatchCtl* ctl, uint32_t n) { ctl->n = n; }

// Zeroes every counter and the length histogram (bins words); `bucket` >= 0 (reads not sorted,
// all of one length) puts every read in that bucket (0 short, 1 mid, 2 long).
__global__ void k_reset(BatchCtl* ctl, int bucket, uint32_t* hist, uint32_t bins) {
  const uint32_t n = ctl->n;
  uint32_t* w = reinterpret_cast<uint32_t*>(ctl);
  constexpr uint32_t words = sizeof(BatchCtl) / 4;
  for (uint32_t i = 1 + threadIdx.x; i < words; i += blockDim.x) w[i] = 0;
  for (uint32_t i = threadIdx.x; i < bins; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  if (threadIdx.x == 0 && bucket >= 0) {
    ctl->seg_mid = bucket == 2 ? n : 0;
    ctl->seg_short = bucket >= 1 ? n : 0;
    ctl->n_short = bucket == 0 ? n : 0;
    ctl->n_mid = bucket <= 1 ? n : 0;
  }
}

// ------------------------------------------------------------------------------ device scans
// Exclusive scan of f(0 .. n-1), n = *n_ptr, over kScanBlocks segments: k_scan_part sums each
// segment, k_scan_out adds the sums of the segments before it and scans the segment in tiles.
// (The CPU emulation, where each block costs fiber set-up, scans in fewer segments.)
#if defined(DN_EMU) && !defined(DN_EMU_GPU_SHAPES)
constexpr int kScanBlocks = 16, kScanThreads = 256;
#else
constexpr int kScanBlocks = 1024, kScanThreads = 256;
#endif

__device__ inline uint32_t scan_seg(uint32_t n) {
  const uint32_t s = (n + kScanBlocks - 1) / kScanBlocks;
  return (s + kScanThreads - 1) / kScanThreads * kScanThreads;
}

// Inclusive scan of v over the block (kScanThreads threads); *tile_total = the block's sum.
__device__ inline uint64_t block_incl_scan(uint64_t v, uint64_t* lds, uint64_t* tile_total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t x = __shfl_up(v, d, 64);
    if (lane >= d) v += x;
  }
  if (lane == 63) lds[wv] = v;
  __syncthreads();
  uint64_t before = 0, all = 0;
#pragma unroll
  for (int x = 0; x < kScanThreads / 64; ++x) {
    before += x < wv ? lds[x] : 0;
    all += lds[x];
  }
  __syncthreads();
  *tile_total = all;
  return v + before;
}

template <class F>
__global__ __launch_bounds__(kScanThreads) void k_scan_part(F f, const uint32_t* n_ptr, uint64_t* part) {
  __shared__ uint64_t lds[kScanThreads / 64];
  const uint32_t n = *n_ptr, seg = scan_seg(n);
  const uint64_t lo = (uint64_t)blockIdx.x * seg, hi = min((uint64_t)n, lo + seg);
  uint64_t s = 0;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += kScanThreads) s += f((uint32_t)i);
  uint64_t tot = 0;
  (void)block_incl_scan(s, lds, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// out[i] = f(0) + ... + f(i-1); the total goes to *total32 and *total64 (either may be null).
template <class F>
__global__ __launch_bounds__(kScanThreads) void k_scan_out(F f, const uint32_t* n_ptr, const uint64_t* part,
                                                           uint64_t* out, uint32_t* total32, uint64_t* total64) {
  __shared__ uint64_t lds[kScanThreads / 64];
  const uint32_t n = *n_ptr, seg = scan_seg(n);
  uint64_t base = 0, all = 0;
  for (uint32_t x = threadIdx.x; x < (uint32_t)kScanBlocks; x += kScanThreads) {
    const uint64_t p = part[x];
    base += x < blockIdx.x ? p : 0;
    all += p;
  }
  uint64_t t0 = 0, t1 = 0;
  (void)block_incl_scan(base, lds, &t0);
  (void)block_incl_scan(all, lds, &t1);
  base = t0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (total32) *total32 = (uint32_t)t1;
    if (total64) *total64 = t1;
  }
  const uint64_t lo = (uint64_t)blockIdx.x * seg, hi = min((uint64_t)n, lo + seg);
  for (uint64_t i0 = lo; i0 < hi; i0 += kScanThreads) {
    const uint64_t i = i0 + threadIdx.x;
    const uint64_t v = i < hi ? f((uint32_t)i) : 0;
    uint64_t tile = 0;
    const uint64_t incl = block_incl_scan(v, lds, &tile);
    if (i < hi) out[i] = base + incl - v;
    base += tile;
  }
}

template <class F>
void enqueue_scan(F f, const uint32_t* n_ptr, uint64_t* part, uint64_t* out, uint32_t* total32, uint64_t* total64,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_scan_part<F>, dim3(kScanBlocks), dim3(kScanThreads), 0, st, f, n_ptr, part);
  hipLaunchKernelGGL(k_scan_out<F>, dim3(kScanBlocks), dim3(kScanThreads), 0, st, f, n_ptr, part, out, total32,
                     total64);
}

// Read id at length-order position p (identity when the reads are not sorted).
__device__ inline uint32_t rid(const uint32_t* order, uint32_t p) { return order ? order[p] : p; }

struct ScanAnchors {
  const uint32_t* cnt;
  __device__ uint64_t operator()(uint32_t i) const { return cnt[i]; }
};
struct ScanRank {  // candidates of reads that reach the rank pass, in length order
  const uint32_t* order;
  const int32_t* n_cand;
  __device__ uint64_t operator()(uint32_t p) const {
    const int nc = n_cand[rid(order, p)];
    return nc >= 2 ? (uint64_t)nc : 0;
  }
};
struct ScanChosen {  // reads with a chosen candidate, in length order
  const uint32_t* order;
  const int32_t* chosen;
  __device__ uint64_t operator()(uint32_t p) const { return chosen[rid(order, p)] >= 0 ? 1 : 0; }
};
struct ScanWaveWords {  // direction words of traceback fill wave w: rows of its longest task
  const SwTask* tasks;
  const uint32_t* n_tasks;
  uint64_t row_words;
  __device__ uint64_t operator()(uint32_t w) const {
    const uint32_t n = *n_tasks, lo = w * 128u, hi = min(n, lo + 128u);
    int q = 0;
    for (uint32_t k = lo; k < hi; ++k) q = max(q, tasks[k].qlen);
    return fill_rows(q) * row_words;
  }
};
struct ScanCigar {
  const uint32_t* len;
  __device__ uint64_t operator()(uint32_t r) const { return len[r]; }
};

// ------------------------------------------------------------------------------ length sort
// Reads ordered longest first by a counting sort over length bins (bin = min(len, bins - 1)),
// and the bucket counts. Order within a bin is arbitrary: each read's records depend on the
// read alone.
constexpr int kHistThreads = 1024;
constexpr uint32_t kMaxLenBins = 8192;  // 32 KiB of LDS histogram per block

__global__ __launch_bounds__(kHistThreads) void k_len_hist(const int32_t* lens, BatchCtl* ctl, uint32_t bins,
                                                           uint32_t* hist, int split0, int split1) {
  DN_DYN_LDS(uint32_t, h);
  __shared__ uint32_t cs, cm;
  const uint32_t n = ctl->n;
  for (uint32_t b = threadIdx.x; b < bins; b += blockDim.x) h[b] = 0;
  if (threadIdx.x == 0) cs = cm = 0;
  __syncthreads();
  uint32_t my_s = 0, my_m = 0;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const int32_t l = lens[r];
    atomicAdd(&h[min((uint32_t)l, bins - 1)], 1u);
    my_s += l <= split0;
    my_m += l <= split1;
  }
  atomicAdd(&cs, my_s);
  atomicAdd(&cm, my_m);
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < bins; b += blockDim.x)
    if (h[b]) atomicAdd(&hist[b], h[b]);
  if (threadIdx.x == 0) {
    atomicAdd(&ctl->n_short, cs);
    atomicAdd(&ctl->n_mid, cm);
  }
}

// Bin starts in longest-first order (hist becomes the scatter cursor); the bucket segments.
__global__ __launch_bounds__(kHistThreads) void k_len_starts(BatchCtl* ctl, uint32_t bins, uint32_t* hist) {
  __shared__ uint32_t part[kHistThreads];
  const uint32_t per = (bins + kHistThreads - 1) / kHistThreads;
  // Thread t owns bins counted from the top: top bin index bins-1-(t*per) down.
  uint32_t s = 0;
  for (uint32_t x = 0; x < per; ++x) {
    const uint32_t k = threadIdx.x * per + x;
    if (k < bins) s += hist[bins - 1 - k];
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t d = 1; d < kHistThreads; d <<= 1) {
    const uint32_t v = threadIdx.x >= d ? part[threadIdx.x - d] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = part[threadIdx.x] - s;
  for (uint32_t x = 0; x < per; ++x) {
    const uint32_t k = threadIdx.x * per + x;
    if (k < bins) {
      const uint32_t b = bins - 1 - k, c = hist[b];
      hist[b] = run;
      run += c;
    }
  }
  if (threadIdx.x == 0) {
    const uint32_t n = ctl->n;
    ctl->seg_mid = n - ctl->n_mid;
    ctl->seg_short = n - ctl->n_short;
  }
}

__global__ void k_len_scatter(const int32_t* lens, const BatchCtl* ctl, uint32_t bins, uint32_t* cursor,
                              uint32_t* order) {
  const uint32_t n = ctl->n;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
    order[atomicAdd(&cursor[min((uint32_t)lens[r], bins - 1)], 1u)] = r;
}

// ------------------------------------------------------------------------------ seed + chain
// The fused kernel of the batched form over a range that only the device knows: bucket 0 / 1 are
// positions [seg_short, n) / [seg_mid, seg_short) of the length order, bucket 2 the reads the
// first pass declined. One read per lane, 64 per block; blocks past the range exit. Declined
// reads go to `decl` (ctl->novf) or, from the last pass, to the global list (ctl->nglobal).
template <int WT>
__global__ __launch_bounds__(kFusedBlock) void k_fused_g(SeedArgs s, ChainArgs c, int cap, int maxmin, uint64_t* mins,
                                                         BatchCtl* ctl, int bucket, const uint32_t* order,
                                                         const uint32_t* in_list, uint32_t* decl, bool last) {
  DN_DYN_LDS(uint64_t, lds);
  const int t = threadIdx.x;
  uint32_t lo, hi;
  if (bucket == 0) lo = ctl->seg_short, hi = ctl->n;
  else if (bucket == 1) lo = ctl->seg_mid, hi = ctl->seg_short;
  else lo = 0, hi = ctl->novf;
  const uint32_t i = lo + blockIdx.x * kFusedBlock + t;
  if (i >= hi) return;
  uint64_t* keys = lds;
  Ring ring;
  ring.h = keys + t;
  ring.ps = reinterpret_cast<uint32_t*>(keys + (size_t)s.w * kFusedBlock) + t;
  ring.stride = kFusedBlock;
  LdsView v;
  v.keys = keys + t;
  v.stride = kFusedBlock;
  uint64_t* my_mins = mins + (size_t)blockIdx.x * maxmin * kFusedBlock + t;
  const uint32_t r = bucket == 2 ? in_list[i] : rid(order, i);
  const int nc = s.ctab ? (cap == 32 ? seed_chain_lds<8, 10, true, 32>(s, c, r, v, my_mins, kFusedBlock, maxmin, ring)
                                     : seed_chain_lds<8, 10, true, 64>(s, c, r, v, my_mins, kFusedBlock, maxmin, ring))
                        : (cap == 32 ? seed_chain_lds<8, 10, false, 32>(s, c, r, v, my_mins, kFusedBlock, maxmin, ring)
                                     : seed_chain_lds<8, 10, false, 64>(s, c, r, v, my_mins, kFusedBlock, maxmin, ring));
  if (nc < 0) {
    c.n_cand[r] = 0;
    if (last) decl[atomicAdd(&ctl->nglobal, 1u)] = r;
    else decl[atomicAdd(&ctl->novf, 1u)] = r;
  } else {
    c.n_cand[r] = nc;
    if (bucket == 2) atomicAdd(&ctl->retry, 1u);
  }
}

// Long reads (positions [0, seg_mid) of the length order) join the global list.
__global__ void k_long_to_global(const BatchCtl* ctl, const uint32_t* order, uint32_t* glist, uint32_t* nglobal) {
  const uint32_t nl = ctl->seg_mid;
  __shared__ uint32_t base;
  // One block reserves the range; the grid is one block (the copy is small and coalesced).
  if (threadIdx.x == 0) base = nl ? atomicAdd(nglobal, nl) : 0;
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < nl; p += blockDim.x) glist[base + p] = rid(order, p);
}

// Global path over glist[0, ctl->nglobal), one read per thread: anchors per read, then (after
// the scan) the anchors themselves, when they fit the buffer.
// The thread's ring of the streamed sketch: in LDS, or (large w) in global memory, one slot per
// thread of the grid.
__device__ inline Ring seed_ring(uint64_t* lds, int w, uint64_t* ring_h, uint32_t* ring_ps) {
  Ring ring;
  if (!ring_h) {
    ring.h = lds + threadIdx.x;
    ring.ps = reinterpret_cast<uint32_t*>(lds + (size_t)w * blockDim.x) + threadIdx.x;
    ring.stride = blockDim.x;
  } else {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    ring.h = ring_h + tid;
    ring.ps = ring_ps + tid;
    ring.stride = (int)(gridDim.x * blockDim.x);
  }
  return ring;
}

__global__ __launch_bounds__(kSeedBlock) void k_gseed_count(SeedArgs a, const BatchCtl* ctl, uint64_t* ring_h,
                                                            uint32_t* ring_ps) {
  DN_DYN_LDS(uint64_t, lds);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ctl->nglobal) return;
  a.anchor_count[i] = seed_count_read(a, i, seed_ring(lds, a.w, ring_h, ring_ps));
}

__global__ __launch_bounds__(kSeedBlock) void k_gseed_emit(SeedArgs a, BatchCtl* ctl, uint64_t cap, uint64_t* ring_h,
                                                           uint32_t* ring_ps) {
  DN_DYN_LDS(uint64_t, lds);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (ctl->anchors > cap) {
    if (i == 0) ctl->anchor_overflow = 1;
    return;
  }
  if (i >= ctl->nglobal) return;
  seed_emit_read(a, i, seed_ring(lds, a.w, ring_h, ring_ps));
}

// A run whose anchors overflowed the buffer is run again by the host, but every later stage of
// this run still reads n_cand of every read: the reads of the global list get none.
__device__ inline bool anchors_overflowed(const ChainArgs& a, const BatchCtl* ctl, uint64_t cap, uint32_t li) {
  if (ctl->anchors <= cap) return false;
  if (li < ctl->nglobal) a.n_cand[read_id(a.list, li)] = 0;
  return true;
}

// One thread per read of the global list (lookback or max_cand above 64, where the wave chain
// does not apply).
__global__ __launch_bounds__(kSeedBlock) void k_gchain_thread(ChainArgs a, const BatchCtl* ctl, uint64_t cap) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (anchors_overflowed(a, ctl, cap, i) || i >= ctl->nglobal) return;
  chain_read(a, i);
}

// One wave per read of the global list (chain_wave.hpp), four per block.
__global__ __launch_bounds__(256) void k_gchain(ChainArgs a, BatchCtl* ctl, uint64_t cap) {
  DN_DYN_LDS(uint64_t, lds);
  if (ctl->anchors > cap) {
    if ((threadIdx.x & 63) == 0) (void)anchors_overflowed(a, ctl, cap, blockIdx.x * 4 + (threadIdx.x >> 6));
    return;
  }
  int32_t* cost = reinterpret_cast<int32_t*>(lds);
  for (int x = threadIdx.x; x <= a.band; x += blockDim.x) cost[x] = a.cost[x];
  __syncthreads();
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const uint32_t li = blockIdx.x * 4 + wv;
  if (li >= ctl->nglobal) return;
  const uint32_t r = read_id(a.list, li);
  const int na = (int)a.anchor_count[li];
  const uint64_t off = a.anchor_off[li];
  WaveChainMem m;
  if (na <= kWaveChainLds && a.min_anchors >= 2) {
    uint8_t* base = reinterpret_cast<uint8_t*>(cost + ((a.band + 4) & ~3)) + wv * kWaveChainBytes;
    m.keys = reinterpret_cast<uint64_t*>(base);
    m.order = m.keys + kWaveChainLds;
    m.f = reinterpret_cast<int32_t*>(m.order + kWaveChainLds);
    m.prev = reinterpret_cast<uint32_t*>(m.f + kWaveChainLds);
    m.chains = m.f + 2 * kWaveChainLds;
    for (int x = lane; x < na; x += 64) m.keys[x] = a.keys[off + x];
    wave_sync();
    chain_wave_read(a, r, na, m, cost, lane);
  } else {
    m.keys = a.keys + off;
    m.order = a.order + off;
    m.f = a.f + off;
    m.prev = reinterpret_cast<uint32_t*>(a.prev + off);
    m.chains = reinterpret_cast<int32_t*>(a.chains + off);
    chain_wave_read(a, r, na, m, cost, lane);
  }
}

// ------------------------------------------------------------------------------ SW stages
// Rank tasks of the read at each length-order position, and the read's first task index.
__global__ void k_grank_tasks(ReadView rv, MapGeom g, const Cand* cands, const int32_t* n_cand, int max_cand,
                              const uint32_t* order, const uint64_t* off_pos, uint64_t* rank_off, SwTask* tasks,
                              const BatchCtl* ctl) {
  const uint32_t n = ctl->n;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const uint32_t r = rid(order, p);
    const int nc = n_cand[r];
    if (nc < 2) continue;
    const uint64_t o = off_pos[p];
    rank_off[r] = o;
    const uint64_t q = (uint64_t)(uintptr_t)(rv.codes + rv.offs[r]);
    for (int c = 0; c < nc; ++c) tasks[o + c] = make_task(g, q, rv.lens[r], cands[(uint64_t)r * max_cand + c]);
  }
}

__global__ void k_gselect(ReadView rv, const int32_t* n_cand, const uint64_t* rank_off, const SwResult* rank_res,
                          int32_t* chosen, int32_t* second, const BatchCtl* ctl) {
  const uint32_t n = ctl->n;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const int nc = n_cand[r];
    const SwResult* res = rank_res + (nc >= 2 ? rank_off[r] : 0);
    int32_t ch, s2;
    select_read(nc, [res](int c) { return res[c].score; }, rv.fnv[r], ch, s2);
    chosen[r] = ch;
    second[r] = s2;
  }
}

// Traceback task of every read with a chosen candidate, in length order; tb_idx[r] = its index.
__global__ void k_gtb_tasks(ReadView rv, MapGeom g, const Cand* cands, int max_cand, const int32_t* chosen,
                            const uint32_t* order, const uint64_t* off_pos, int64_t* tb_idx, SwTask* tasks,
                            BatchCtl* ctl) {
  const uint32_t n = ctl->n;
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid == 0) ctl->tb_waves = (ctl->nt + 127) / 128;
  for (uint32_t p = tid; p < n; p += gridDim.x * blockDim.x) {
    const uint32_t r = rid(order, p);
    const int ch = chosen[r];
    if (ch < 0) {
      tb_idx[r] = -1;
      continue;
    }
    const uint64_t o = off_pos[p];
    tb_idx[r] = (int64_t)o;
    const uint64_t q = (uint64_t)(uintptr_t)(rv.codes + rv.offs[r]);
    tasks[o] = make_task(g, q, rv.lens[r], cands[(uint64_t)r * max_cand + ch]);
  }
}

// The packed fill over tasks [first, min(first + a.n, *n_ptr)) of a launch whose grid covers the
// most such tasks; a.tasks, a.res, a.cigar and a.woff already point at the cut's first task and
// wave, so the slot of wave w is dir + woff[w] - woff[0].
template <int W, bool TB>
__global__ __launch_bounds__(kSwBlock) DN_WAVES_PER_EU(sw_min_waves(W, TB)) void k_sw_packed_g(
    SwArgs a, const uint32_t* n_ptr, uint32_t first) {
  const int lane = threadIdx.x & 63;
  const uint32_t total = *n_ptr;
  const uint32_t n = total > first ? min(a.n, total - first) : 0u;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t base = (uint64_t)wave * 128;
  if (base >= n) return;
  if (TB && a.woff) a.woff_base = a.woff[0];
  const uint32_t ka = (uint32_t)base + lane, kb = (uint32_t)base + 64 + lane;
  const bool va = ka < n, vb = kb < n;
  int ql = max(va ? a.tasks[ka].qlen : 0, vb ? a.tasks[kb].qlen : 0);
  ql = wave_max(ql);
  uint64_t* slot = TB ? dir_slot(a, wave) : nullptr;
  if (TB && W > kBestRowInRegs) {
    DN_DYN_LDS(uint16_t, bh_rows);
    uint16_t* bh = bh_rows + (size_t)(threadIdx.x >> 6) * best_row_lds_words(W) * 2 + 2 * lane;
    sw_pair<W, TB>(a, ka, kb, va, vb, ql, slot, lane, bh, 64);
  } else if (TB && flag_granules<W, 1>()) {
    DN_DYN_LDS(uint32_t, stage);
    sw_pair<W, TB>(a, ka, kb, va, vb, ql, slot, lane, nullptr, 0,
                   stage + (size_t)(threadIdx.x >> 6) * fill_lds_words(W) + lane, 64);
  } else {
    sw_pair<W, TB>(a, ka, kb, va, vb, ql, slot, lane);
  }
}

// Traceback walk, one thread per task of the cut.
template <int W>
__global__ __launch_bounds__(256) void k_walk_g(SwArgs a, const uint32_t* n_ptr, uint32_t first) {
  __shared__ uint32_t seqbuf[kWalkSeqDwords * 256];
  const uint32_t total = *n_ptr;
  const uint32_t n = total > first ? min(a.n, total - first) : 0u;
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  if (a.woff) a.woff_base = a.woff[0];
  WalkSeq seq(seqbuf + threadIdx.x, 256);
  walk_task<W>(a, k, seq);
}

// The walk through LDS windows of flag rows (k_walk_win), one block per fill wave of the cut.
template <int W>
__global__ __launch_bounds__(128) void k_walk_win_g(SwArgs a, int rows, const uint32_t* n_ptr, uint32_t first) {
  constexpr int G = (W + 7) / 8;
  constexpr bool kGran = flag_granules<W, 1>();
  DN_DYN_LDS(uint64_t, win);
  __shared__ int s_top;
  __shared__ uint32_t seqbuf[kWalkSeqDwords * 128];
  const int t = threadIdx.x, lane = t & 63, half = t >> 6;
  const uint32_t total = *n_ptr;
  const uint32_t n = total > first ? min(a.n, total - first) : 0u;
  const uint32_t wave = blockIdx.x;
  if ((uint64_t)wave * 128 >= n) return;
  if (a.woff) a.woff_base = a.woff[0];
  const uint32_t k = wave * 128 + t;
  bool active = false;
  int score = 0, i_end = 0, d_end = 0;
  SwTask tk{};
  Walk w{};
  WalkSeq seq(seqbuf + t, 128);
  uint32_t* out = nullptr;
  if (k < n) {
    SwResult& r = a.res[k];
    score = r.score;
    i_end = r.qe;
    d_end = r.te;
    if (score <= 0) {
      r.qb = r.qe = r.tb = r.te = r.nm = r.n_cigar = r.status = 0;
    } else {
      tk = a.tasks[k];
      seq.begin(tk);
      out = a.cigar + (uint64_t)k * (uint64_t)a.cigar_cap;
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
      const uint4 vv = *reinterpret_cast<const uint4*>(src + x);
      *reinterpret_cast<uint4*>(win + x) = vv;
    }
    __syncthreads();
    while (active && !w.done && (w.i >= lo || w.i < 0))
      walk_step<W>(w, tk, reinterpret_cast<const uint32_t*>(win) + (kGran ? (half * 64 + lane) * 4 : half * 64 + lane),
                   a.sp, out, a.cigar_cap, seq, lo);
    __syncthreads();
  }
  if (active) walk_finish(w, i_end, d_end, score, a.cigar_cap, a.res[k]);
}

// int32 fill + walk for bands or lengths the packed kernel does not take, one task per thread
// of a fixed grid (rows and direction bytes are per thread).
__global__ void k_sw_generic_g(SwGenericArgs g, const uint32_t* n_ptr) {
  const uint32_t n = *n_ptr;
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t i = tid; i < n; i += gridDim.x * blockDim.x) sw_generic_one(g, i, tid);
}

// ------------------------------------------------------------------------------ records
__global__ void k_grecords(ReadView rv, MapGeom g, const Cand* cands, const int32_t* n_cand, int max_cand,
                           const int32_t* chosen, const int32_t* second, const int64_t* tb_idx, const SwResult* tb_res,
                           int min_aln_score, const uint8_t* lut, int lut_len, dnaln_record* rec, uint32_t* cig_len,
                           BatchCtl* ctl) {
  const uint32_t n = ctl->n;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const int ch = chosen[r];
    const SwResult* res = ch >= 0 ? &tb_res[tb_idx[r]] : nullptr;
    if (res && (res->n_cigar < 0 || res->status != 0)) atomicOr(&ctl->err, res->n_cigar < 0 ? 1u : 2u);
    dnaln_record m;
    record_read(g, rv.lens[r], n_cand[r], cands + (uint64_t)r * max_cand, ch, second[r], res, min_aln_score, lut,
                lut_len, m);
    if (res && res->n_cigar < 0) m.cigar_len = 0;
    rec[r] = m;
    cig_len[r] = m.cigar_len;
  }
}

__global__ void k_gcigar(ReadView rv, const int32_t* chosen, const int64_t* tb_idx, const SwResult* tb_res,
                         const uint32_t* slots, int cap, const uint64_t* cig_off, dnaln_record* rec, uint32_t* pool,
                         const BatchCtl* ctl) {
  const uint32_t n = ctl->n;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    rec[r].cigar_off = cig_off[r];
    if (rec[r].cigar_len == 0) continue;
    const uint64_t t = (uint64_t)tb_idx[r];
    gather_cigar(rv.lens[r], tb_res[t], slots + t * (uint64_t)cap, pool + cig_off[r]);
  }
}

// ------------------------------------------------------------------------------ launchers
// waves: fill waves the launch covers (the grid); the kernel reads the task count on the device.
template <bool TB>
bool launch_packed_g(int bw, const SwArgs& a, uint64_t waves, const uint32_t* n_ptr, uint32_t first, hipStream_t st) {
  switch (bw) {
#define DN_CASE(B)                                                                                                \
  case B: {                                                                                                       \
    constexpr int kW = 2 * B + 1;                                                                                 \
    const unsigned threads = (unsigned)sw_block(kW, TB);                                                          \
    const unsigned blocks = (unsigned)std::max<uint64_t>(1, (waves * 64 + threads - 1) / threads);                \
    hipLaunchKernelGGL((k_sw_packed_g<kW, TB>), dim3(blocks), dim3(threads),                                      \
                       TB ? (size_t)fill_lds_words(kW) * 4 * (threads / 64) : 0, st, a, n_ptr, first);             \
    return true;                                                                                                  \
  }
    DN_PACKED_BW_LIST(DN_CASE)
#undef DN_CASE
    default:
      return false;
  }
}

// tasks: tasks the launch covers.
bool launch_walk_g(int bw, const SwArgs& a, int walk_kind, uint64_t tasks, const uint32_t* n_ptr, uint32_t first,
                   hipStream_t st) {
  const int W = 2 * bw + 1;
  const int rows = walk_rows(W);
  const size_t lds = (size_t)rows * ((W + 7) / 8) * 512;
  const unsigned waves = (unsigned)std::max<uint64_t>(1, (tasks + 127) / 128);
  const unsigned blocks = (unsigned)std::max<uint64_t>(1, (tasks + 255) / 256);
  switch (bw) {
#define DN_CASE(B)                                                                                                 \
  case B:                                                                                                          \
    if (walk_kind == 1)                                                                                            \
      hipLaunchKernelGGL((k_walk_win_g<2 * B + 1>), dim3(waves), dim3(128), lds, st, a, rows, n_ptr, first);       \
    else                                                                                                           \
      hipLaunchKernelGGL((k_walk_g<2 * B + 1>), dim3(blocks), dim3(256), 0, st, a, n_ptr, first);                  \
    return true;
    DN_PACKED_BW_LIST(DN_CASE)
#undef DN_CASE
    default:
      return false;
  }
}

}  // namespace
