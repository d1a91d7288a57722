// hip_runtime.h (CPU emulation) — the HIP runtime and the gfx950 execution model on the host, so
// that the library's kernels and the host code that launches them run on a CPU under
// AddressSanitizer. Test infrastructure: only the emulated build (DN_EMU) includes this header;
// the product is built with hipcc against the real runtime.
//
// Model: a kernel launch runs its blocks one after another; the threads of a block run as fibers
// on the launching thread, with __syncthreads a block barrier. A wave is 64 consecutive threads; wave
// operations (shuffles, DPP, ballot, readlane, ds_bpermute) are collectives over the wave's
// running lanes: every lane deposits its value, a wave barrier, every lane reads, a barrier.
// __shared__ variables are function statics (blocks never overlap); a kernel's dynamic LDS is
// a heap block of exactly the launch's size. Device memory is host memory of exactly the
// requested size (hipMalloc is malloc), so an access outside a buffer is an ASan report.
// Stream capture records launches; a graph launch replays them. A memset or copy recorded
// during capture is an error, as the library's graphs must hold kernels only.
#pragma once

#include <algorithm>
#include <atomic>
#include <barrier>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <type_traits>
#include <vector>

#include <ucontext.h>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) (reinterpret_cast<const void*>(&(x)))
#define DN_WAVES_PER_EU(n)

struct dim3 {
  uint32_t x, y, z;
  constexpr dim3(uint32_t a = 1, uint32_t b = 1, uint32_t c = 1) : x(a), y(b), z(c) {}
};
struct uint4 {
  uint32_t x, y, z, w;
};

typedef int hipError_t;
enum : int { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorNotReady = 600, hipErrorUnknown = 999 };
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 1 };
#define hipStreamNonBlocking 1u
#define hipHostMallocDefault 0u

namespace hip_emu {

struct Stream {
  bool capturing = false;
  std::vector<std::function<void()>> rec;
};
struct Graph {
  std::vector<std::function<void()>> ops;  // kernel launches, in capture order (a chain)
};
struct GraphNodeRef {
  Graph* g;
  size_t i;
};
struct Event {
  int recorded = 0;
};

// ------------------------------------------------------------------------------ the SIMT model
// The threads of a block run as fibers (ucontext) on the launching thread: a fiber runs until it
// waits at a barrier or returns, and the scheduler resumes the others in turn. A barrier phase
// completes when every live member has arrived; a thread that returns leaves its barriers.
struct Barrier {
  unsigned expected = 0, arrived = 0;
  uint64_t phase = 0;
};
struct Wave {
  Barrier bar;
  uint64_t slot[64];
  uint64_t active = 0;
};
struct Fiber {
  ucontext_t ctx;
  bool done = false;
  uint64_t wait_phase = ~0ull;  // phase awaited (~0: runnable)
  Barrier* wait_on = nullptr;
};
struct Block {
  Barrier bar;
  std::vector<Wave> waves;
  std::vector<Fiber> fib;
  void* lds = nullptr;
  ucontext_t sched;
  unsigned cur = 0;
  bool direct = false;  // threads run one after another without fibers (no synchronisation seen)
  bool synced = false;  // a thread called a barrier or a wave operation
};
inline thread_local dim3 tidx, bidx, bdim, gdim;
inline thread_local Block* blk = nullptr;

inline Wave& my_wave() { return blk->waves[tidx.x >> 6]; }
inline unsigned my_lane() { return tidx.x & 63; }
inline void* dyn_lds() { return blk->lds; }

// Arrive at b and wait for the phase to complete (yield to the scheduler meanwhile).
inline void arrive_and_wait(Barrier& b) {
  blk->synced = true;
  if (blk->direct) {
    std::fprintf(stderr, "hip_emu: a block synchronised where the kernel's first block did not\n");
    std::abort();
  }
  const uint64_t ph = b.phase;
  if (++b.arrived == b.expected) {
    b.arrived = 0;
    ++b.phase;
    return;
  }
  Fiber& f = blk->fib[blk->cur];
  f.wait_on = &b;
  f.wait_phase = ph;
  while (b.phase == ph) swapcontext(&f.ctx, &blk->sched);
  f.wait_on = nullptr;
  f.wait_phase = ~0ull;
}
inline void leave(Barrier& b) {
  --b.expected;
  if (b.expected && b.arrived == b.expected) {
    b.arrived = 0;
    ++b.phase;
  } else if (!b.expected) {
    b.arrived = 0;
    ++b.phase;
  }
}

// Collective over the calling thread's wave: deposit v, then f(slots, active) for the result.
template <class F>
inline uint64_t wave_collect(uint64_t v, F&& f) {
  Wave& w = my_wave();
  w.slot[my_lane()] = v;
  arrive_and_wait(w.bar);
  const uint64_t r = f(w.slot, w.active);
  arrive_and_wait(w.bar);
  return r;
}
template <class T>
inline uint64_t to_bits(T v) {
  uint64_t b = 0;
  std::memcpy(&b, &v, sizeof(T) < 8 ? sizeof(T) : 8);
  return b;
}
template <class T>
inline T from_bits(uint64_t b) {
  T v;
  std::memcpy(&v, &b, sizeof(T));
  return v;
}
template <class T>
inline T shfl_from(T v, int src) {
  const unsigned s = (unsigned)src & 63u;
  return from_bits<T>(wave_collect(to_bits(v), [s](const uint64_t* sl, uint64_t) { return sl[s]; }));
}

inline std::atomic<uint64_t> launches{0};

inline bool trace() {
  static const bool on = std::getenv("HIP_EMU_TRACE") != nullptr;
  return on;
}

constexpr size_t kFiberStack = 64 * 1024;
inline std::vector<char*>& stacks() {
  static thread_local std::vector<char*> s;
  return s;
}

struct FiberStart {
  std::function<void()>* body;
};
inline void fiber_entry(unsigned lo, unsigned hi) {
  const uintptr_t p = (uintptr_t)lo | ((uintptr_t)hi << 32);
  FiberStart* fs = reinterpret_cast<FiberStart*>(p);
  (*fs->body)();
}

// Runs kernel k over the grid: blocks in order, the threads of a block as fibers.
template <class... P, class Tup>
inline void run_kernel(void (*k)(P...), dim3 g, dim3 b, size_t lds, const Tup& args) {
  launches.fetch_add(1);
  if (trace()) std::fprintf(stderr, "hip_emu: launch %p grid %u block %u lds %zu\n", (void*)k, g.x, b.x, lds);
  const unsigned nt = b.x * b.y * b.z;
  if (b.y != 1 || b.z != 1 || g.y != 1 || g.z != 1 || nt == 0 || nt > 1024) {
    std::fprintf(stderr, "hip_emu: unsupported launch shape\n");
    std::abort();
  }
  auto& st = stacks();
  while (st.size() < nt) st.push_back(static_cast<char*>(std::malloc(kFiberStack)));
  // Blocks after the first run their threads directly when the first block never synchronised
  // (kernels whose threads are independent), which keeps large grids quick.
  bool independent = false;
  for (uint32_t bx = 0; bx < g.x; ++bx) {
    if (independent) {
      Block B;
      B.direct = true;
      B.lds = lds ? std::malloc(lds) : nullptr;
      if (B.lds) std::memset(B.lds, 0xA5, lds);
      const unsigned nw = (nt + 63) / 64;
      B.waves.resize(nw);
      blk = &B;
      bidx = dim3(bx, 0, 0);
      bdim = b;
      gdim = g;
      for (unsigned t = 0; t < nt; ++t) {
        tidx = dim3(t, 0, 0);
        Tup mine = args;
        std::apply(k, mine);
      }
      blk = nullptr;
      std::free(B.lds);
      continue;
    }
    Block B;
    B.bar.expected = nt;
    const unsigned nw = (nt + 63) / 64;
    B.waves.resize(nw);
    for (unsigned w = 0; w < nw; ++w) {
      const unsigned lanes = std::min(64u, nt - 64 * w);
      B.waves[w].bar.expected = lanes;
      B.waves[w].active = lanes == 64 ? ~0ull : ((1ull << lanes) - 1);
    }
    B.lds = lds ? std::malloc(lds) : nullptr;
    if (B.lds) std::memset(B.lds, 0xA5, lds);
    B.fib.resize(nt);
    std::vector<std::function<void()>> bodies(nt);
    std::vector<FiberStart> starts(nt);
    blk = &B;
    bidx = dim3(bx, 0, 0);
    bdim = b;
    gdim = g;
    for (unsigned t = 0; t < nt; ++t) {
      bodies[t] = [&, t] {
        Tup mine = args;  // by-value parameters are private to each thread
        std::apply(k, mine);
        Wave& w = B.waves[t >> 6];
        w.active &= ~(1ull << (t & 63));
        leave(w.bar);
        leave(B.bar);
        B.fib[t].done = true;
      };
      starts[t].body = &bodies[t];
      Fiber& f = B.fib[t];
      getcontext(&f.ctx);
      f.ctx.uc_stack.ss_sp = st[t];
      f.ctx.uc_stack.ss_size = kFiberStack;
      f.ctx.uc_link = &B.sched;
      const uintptr_t p = reinterpret_cast<uintptr_t>(&starts[t]);
      makecontext(&f.ctx, (void (*)())fiber_entry, 2, (unsigned)(p & 0xffffffffu), (unsigned)(p >> 32));
    }
    unsigned live = nt;
    while (live) {
      bool progress = false;
      for (unsigned t = 0; t < nt; ++t) {
        Fiber& f = B.fib[t];
        if (f.done) continue;
        if (f.wait_on && f.wait_on->phase == f.wait_phase) continue;  // still waiting
        B.cur = t;
        tidx = dim3(t, 0, 0);
        swapcontext(&B.sched, &f.ctx);
        progress = true;
        if (f.done) --live;
      }
      if (!progress) {
        std::fprintf(stderr, "hip_emu: block %u of kernel %p deadlocked at a barrier\n", bx, (void*)k);
        std::abort();
      }
    }
    independent = !B.synced;
    blk = nullptr;
    std::free(B.lds);
  }
}

template <class... P, class... A>
inline void launch(Stream* s, void (*k)(P...), dim3 g, dim3 b, size_t lds, A&&... a) {
  std::tuple<std::decay_t<P>...> args{static_cast<std::decay_t<P>>(a)...};
  auto op = [k, g, b, lds, args] { run_kernel(k, g, b, lds, args); };
  if (s && s->capturing) s->rec.push_back(op);
  else op();
}

inline bool capturing(Stream* s) { return s && s->capturing; }

}  // namespace hip_emu

#define threadIdx (hip_emu::tidx)
#define blockIdx (hip_emu::bidx)
#define blockDim (hip_emu::bdim)
#define gridDim (hip_emu::gdim)
#define DN_DYN_LDS(T, name) T* name = reinterpret_cast<T*>(hip_emu::dyn_lds())
#define hipLaunchKernelGGL(k, g, b, lds, s, ...) hip_emu::launch((s), k, dim3(g), dim3(b), (size_t)(lds), __VA_ARGS__)

typedef hip_emu::Stream* hipStream_t;
typedef hip_emu::Event* hipEvent_t;
typedef hip_emu::Graph* hipGraph_t;
typedef hip_emu::Graph* hipGraphExec_t;
typedef hip_emu::GraphNodeRef* hipGraphNode_t;
enum hipGraphNodeType { hipGraphNodeTypeKernel = 0, hipGraphNodeTypeEventRecord = 7 };

// ------------------------------------------------------------------------------ device functions
inline void __syncthreads() { hip_emu::arrive_and_wait(hip_emu::blk->bar); }
inline void __builtin_amdgcn_wave_barrier_emu() { hip_emu::arrive_and_wait(hip_emu::my_wave().bar); }
#define __builtin_amdgcn_wave_barrier() __builtin_amdgcn_wave_barrier_emu()
#define __builtin_amdgcn_fence(order, scope) ((void)0)

template <class T>
inline T __shfl(T v, int src, int width = 64) {
  return hip_emu::shfl_from(v, src);
}
template <class T>
inline T __shfl_xor(T v, int mask, int width = 64) {
  return hip_emu::shfl_from(v, (int)(hip_emu::my_lane() ^ (unsigned)mask));
}
template <class T>
inline T __shfl_up(T v, unsigned d, int width = 64) {
  const unsigned l = hip_emu::my_lane();
  return hip_emu::shfl_from(v, (int)(l >= d ? l - d : l));
}
inline int __builtin_amdgcn_readlane_emu(int v, int l) { return hip_emu::shfl_from(v, l); }
#define __builtin_amdgcn_readlane(v, l) __builtin_amdgcn_readlane_emu((v), (l))
// The first active lane's value. A collective, as the walks that take memory state through it
// rely on the wave's lanes stepping together (lanes here run one by one between collectives).
inline int __builtin_amdgcn_readfirstlane_emu(int v) {
  return (int)hip_emu::wave_collect((uint64_t)(uint32_t)v, [](const uint64_t* sl, uint64_t act) {
    return sl[act ? __builtin_ctzll(act) : 0];
  });
}
#define __builtin_amdgcn_readfirstlane(v) __builtin_amdgcn_readfirstlane_emu((int)(v))
inline int __builtin_amdgcn_ds_bpermute_emu(int addr, int v) { return hip_emu::shfl_from(v, (addr >> 2) & 63); }
#define __builtin_amdgcn_ds_bpermute(a, v) __builtin_amdgcn_ds_bpermute_emu((a), (v))
inline uint64_t __builtin_amdgcn_ballot_w64_emu(bool p) {
  return hip_emu::wave_collect(p ? 1 : 0, [](const uint64_t* sl, uint64_t act) {
    uint64_t m = 0;
    for (int l = 0; l < 64; ++l)
      if (((act >> l) & 1) && sl[l]) m |= 1ull << l;
    return m;
  });
}
#define __builtin_amdgcn_ballot_w64(p) __builtin_amdgcn_ballot_w64_emu((p))
// DPP moves the library uses: row_shr:1..15 (0x111..0x11f), row_bcast:15 (0x142), row_bcast:31
// (0x143), with the row mask selecting the rows that take the moved value.
inline int __builtin_amdgcn_update_dpp_emu(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl) {
  const int lane = (int)hip_emu::my_lane(), row = lane >> 4;
  int from = -1;
  if (ctrl >= 0x111 && ctrl <= 0x11f) {
    const int sh = ctrl - 0x110;
    if ((lane & 15) >= sh) from = lane - sh;
  } else if (ctrl == 0x142) {
    if (row >= 1) from = 16 * row - 1;
  } else if (ctrl == 0x143) {
    if (row >= 2) from = 31;
  } else {
    std::fprintf(stderr, "hip_emu: DPP control %#x not emulated\n", ctrl);
    std::abort();
  }
  const int got = hip_emu::shfl_from(src, from < 0 ? lane : from);
  const bool on = ((row_mask >> row) & 1) && ((bank_mask >> ((lane & 15) >> 2)) & 1);
  if (!on) return old;
  if (from < 0) return bound_ctrl ? 0 : old;
  return got;
}
#define __builtin_amdgcn_update_dpp(o, s, c, r, b, bc) __builtin_amdgcn_update_dpp_emu((o), (s), (c), (r), (b), (bc))

template <class T, class U>
inline T atomicAdd(T* p, U v) {
  return std::atomic_ref<T>(*p).fetch_add((T)v);
}
template <class T, class U>
inline T atomicOr(T* p, U v) {
  return std::atomic_ref<T>(*p).fetch_or((T)v);
}
template <class T, class U>
inline T atomicMax(T* p, U v) {
  std::atomic_ref<T> a(*p);
  T cur = a.load();
  while (cur < (T)v && !a.compare_exchange_weak(cur, (T)v)) {
  }
  return cur;
}
template <class T, class U, class V>
inline T atomicCAS(T* p, U expected, V desired) {
  T e = (T)expected;
  std::atomic_ref<T>(*p).compare_exchange_strong(e, (T)desired);
  return e;
}
template <class T>
inline T min(T a, T b) {
  return a < b ? a : b;
}
template <class T>
inline T max(T a, T b) {
  return a < b ? b : a;
}

// ------------------------------------------------------------------------------ runtime API
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "emulated error"; }
// Device memory starts undefined: every byte of a new allocation is HIP_EMU_POISON (default 0xCD,
// a negative int32; 0x01 makes stale counts large and positive, as reused device memory can be).
inline int hip_emu_poison() {
  static const int v = [] {
    const char* e = std::getenv("HIP_EMU_POISON");
    return e ? (int)(std::strtol(e, nullptr, 0) & 0xff) : 0xCD;
  }();
  return v;
}
// HIP_EMU_STALE=1: a new allocation starts with the bytes of earlier freed allocations instead
// (a device allocator hands out reused memory), so stale counts, offsets and addresses of freed
// buffers turn up where code reads memory it has not written.
namespace hip_emu {
struct StalePool {
  std::mutex mu;
  std::vector<unsigned char> bytes;  // contents of freed allocations, oldest overwritten first
  size_t head = 0, fill = 0;
  std::vector<std::pair<void*, size_t>> live;
  static bool on() {
    static const bool v = std::getenv("HIP_EMU_STALE") != nullptr;
    return v;
  }
};
inline StalePool& stale_pool() {
  static StalePool p;
  return p;
}
}  // namespace hip_emu
inline hipError_t hipMalloc(void** p, size_t n) {
  *p = std::malloc(n ? n : 1);
  if (!*p) return hipErrorInvalidValue;
  std::memset(*p, hip_emu_poison(), n);
  if (hip_emu::StalePool::on()) {
    auto& sp = hip_emu::stale_pool();
    std::lock_guard<std::mutex> lk(sp.mu);
    unsigned char* d = static_cast<unsigned char*>(*p);
    for (size_t i = 0; sp.fill && i < n;) {
      const size_t off = (sp.head + i) % sp.fill, k = std::min(n - i, sp.fill - off);
      std::memcpy(d + i, sp.bytes.data() + off, k);
      i += k;
    }
    sp.head = sp.fill ? (sp.head + n) % sp.fill : 0;
    sp.live.emplace_back(*p, n);
  }
  return hipSuccess;
}
inline hipError_t hipFree(void* p) {
  if (p && hip_emu::StalePool::on()) {
    auto& sp = hip_emu::stale_pool();
    std::lock_guard<std::mutex> lk(sp.mu);
    for (size_t i = 0; i < sp.live.size(); ++i)
      if (sp.live[i].first == p) {
        constexpr size_t kCap = 64u << 20;
        if (sp.bytes.size() < kCap) sp.bytes.resize(kCap);
        const unsigned char* s = static_cast<const unsigned char*>(p);
        for (size_t j = 0; j < sp.live[i].second && j < kCap; ++j) sp.bytes[(sp.fill + j) % kCap] = s[j];
        sp.fill = std::min(kCap, sp.fill + sp.live[i].second);
        sp.live.erase(sp.live.begin() + (long)i);
        break;
      }
  }
  std::free(p);
  return hipSuccess;
}
inline hipError_t hipHostMalloc(void** p, size_t n, unsigned) { return hipMalloc(p, n); }
inline hipError_t hipHostFree(void* p) { return hipFree(p); }
inline hipError_t hipMemGetInfo(size_t* f, size_t* t) {
  // 64 GiB free by default; HIP_EMU_FREE_GB sets another (288 for the planning of an MI355X).
  const char* e = std::getenv("HIP_EMU_FREE_GB");
  *f = (e ? std::strtoull(e, nullptr, 10) : 64ull) << 30;
  *t = *f;
  return hipSuccess;
}
inline hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) {
  // Eight "compute units" keep grids that scale with the CU count small; HIP_EMU_CUS sets another.
  const char* e = std::getenv("HIP_EMU_CUS");
  *v = e ? std::atoi(e) : 8;
  return hipSuccess;
}
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  *s = new hip_emu::Stream;
  return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t s) {
  delete s;
  return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline void hip_emu_no_capture(hipStream_t s, const char* what) {
  if (hip_emu::capturing(s)) {
    std::fprintf(stderr, "hip_emu: %s recorded into a captured graph\n", what);
    std::abort();
  }
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t st) {
  hip_emu_no_capture(st, "hipMemcpyAsync");
  if (n) std::memmove(d, s, n);
  return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  if (n) std::memmove(d, s, n);
  return hipSuccess;
}
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t st) {
  hip_emu_no_capture(st, "hipMemsetAsync");
  std::memset(d, v, n);
  return hipSuccess;
}
inline hipError_t hipEventCreate(hipEvent_t* e) {
  *e = new hip_emu::Event;
  return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  hip_emu_no_capture(s, "hipEventRecord");  // the library's graphs hold kernels only
  e->recorded = 1;
  return hipSuccess;
}
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  *ms = 0.001f;
  return hipSuccess;
}
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) {
  *p = const_cast<void*>(sym);
  return hipSuccess;
}
inline hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  s->capturing = true;
  s->rec.clear();
  return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
  s->capturing = false;
  *g = new hip_emu::Graph{std::move(s->rec)};
  s->rec.clear();
  return hipSuccess;
}
// Nodes of a captured graph: its launches, each depending on the one before.
inline hipError_t hipGraphGetNodes(hipGraph_t g, hipGraphNode_t* nodes, size_t* n) {
  static thread_local std::vector<hip_emu::GraphNodeRef> refs;
  if (nodes) {
    refs.resize(g->ops.size());
    for (size_t i = 0; i < *n && i < g->ops.size(); ++i) {
      refs[i] = hip_emu::GraphNodeRef{g, i};
      nodes[i] = &refs[i];
    }
  }
  *n = g->ops.size();
  return hipSuccess;
}
inline hipError_t hipGraphNodeGetType(hipGraphNode_t, hipGraphNodeType* t) {
  *t = hipGraphNodeTypeKernel;
  return hipSuccess;
}
inline hipError_t hipGraphNodeGetDependencies(hipGraphNode_t nd, hipGraphNode_t*, size_t* n) {
  *n = nd->i > 0 ? 1 : 0;
  return hipSuccess;
}
inline hipError_t hipGraphNodeGetDependentNodes(hipGraphNode_t nd, hipGraphNode_t*, size_t* n) {
  *n = nd->i + 1 < nd->g->ops.size() ? 1 : 0;
  return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, void*, size_t) {
  *e = new hip_emu::Graph{g->ops};
  return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  delete e;
  return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s) {
  hip_emu_no_capture(s, "hipGraphLaunch");
  for (auto& op : e->ops) op();
  return hipSuccess;
}
