// chain_wave.hpp — chaining with one wave per read (the batched form's chain stage for reads
// that leave the fused LDS kernel: long reads and reads with many anchors).
//
// Same results as chain_read (src/chain.cpp:12-71, src/mapper.cpp:44-62), organised for a wave:
//   * the read's anchors are sorted with a wave-wide bitonic network whose comparators all put
//     the smaller key first, so a virtual +inf pad beyond n never moves into range;
//   * DP: the last 64 anchors sit in a circular register window, anchor j in lane j % 64 (its
//     key, f and prev), and each lane scores its anchor as predecessor of anchor i. The
//     reference's scan stops at the first j that changes strand or lies more than max_gap
//     behind; both conditions are monotone in j on sorted anchors, so each lane tests its own
//     j. A DPP max-reduction over (score << 6 | 63 - (i-1-j)) picks the best score and, on ties,
//     the nearest j, as the reference's strict > scan from j = i-1 down does. Lane i % 64 then
//     takes anchor i's f and prev, which reach memory once per 64 anchors;
//   * extraction walks ends in (f desc, index asc) order, a second bitonic sort of packed
//     keys; chains go to a list that is sorted (score desc, extraction order asc) the same way.
//     prev is stored as prev + 1 with the used flag in bit 31, so a walk step is one dependent
//     load, and ends already used when their block of 64 starts are skipped by a ballot;
//   * candidates: lane x keeps accepted candidate x's diagonal, so the dedupe test is a ballot.
// Every lane writes the same value wherever all lanes later read (prev during extraction), so
// each lane reads its own writes. In the CPU emulation, where the lanes of a wave run one after
// another between collectives, the walk takes prev through readfirstlane (wave_same), so every
// lane follows the first lane's value.
// Lookback <= 64 and max_cand <= 64 (checked by the host).
#pragma once

#include "seed_chain.hpp"

namespace dn {
// A value every lane of the wave holds alike: as loaded on the device; through readfirstlane in
// the CPU emulation.
DN_HD uint32_t wave_same(uint32_t v) {
#if defined(DN_EMU)
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
#else
  return v;
#endif
}
}  // namespace dn

namespace dn {

// Makes a wave's earlier LDS and global writes visible to its later reads.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Maximum over the 64 lanes (all lanes active), returned to every lane.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true));  // row_shr:1
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true));  // row_shr:2
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true));  // row_shr:4
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true));  // row_shr:8
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false)); // row_bcast:15
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false)); // row_bcast:31
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Ascending sort of a[0..n) by the wave (the 64 lanes take the comparators of each step).
template <class T>
__device__ __forceinline__ void wave_sort(T* a, int n, int lane) {
  int N = 1;
  while (N < n) N <<= 1;
  for (int k = 2; k <= N; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < N; i += 64) {
        const int p = j == (k >> 1) ? (i ^ (k - 1)) : (i ^ j);
        if (p > i && p < n) {
          const T x = a[i], y = a[p];
          if (y < x) {
            a[i] = y;
            a[p] = x;
          }
        }
      }
      wave_sync();
    }
  }
}

// Lane lane ^ mask's value (all lanes active).
__device__ __forceinline__ uint64_t shfl_xor64(uint64_t x, int mask, int lane) {
  const int addr = (lane ^ mask) << 2;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)(uint32_t)x);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)(uint32_t)(x >> 32));
  return (uint64_t)hi << 32 | lo;
}

// The network of wave_sort over 64 * R keys held in registers: key e in v[e / 64] of lane
// e % 64. A comparator's partner e ^ m sits in slot (e / 64) ^ (m / 64) of lane
// (e % 64) ^ (m % 64), fetched with ds_bpermute; the lower index keeps the smaller key.
template <int R>
__device__ __forceinline__ void reg_sort(uint64_t (&v)[R], int lane) {
  for (int k = 2; k <= 64 * R; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int m = j == (k >> 1) ? k - 1 : j;
      const int ms = m >> 6, ml = m & 63;
      uint64_t pv[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        uint64_t src = v[0];
#pragma unroll
        for (int q = 1; q < R; ++q) src = (r ^ ms) == q ? v[q] : src;
        pv[r] = ml ? shfl_xor64(src, ml, lane) : src;
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const bool lower = j >= 64 ? (r & (j >> 6)) == 0 : (lane & j) == 0;
        const bool lt = v[r] < pv[r];
        v[r] = lower == lt ? v[r] : pv[r];
      }
    }
  }
}

// Ascending sort of a[0..n) through registers (n <= 64 * R), padded with ~0 keys.
template <int R>
__device__ __forceinline__ void reg_sort_array(uint64_t* a, int n, int lane) {
  uint64_t v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = 64 * r + lane < n ? a[64 * r + lane] : ~0ull;
  reg_sort<R>(v, lane);
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (64 * r + lane < n) a[64 * r + lane] = v[r];
  wave_sync();
}

// Sorts a[0..n) ascending: in registers up to 256 keys, else the memory network. No key of
// the callers is ~0, so the padding stays beyond n.
__device__ __forceinline__ void sort_keys(uint64_t* a, int n, int lane) {
  if (n <= 64) reg_sort_array<1>(a, n, lane);
  else if (n <= 128) reg_sort_array<2>(a, n, lane);
  else if (n <= 256) reg_sort_array<4>(a, n, lane);
  else wave_sort(a, n, lane);
}

// Per-read working arrays (LDS, or global scratch for reads too large for LDS).
struct WaveChainMem {
  uint64_t* keys;     // n anchors
  uint64_t* order;    // n sort keys
  int32_t* f;         // n
  uint32_t* prev;     // n: prev + 1 in bits 0..30, bit 31 set once extraction has used the anchor
  int32_t* chains;    // chain_cap x {score, rpos | strand << 31, dmin, dmax}
};

__device__ __forceinline__ void chain_wave_read(const ChainArgs& a, uint32_t r, int n, const WaveChainMem& s,
                                                const int32_t* cost, int lane) {
  Cand* cands = a.cands + (uint64_t)r * (uint64_t)a.max_cand;
  if (n == 0) {
    if (lane == 0) a.n_cand[r] = 0;
    return;
  }
  // 1. anchors in (strand, rpos, qpos) order
  sort_keys(s.keys, n, lane);

  // 2. DP (src/chain.cpp:21-45). 32-bit arithmetic: rpos < 2^31 and qpos < 2^31, and
  // dr, dq and their difference are only used where 0 < dr <= max_gap and 0 < dq. With
  // a.cost_nib_ok the gap cost comes from the four-bit table in the kernel arguments (scalar
  // registers), which keeps a memory round trip off the serial chain of steps.
  const int k = a.k;
  const uint32_t max_gap = (uint32_t)a.max_gap;
  const int band = a.band;
  const bool nib = a.cost_nib_ok;
  constexpr int kBias = 1 << 20;
  // Keys of anchors i0 .. i0 + 63 (kblk) and of the block before (kprv), one per lane, with
  // the next block loaded at a block's start and read only at its end, so no step waits for
  // memory: a step reads anchor i's key from lane i % 64 (readlane) and each lane's window
  // anchor from kblk or kprv. fw and pw hold the f and prev + 1 of the lane's window anchor;
  // after the block's last step they hold the whole block, which goes to memory in one store.
  uint64_t kprv = 0, kblk = lane < n ? s.keys[lane] : 0;
  int fw = 0;
  uint32_t pw = 0;
  for (int i0 = 0; i0 < n; i0 += 64) {
    const uint64_t knxt = i0 + 64 + lane < n ? s.keys[i0 + 64 + lane] : 0;
    const int iend = n < i0 + 64 ? n : i0 + 64;
    for (int i = i0; i < iend; ++i) {
      const int ri = i - i0;  // lane of anchor i
      const uint64_t ki = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(kblk >> 32), ri) << 32 |
                          (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)kblk, ri);
      const uint32_t hi_i = (uint32_t)(ki >> 32);
      const int qi = key_qpos(ki);
      const uint64_t kw = lane < ri ? kblk : kprv;  // anchor j = i - 1 - dist
      const int dist = (ri - 1 - lane) & 63;
      const int reach = a.lookback < i ? a.lookback : i;  // j >= 0 and within the lookback
      // Every lane computes its score and one select keeps it, so a step has no divergent
      // branch; the byte table is read at min(dd, band), inside the table for every lane.
      const uint32_t hj = (uint32_t)(kw >> 32);
      const uint32_t dr = hi_i - hj;  // same strand: the strand bits cancel
      const int dq = qi - key_qpos(kw);
      const uint32_t t = dr - (uint32_t)dq;
      const uint32_t dd = (int32_t)t < 0 ? 0u - t : t;
      const bool ok = dist < reach && ((hi_i ^ hj) >> 31) == 0 && dr <= max_gap && dr > 0 && dq > 0 &&
                      dd <= (uint32_t)band;
      const int m = dq < (int)dr ? dq : (int)dr;
      const int c = nib ? (dd == (uint32_t)band ? a.cost_at_band : (int)((a.cost_nib >> (4 * (dd & 15))) & 15u))
                        : cost[dd < (uint32_t)band ? dd : (uint32_t)band];
      const int sc = fw + (m < k ? m : k) - c;
      const uint32_t cand = ok ? (uint32_t)(sc + kBias) << 6 | (uint32_t)(63 - dist) : 0u;
      const uint32_t best = wave_max_u32(cand);
      int fi = k, pi = -1;
      if (best) {
        const int sc = (int)(best >> 6) - kBias;
        if (sc > k) {
          fi = sc;
          pi = i - 1 - (63 - (int)(best & 63));
        }
      }
      if (lane == ri) {
        fw = fi;
        pw = (uint32_t)(pi + 1);
      }
    }
    if (lane < iend - i0) {
      s.f[i0 + lane] = fw;
      s.prev[i0 + lane] = pw;
    }
    kprv = kblk;
    kblk = knxt;
  }
  wave_sync();

  // 3. extraction order: f desc, index asc (src/chain.cpp:47-52)
  for (int x = lane; x < n; x += 64) s.order[x] = (uint64_t)(uint32_t)(0x7fffffff - s.f[x]) << 32 | (uint32_t)x;
  wave_sync();
  sort_keys(s.order, n, lane);

  // 4. chains (src/chain.cpp:53-68), lane-uniform walks. Ends are taken 64 at a time: those
  // unused at the block's start are visited in order and checked again before their walk,
  // since an earlier walk of the block may have used them.
  int m = 0;
  for (int o0 = 0; o0 < n; o0 += 64) {
    const uint32_t my_end = o0 + lane < n ? (uint32_t)s.order[o0 + lane] : 0u;
    uint64_t open = __builtin_amdgcn_ballot_w64(o0 + lane < n && !(s.prev[my_end] >> 31));
    while (open) {
      const int end = __builtin_amdgcn_readlane((int)my_end, (int)__builtin_ctzll(open));
      open &= open - 1;
      uint32_t pu = wave_same(s.prev[end]);
      if (pu >> 31) continue;
      int i = end, cnt = 0, stop = -1;
      int64_t dmin = INT64_MAX, dmax = INT64_MIN;
      for (;;) {
        const uint64_t kk = s.keys[i];
        s.prev[i] = pu | 0x80000000u;
        ++cnt;
        const int64_t d = (int64_t)key_rpos(kk) - key_qpos(kk);
        dmin = d < dmin ? d : dmin;
        dmax = d > dmax ? d : dmax;
        const int nx = (int)(pu & 0x7fffffffu) - 1;
        if (nx < 0) break;
        pu = wave_same(s.prev[nx]);
        if (pu >> 31) {
          stop = nx;
          break;
        }
        i = nx;
      }
      const int score = s.f[end] - (stop >= 0 ? s.f[stop] : 0);
      if (score < a.min_score || cnt < a.min_anchors) continue;
      const uint64_t ke = s.keys[end];
      int32_t* c = s.chains + 4 * m;
      c[0] = score;
      c[1] = (int32_t)(key_rpos(ke) | key_strand(ke) << 31);
      c[2] = (int32_t)dmin;
      c[3] = (int32_t)dmax;
      ++m;
    }
  }
  wave_sync();
  if (m == 0) {
    if (lane == 0) a.n_cand[r] = 0;
    return;
  }

  // 5. chains by score desc, extraction order asc (the reference's stable sort)
  for (int x = lane; x < m; x += 64)
    s.order[x] = (uint64_t)(uint32_t)(0x7fffffff - s.chains[4 * x]) << 32 | (uint32_t)x;
  wave_sync();
  sort_keys(s.order, m, lane);

  // 6. candidates (src/mapper.cpp:44-62)
  const double threshold = s.chains[4 * (int)(uint32_t)s.order[0]] * a.pri_ratio;
  const int64_t bw = a.bw;
  int nc = 0;
  int64_t my_diag = 0;   // lane x < nc: candidate x
  uint32_t my_strand = 2;
  for (int o = 0; o < m; ++o) {
    const int32_t* c = s.chains + 4 * (int)(uint32_t)s.order[o];
    if (c[0] < threshold || nc >= a.max_cand) break;
    const uint32_t strand = (uint32_t)c[1] >> 31;
    const int64_t diag = (int64_t)c[2] + ((int64_t)c[3] - c[2]) / 2;
    const int64_t dd = my_diag - diag;
    const bool dup_here = lane < nc && my_strand == strand && (dd < 0 ? -dd : dd) <= bw;
    if (__builtin_amdgcn_ballot_w64(dup_here)) continue;
    if (lane == nc) {
      my_diag = diag;
      my_strand = strand;
      Cand cd;
      cd.rev = strand;
      cd.contig = contig_of(a.contig_off, a.n_contigs, (uint32_t)c[1] & 0x7fffffffu);
      cd.diag = diag;
      cd.score = c[0];
      cd.reserved = 0;
      cands[nc] = cd;
    }
    ++nc;
  }
  if (lane == 0) a.n_cand[r] = nc;
}

}  // namespace dn
