// sketch_vh.hpp — the minimizer sketch of an N-free read with the window length fixed at compile
// time, branch-free per k-mer (the fused seed + chain kernel and the global seed path).
#pragma once

#include "common.hpp"

namespace dn {

// Minimizer list entry: hash << 17 | pos << 1 | strand (2k <= 46 bits, reads < 65536 bp).
DN_HD uint64_t pack_min(uint64_t h, uint32_t pos, uint32_t strand) { return h << 17 | (uint64_t)pos << 1 | strand; }

// hash64 (src/minimizer.hpp:21-30) in the k-mer's own word type. Every step either masks to 2k
// bits or shifts a masked value right, so for 2k <= 32 the 32-bit form gives the same hash.
template <class T>
DN_HD T hash_kmer(T key, T mask) {
  key = (~key + (key << 21)) & mask;
  key = key ^ key >> 24;
  key = ((key + (key << 3)) + (key << 8)) & mask;
  key = key ^ key >> 14;
  key = ((key + (key << 2)) + (key << 4)) & mask;
  key = key ^ key >> 28;
  key = (key + (key << 31)) & mask;
  return key;
}
// The 32-bit form, with shift amounts offset by z, a value that is 0 at run time but unknown to
// the compiler: it then keeps every a + (a << s) as one v_lshl_add_u32 instead of turning the
// shift-add chains into v_mul_lo_u32, which issues at a quarter of the rate.
DN_HD uint32_t hash_kmer32(uint32_t key, uint32_t mask, uint32_t z) {
  key = (~key + (key << (21 + z))) & mask;
  key = key ^ key >> 24;
  key = ((key + (key << (3 + z))) + (key << (8 + z))) & mask;
  key = key ^ key >> 14;
  key = ((key + (key << (2 + z))) + (key << (4 + z))) & mask;
  key = key ^ key >> 28;
  key = (key + (key << (31 + z))) & mask;
  return key;
}
template <class T>
DN_HD T hash_k(T key, T mask, uint32_t) {
  return hash_kmer<T>(key, mask);
}
template <>
DN_HD uint32_t hash_k<uint32_t>(uint32_t key, uint32_t mask, uint32_t z) {
  return hash_kmer32(key, mask, z);
}

// The reference's streamed van Herk / Gil-Werman window minimum (src/minimizer.cpp:5-68) with
// the window length fixed at compile time, for a read without N: block and suffix minima are
// registers, and the step of every k-mer is the same branch-free code on every lane (emission is
// a predicated store). The minimizer of a window is its leftmost smallest hash, as the
// reference's tie rules give: strict < within a block prefix, <= within a suffix, and the suffix
// (the left part) wins a tie between the two. T is uint32_t for k <= 15 (the hash then stays
// below the all-ones sentinel) and uint64_t otherwise. Calls emit(hash, pos, strand) for every
// minimizer in position order (a predicated call: every lane runs the same steps). The reader
// is called in increasing position order, for up to WT - 1 positions past the read; what it
// returns there is never used.
template <int WT, class T, class Reader, class Emit>
DN_HD void sketch_vh(Reader&& rd, int len, int k, Emit&& emit) {
  const T mask = (T)((1ULL << (2 * k)) - 1);
  const int shift = 2 * (k - 1);
  const T kInvalid = ~(T)0;
  const uint32_t z = (uint32_t)k >> 8;  // 0 (k <= 28)
  T fwd = 0, rev = 0;
  if (len < k) return;
  for (int i = 0; i < k - 1; ++i) {
    const uint32_t c = rd(i);
    fwd = ((fwd << 2) | c) & mask;
    rev = (rev >> 2) | ((T)(3 - c) << shift);
  }
  T bh[WT], sh[WT];
  uint32_t bp[WT], sp[WT];
#pragma unroll
  for (int j = 0; j < WT; ++j) bh[j] = sh[j] = kInvalid, bp[j] = sp[j] = 0;
  T ph = kInvalid;
  uint32_t pp = 0, last = ~0u;
  const int nk = len - k + 1;
  for (int x0 = 0; x0 < nk; x0 += WT) {
#pragma unroll
    for (int j = 0; j < WT; ++j) {
      const int x = x0 + j;
      const uint32_t c = rd(x + k - 1);  // past the read (x >= nk) the k-mer is never used
      fwd = ((fwd << 2) | c) & mask;
      rev = (rev >> 2) | ((T)(3 - c) << shift);
      const bool strand = rev < fwd;
      const T h = fwd != rev ? hash_k<T>(strand ? rev : fwd, mask, z) : kInvalid;
      const uint32_t ps = (uint32_t)x << 1 | (uint32_t)strand;
      bh[j] = h;
      bp[j] = ps;
      if (j == 0) {
        ph = h;
        pp = ps;
      } else {
        const bool lt = h < ph;
        ph = lt ? h : ph;
        pp = lt ? ps : pp;
      }
      // Window [x - WT + 1, x]: the suffix of the previous block, then this block's prefix.
      T mh = ph;
      uint32_t mps = pp;
      if (j + 1 < WT) {
        const bool suf = !(ph < sh[j + 1]);
        mh = suf ? sh[j + 1] : ph;
        mps = suf ? sp[j + 1] : pp;
      }
      const bool live = x + 1 >= WT && x < nk && mh != kInvalid;
      const uint32_t p = mps >> 1;
      if (live && p != last) emit((uint64_t)mh, p, mps & 1);
      last = live ? p : last;
    }
    // block complete: suffix minima, right to left (ties keep the left k-mer)
    sh[WT - 1] = bh[WT - 1];
    sp[WT - 1] = bp[WT - 1];
#pragma unroll
    for (int r = WT - 2; r >= 0; --r) {
      const bool left = bh[r] <= sh[r + 1];
      sh[r] = left ? bh[r] : sh[r + 1];
      sp[r] = left ? bp[r] : sp[r + 1];
    }
  }
}

// True when seq[0, len) holds a code above 3 (an N), read as 16-byte aligned chunks with all
// loads of a batch of eight issued together. Touches up to 15 bytes before and after the read:
// the read buffer carries guard bytes.
DN_HD bool read_has_n(const uint8_t* seq, int len) {
  const uintptr_t a = (uintptr_t)seq;
  const uint32_t off = (uint32_t)(a & 15);
  const int nch = (int)((off + (uint32_t)len + 15) >> 4);
  const uint64_t src = (uint64_t)(a - off);
  uint32_t bad = 0;
  for (int b0 = 0; b0 < nch; b0 += 8) {
    Quad c[8];
#pragma unroll
    for (int x = 0; x < 8; ++x) c[x] = load16(src + 16 * (uint64_t)(b0 + x < nch ? b0 + x : nch - 1));
#pragma unroll
    for (int x = 0; x < 8; ++x) {
      if (b0 + x >= nch) continue;
      const int q = 16 * (b0 + x) - (int)off;  // read position of the chunk's byte 0
      const uint32_t dw[4] = {c[x].x, c[x].y, c[x].z, c[x].w};
#pragma unroll
      for (int d = 0; d < 4; ++d) bad |= dw[d] & bytes_from(-(q + 4 * d)) & bytes_below(len - (q + 4 * d));
    }
  }
  return (bad & 0xFCFCFCFCu) != 0;
}

}  // namespace dn
