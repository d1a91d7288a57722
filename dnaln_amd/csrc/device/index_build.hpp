// index_build.hpp — MinimizerIndex::build (src/index.cpp:10-45) split into independent pieces
// of work, for the GPU build (one thread per genome segment, one per key).
//
// The reference sketches each contig front to back. Its emitted positions have a closed
// description that needs no streaming state: a k-mer start p is emitted exactly when p is the
// leftmost smallest hash of at least one valid window (w k-mers, no N) of the contig. Each
// distinct window minimum occupies one contiguous run of windows (a k-mer in two windows is in
// every window between them), the emitted positions strictly increase, and `last` can never
// equal the first minimum after an N. So a segment of k-mer starts [s, e) is sketched by running
// the reference's sketch over the bases of every window that contains one of them and keeping the
// emissions that fall in [s, e); the segments of a contig together give the contig's list.
#pragma once

#include "seed_chain.hpp"
#include "sketch_vh.hpp"

namespace dn {

// One thread's share of the genome: k-mer starts [s, e) of the contig whose codes begin at
// `base` (an offset into the reference) and hold `clen` bases.
struct IndexSeg {
  uint64_t base;
  uint32_t clen;
  uint32_t s, e;
  uint32_t pad;
};

// Calls emit(hash, p, strand) for every minimizer of the contig with p in [g.s, g.e), p counted
// from the contig's start, in increasing p. `ref` must allow 16-byte chunk reads up to 16 bytes
// before and 32 + w bytes after each contig (guard bytes). The window length 10 without N takes
// the branch-free register sketch; anything else the streamed sketch over `ring`.
template <class Emit>
DN_HD void sketch_segment(const uint8_t* ref, const IndexSeg& g, int k, int w, Ring ring, Emit&& emit) {
  if (g.clen < (uint32_t)k || g.s >= g.e) return;
  const uint32_t nk = g.clen - (uint32_t)k + 1;  // k-mer starts of the contig
  // k-mers of the windows that contain s .. e-1: [s - w + 1, e - 1 + w - 1], inside the contig
  const uint32_t k0 = g.s >= (uint32_t)(w - 1) ? g.s - (uint32_t)(w - 1) : 0;
  const uint32_t k1 = g.e + (uint32_t)(w - 1) < nk ? g.e + (uint32_t)(w - 1) : nk;  // exclusive
  const int len = (int)(k1 - k0) + k - 1;
  const uint8_t* seq = ref + g.base + k0;
  auto keep = [&](uint64_t h, uint32_t p, uint32_t st) {
    const uint32_t q = p + k0;
    if (q >= g.s && q < g.e) emit(h, q, st);
  };
  if (w == 10 && !read_has_n(seq, len)) {
    if (k <= 15) sketch_vh<10, uint32_t>(ChunkReader(seq), len, k, keep);
    else sketch_vh<10, uint64_t>(ChunkReader(seq), len, k, keep);
  } else {
    sketch_stream(ChunkReader(seq), len, k, w, ring, keep);
  }
}

// Sort key of an occurrence: hash above (contig offset + p) << 1 | strand, so an ascending sort
// gives the reference's (hash, pos) order (src/index.cpp:20-24). Needs 2k <= 32.
DN_HD uint64_t occ_key(uint64_t h, uint64_t base, uint32_t p, uint32_t st) {
  return h << 32 | (uint32_t)((base + p) << 1 | st);
}

}  // namespace dn
