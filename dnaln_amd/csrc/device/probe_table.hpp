// probe_table.hpp — the index's bucket table in 8-byte entries, for the seed probes (k <= 15).
//
// The reference's table (src/index.hpp:49-56) is 16-byte buckets {u64 key, u32 offset, u32 count}
// probed linearly from hash & mask (src/index.hpp:24-38). A minimizer hash has 2k bits, so for
// k <= 15 it fits 30 bits and a bucket fits 8 bytes, kept slot for slot:
//   low word   hash << 2 | tag, or kCEmpty for an empty slot
//   high word  tag kCOne:  the key's single occurrence (rpos << 1 | strand)
//              tag kCMany: 2 .. max_occ occurrences; the index of the key's block in `pos2`:
//                          pos2[val] is the count, pos2[val + 1 ..] the occurrences in the
//                          index's order (ascending position)
//              tag kCSkip: more than max_occ occurrences, which the mapper skips
//                          (src/mapper.cpp:27-28); unused
// The slots are the 16-byte table's, so a probe visits the same slots and stops at the same key
// or empty slot as the reference's: every lookup yields the same occurrences. kCTail entries past
// the last slot repeat slots 0, 1, 2, so the four slots from any home slot are read without a
// wrap. A probe reads those four together (one 32-byte span, usually one cache line, where a
// 16-byte bucket needs a line per slot), and a key with several occurrences then needs one more
// access for its count and first occurrences together.
#pragma once

#include "common.hpp"

namespace dn {

// A bucket of the 16-byte table (src/index.hpp:49-52).
struct Bucket {
  uint64_t key;
  uint32_t offset, count;
};

constexpr uint32_t kCEmpty = 0xffffffffu;
constexpr int kCTail = 3;
constexpr uint32_t kCOne = 0, kCMany = 1, kCSkip = 2;

DN_HD uint32_t c_key(uint64_t e) { return (uint32_t)e; }
DN_HD uint32_t c_val(uint64_t e) { return (uint32_t)(e >> 32); }

// Words of pos2 that bucket b needs: the count and the occurrences, for a key with 2 .. max_occ.
DN_HD uint32_t c_need(const Bucket& b, uint32_t max_occ) {
  return b.key != ~0ull && b.count >= 2 && b.count <= max_occ ? b.count + 1 : 0u;
}

// Entry of bucket b, whose block in pos2 (when it needs one) starts at `at`. `one` is the
// occurrence of a key with count 1.
DN_HD uint64_t c_entry(const Bucket& b, uint32_t max_occ, uint32_t one, uint32_t at) {
  if (b.key == ~0ull) return kCEmpty;
  const uint32_t tag = b.count > max_occ ? kCSkip : b.count == 1 ? kCOne : kCMany;
  const uint32_t val = tag == kCOne ? one : tag == kCMany ? at : 0u;
  return (uint64_t)val << 32 | (uint32_t)(b.key << 2 | tag);
}

// Looks hash h up among the entries of slots home .. home + 3: 1 (ent set) when the key comes
// before any empty slot, 0 when an empty slot comes first, -1 when all four hold other keys.
DN_HD int c_find4(const uint64_t (&e)[4], uint32_t h, uint64_t& ent) {
  int r = -1;
#pragma unroll
  for (int x = 3; x >= 0; --x) {  // the nearest decisive slot is written last
    const uint32_t k = c_key(e[x]);
    const bool empty = k == kCEmpty, hit = !empty && (k >> 2) == h;
    r = empty ? 0 : hit ? 1 : r;
    ent = hit ? e[x] : ent;
  }
  return r;
}

// Continues the probe for h from slot i, one slot at a time (the rare chains longer than four).
// The table's load is at most one half, so an empty slot ends every probe.
DN_HD int c_probe_from(const uint64_t* ctab, uint64_t mask, uint32_t h, uint64_t i, uint64_t& ent) {
  for (;;) {
    const uint64_t e = ctab[i];
    const uint32_t k = c_key(e);
    if (k == kCEmpty) return 0;
    if ((k >> 2) == h) {
      ent = e;
      return 1;
    }
    i = (i + 1) & mask;
  }
}

// The four entries from slot `home`.
DN_HD void c_load4(const uint64_t* ctab, uint64_t home, uint64_t (&e)[4]) {
#pragma unroll
  for (int x = 0; x < 4; ++x) e[x] = ctab[home + x];
}

}  // namespace dn
