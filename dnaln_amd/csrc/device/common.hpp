// common.hpp — types and integer primitives shared by the gfx950 kernels.
//
// The per-lane programs in sw_packed.hpp, sw_generic.hpp and seed_chain.hpp are written against
// the primitives below. Under hipcc's device pass they map to single CDNA4 instructions
// (v_perm_b32, v_pk_*_i16, v_bfi_b32); under a host compiler they are plain C++ with the same
// bit-exact semantics, which lets the CPU test build run the exact kernel logic lane by lane.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define DN_HD __host__ __device__ inline __attribute__((always_inline))
#define DN_DEV __device__ inline __attribute__((always_inline))
#else
#define DN_HD inline __attribute__((always_inline))
#define DN_DEV inline
#endif

// The dynamic LDS array of a kernel (sized at launch), and the waves per SIMD a kernel's
// register allocation must allow.
#if !defined(DN_DYN_LDS)
#define DN_DYN_LDS(T, name) extern __shared__ T name[]
#endif
#if !defined(DN_WAVES_PER_EU)
#define DN_WAVES_PER_EU(n) __attribute__((amdgpu_waves_per_eu(n)))
#endif

namespace dn {

// ------------------------------------------------------------------------------ task descriptors
// One banded alignment. The query is read forward (q[i]) or reverse-complemented
// (comp(q[qlen-1-i])); the target window t[0, qlen+2bw) reads as N outside [tlo, thi).
struct SwTask {
  uint64_t q;      // device address of the query codes (0..4)
  uint64_t t;      // device address of target position 0 (may lie outside the buffer; only [tlo, thi) is read)
  int32_t qlen;
  int32_t tlo, thi;
  uint32_t flags;  // bit 0: reverse-complement the query
};

struct SwResult {
  int32_t score, qb, qe, tb, te, nm;
  int32_t n_cigar;   // ops written (reversed order), -1 on overflow of the per-task slot
  int32_t status;    // 0 ok; nonzero: the walk hit an inconsistent cell (never expected)
};

struct ScoreP {
  int32_t match, mismatch, gap_open, gap_ext, ambig, bw;
};

DN_HD uint8_t comp_code(uint8_t c) { return c < 4 ? (uint8_t)(3 - c) : (uint8_t)4; }

DN_HD int sub_score(const ScoreP& sp, uint8_t a, uint8_t b) {
  return (a > 3 || b > 3) ? -sp.ambig : (a == b ? sp.match : -sp.mismatch);
}

// ------------------------------------------------------------------------------ 32-bit primitives
// v_perm_b32: each selector byte picks a byte of {s0:s1} (s1 = bytes 0..3, s0 = bytes 4..7);
// 8..11 replicate the sign of bytes 1/3/5/7, 12 gives 0x00, 13..255 give 0xFF.
DN_HD uint32_t perm_b32(uint32_t s0, uint32_t s1, uint32_t sel) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_perm(s0, s1, sel);
#else
  const uint64_t in = ((uint64_t)s0 << 32) | s1;
  uint32_t out = 0;
  for (int b = 0; b < 4; ++b) {
    const uint32_t s = (sel >> (8 * b)) & 0xff;
    uint32_t v;
    if (s >= 13) v = 0xff;
    else if (s == 12) v = 0x00;
    else if (s >= 8) v = ((in >> (8 * (2 * (s - 8) + 1) + 7)) & 1) ? 0xff : 0x00;
    else v = (uint32_t)(in >> (8 * s)) & 0xff;
    out |= v << (8 * b);
  }
  return out;
#endif
}

// 16 aligned bytes, loaded or stored with one 16-byte access.
struct alignas(16) Quad {
  uint32_t x, y, z, w;
};

// 16 bytes at the 16-aligned global address a (one global_load_dwordx4; an address held as an
// integer would otherwise be read with a generic flat load).
DN_HD Quad load16(uint64_t a) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef const __attribute__((address_space(1))) Quad* gquad;
  return *(gquad)a;
#else
  return *(const Quad*)(uintptr_t)a;
#endif
}

// Store of data the kernel itself never reads again (traceback flags, read back by the walk
// much later): non-temporal on the device, so it streams past the caches that hold the
// sequences.
DN_HD void store_stream(uint64_t* p, uint64_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}
// 16 bytes at the 16-aligned p, one non-temporal 16-byte store on the device.
DN_HD void store_stream16(uint32_t* p, const Quad& q) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef uint32_t v4u __attribute__((ext_vector_type(4)));
  const v4u v = {q.x, q.y, q.z, q.w};
  __builtin_nontemporal_store(v, reinterpret_cast<v4u*>(p));
#else
  p[0] = q.x;
  p[1] = q.y;
  p[2] = q.z;
  p[3] = q.w;
#endif
}
DN_HD void store_stream(uint32_t* p, uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}

// Packed pairs of int16 in one 32-bit word: lane pair (A = bits 0..15, B = bits 16..31).
#if defined(__HIP_DEVICE_COMPILE__)
typedef short pk16 __attribute__((ext_vector_type(2)));
DN_HD pk16 pk_from(uint32_t x) { return __builtin_bit_cast(pk16, x); }
DN_HD uint32_t pk_bits(pk16 x) { return __builtin_bit_cast(uint32_t, x); }
DN_HD pk16 pk_add(pk16 a, pk16 b) { return a + b; }
DN_HD pk16 pk_sub(pk16 a, pk16 b) { return a - b; }
DN_HD pk16 pk_max(pk16 a, pk16 b) { return __builtin_elementwise_max(a, b); }
// max(a - b, 0) for halves in [0, 32767]: one v_pk_sub_u16 with clamp.
DN_HD pk16 pk_subsat(pk16 a, pk16 b) {
  typedef unsigned short upk16 __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(pk16, __builtin_elementwise_sub_sat(__builtin_bit_cast(upk16, a),
                                                                __builtin_bit_cast(upk16, b)));
}
#else
struct pk16 {
  int16_t x, y;
};
DN_HD pk16 pk_from(uint32_t v) { return pk16{(int16_t)(v & 0xffff), (int16_t)(v >> 16)}; }
DN_HD uint32_t pk_bits(pk16 a) { return (uint32_t)(uint16_t)a.x | ((uint32_t)(uint16_t)a.y << 16); }
DN_HD pk16 pk_add(pk16 a, pk16 b) { return pk16{(int16_t)(a.x + b.x), (int16_t)(a.y + b.y)}; }
DN_HD pk16 pk_sub(pk16 a, pk16 b) { return pk16{(int16_t)(a.x - b.x), (int16_t)(a.y - b.y)}; }
DN_HD pk16 pk_max(pk16 a, pk16 b) { return pk16{a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y}; }
DN_HD pk16 pk_subsat(pk16 a, pk16 b) {
  return pk16{(int16_t)(a.x > b.x ? a.x - b.x : 0), (int16_t)(a.y > b.y ? a.y - b.y : 0)};
}
#endif
DN_HD pk16 pk_splat(int v) { return pk_from((uint32_t)(uint16_t)v * 0x10001u); }
DN_HD int pk_lo(pk16 a) { return (int16_t)(pk_bits(a) & 0xffff); }
DN_HD int pk_hi(pk16 a) { return (int16_t)(pk_bits(a) >> 16); }

// Sign bits of the two halves as 16-bit masks (0xffff where negative).
DN_HD uint32_t pk_signmask(pk16 a) { return perm_b32(0, pk_bits(a), 0x09090808u); }

// Bitfield insert: m ? a : b, bitwise.
DN_HD uint32_t bfi(uint32_t m, uint32_t a, uint32_t b) { return (a & m) | (b & ~m); }

// (a & b) | c as one v_bitop3_b32 on gfx950 (truth table 0xEA over a = 0xF0, b = 0xCC, c = 0xAA),
// where the compiler would otherwise regroup a chain of such updates into ands and or3s.
DN_HD uint32_t and_or(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0xEA);
#else
  return (a & b) | c;
#endif
}

// Byte masks over a word's four bytes: bytes k..3, and bytes 0..k-1 (k clamped to 0..4).
DN_HD uint32_t bytes_from(int k) { return k <= 0 ? ~0u : k >= 4 ? 0u : ~0u << (8 * k); }
DN_HD uint32_t bytes_below(int k) { return ~bytes_from(k); }

}  // namespace dn
