// host_api.cpp — index build, encoding and SAM formatting for the HIP path (include/dnaln_host.h).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "device/seed_chain.hpp"
#include "dnaln_host.h"
#include "host/c_guard.hpp"
#include "host/sam_text.hpp"

namespace {

uint8_t nt4(unsigned char c) {  // src/common.hpp:23-31
  switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': case 'U': case 'u': return 3;
    default: return 4;
  }
}

struct Occ {
  uint64_t hash;
  uint32_t pos;  // pos << 1 | strand
};

}  // namespace

extern "C" {

int dnaln_index_build(const uint8_t* ref, const dnaln_contig* contigs, int32_t n_contigs, int32_t k, int32_t w,
                      dnaln_index* out) {
  return dnaln_amd::c_guard([&]() -> int {
    if (!out || !contigs || n_contigs < 1 || k < 4 || k > 28 || w < 1 || w > 255) return DNALN_E_INVALID;
    std::memset(out, 0, sizeof *out);
    uint64_t total = 0;
    for (int c = 0; c < n_contigs; ++c) total = std::max<uint64_t>(total, contigs[c].offset + contigs[c].len);
    if (total >= (1ull << 31)) return DNALN_E_INVALID;
    std::vector<Occ> occ;
    occ.reserve(total * 2 / (w + 1) + 1024);
    std::vector<uint64_t> rh(w);
    std::vector<uint32_t> rps(w);
    dn::Ring ring{rh.data(), rps.data(), 1};
    for (int c = 0; c < n_contigs; ++c) {
      const uint32_t base = (uint32_t)contigs[c].offset;
      dn::sketch_stream(dn::PtrReader{ref + contigs[c].offset}, (int)contigs[c].len, k, w, ring,
                        [&](uint64_t h, uint32_t p, uint32_t s) { occ.push_back({h, (p + base) << 1 | s}); });
    }
    std::sort(occ.begin(), occ.end(),
              [](const Occ& a, const Occ& b) { return a.hash != b.hash ? a.hash < b.hash : a.pos < b.pos; });
    uint64_t n_keys = 0;
    for (size_t i = 0; i < occ.size(); ++i) n_keys += i == 0 || occ[i].hash != occ[i - 1].hash;
    uint64_t cap = 16;
    while (cap < n_keys * 2) cap <<= 1;
    std::unique_ptr<dnaln_bucket[]> table(new dnaln_bucket[cap]);
    for (uint64_t s = 0; s < cap; ++s) table[s] = {~0ULL, 0, 0};
    std::unique_ptr<uint32_t[]> pos(new uint32_t[occ.size() + 1]);
    const uint64_t mask = cap - 1;
    uint32_t max_count = 0;
    for (size_t i = 0; i < occ.size();) {
      size_t j = i;
      for (; j < occ.size() && occ[j].hash == occ[i].hash; ++j) pos[j] = occ[j].pos;
      uint64_t s = occ[i].hash & mask;
      while (table[s].key != ~0ULL) s = (s + 1) & mask;
      table[s] = {occ[i].hash, (uint32_t)i, (uint32_t)(j - i)};
      max_count = std::max(max_count, (uint32_t)(j - i));
      i = j;
    }
    out->view.table = table.release();
    out->view.n_buckets = cap;
    out->view.pos = pos.release();
    out->view.n_pos = occ.size();
    out->view.k = k;
    out->view.w = w;
    out->n_keys = n_keys;
    out->max_count = max_count;
    return DNALN_OK;
  });
}

void dnaln_index_free(dnaln_index* idx) {
  if (!idx) return;
  delete[] idx->view.table;
  delete[] idx->view.pos;
  std::memset(idx, 0, sizeof *idx);
}

int dnaln_index_build_gpu(const uint8_t* ref, const dnaln_contig* contigs, int32_t n_contigs, int32_t k, int32_t w,
                          int32_t device, dnaln_index* out, double* build_ms) {
  return dnaln_amd::c_guard([&]() -> int {
    if (!out || !contigs || n_contigs < 1) return DNALN_E_INVALID;
    std::memset(out, 0, sizeof *out);
    uint64_t total = 0;
    for (int c = 0; c < n_contigs; ++c) total = std::max<uint64_t>(total, contigs[c].offset + contigs[c].len);
    dnaln_hip_index* ix = nullptr;
    int rc = dnaln_hip_index_build(ref, total, contigs, n_contigs, k, w, device, &ix);
    if (rc != DNALN_OK) return rc;
    uint64_t nb = 0, np = 0, nk = 0;
    uint32_t mc = 0;
    double ms = 0;
    dnaln_hip_index_info(ix, &nb, &np, &nk, &mc, &ms);
    std::unique_ptr<dnaln_bucket[]> table;
    std::unique_ptr<uint32_t[]> pos;
    try {
      table.reset(new dnaln_bucket[nb]);
      pos.reset(new uint32_t[np + 1]);
    } catch (...) {
      dnaln_hip_index_free(ix);
      throw;
    }
    rc = dnaln_hip_index_download(ix, table.get(), pos.get());
    dnaln_hip_index_free(ix);
    if (rc != DNALN_OK) return rc;
    out->view.table = table.release();
    out->view.n_buckets = nb;
    out->view.pos = pos.release();
    out->view.n_pos = np;
    out->view.k = k;
    out->view.w = w;
    out->n_keys = nk;
    out->max_count = mc;
    if (build_ms) *build_ms = ms;
    return DNALN_OK;
  });
}

// A 64-bit multiply-xorshift digest over 8-byte words (not cryptographic: it catches an index
// file paired with other codes or contigs, not a forgery).
uint64_t dnaln_reference_digest(const uint8_t* ref, const dnaln_contig* contigs, int32_t n_contigs) {
  uint64_t h = 0x243f6a8885a308d3ull ^ (uint64_t)n_contigs;
  auto mix = [&h](uint64_t v) {
    h ^= v;
    h *= 0x9fb21c651e98df25ull;
    h ^= h >> 29;
  };
  for (int c = 0; c < n_contigs; ++c) {
    mix(contigs[c].offset);
    mix(contigs[c].len);
  }
  for (int c = 0; c < n_contigs; ++c) {
    const uint8_t* p = ref + contigs[c].offset;
    const uint64_t n = contigs[c].len;
    uint64_t i = 0;
    for (; i + 8 <= n; i += 8) {
      uint64_t v;
      std::memcpy(&v, p + i, 8);
      mix(v);
    }
    uint64_t tail = 0;
    for (uint64_t j = 0; i + j < n; ++j) tail |= (uint64_t)p[i + j] << (8 * j);
    mix(tail ^ (n - i) << 56);
  }
  return h;
}

namespace {

struct IndexFileHeader {
  char magic[8];
  uint32_t version;
  int32_t k, w;
  uint32_t max_count;
  uint64_t n_buckets, n_pos, n_keys, digest, reserved;
};
static_assert(sizeof(IndexFileHeader) == 64, "index file header is 64 bytes");
constexpr char kIndexMagic[8] = {'D', 'N', 'A', 'L', 'N', 'I', 'X', '1'};

struct File {
  std::FILE* f;
  ~File() {
    if (f) std::fclose(f);
  }
};

}  // namespace

int dnaln_index_save(const dnaln_index* idx, const uint8_t* ref, const dnaln_contig* contigs, int32_t n_contigs,
                     const char* path) {
  if (!idx || !path || !contigs || n_contigs < 1 || !idx->view.table) return DNALN_E_INVALID;
  IndexFileHeader h{};
  std::memcpy(h.magic, kIndexMagic, 8);
  h.version = 1;
  h.k = idx->view.k;
  h.w = idx->view.w;
  h.max_count = idx->max_count;
  h.n_buckets = idx->view.n_buckets;
  h.n_pos = idx->view.n_pos;
  h.n_keys = idx->n_keys;
  h.digest = dnaln_reference_digest(ref, contigs, n_contigs);
  File f{std::fopen(path, "wb")};
  if (!f.f) return DNALN_E_IO;
  if (std::fwrite(&h, sizeof h, 1, f.f) != 1) return DNALN_E_IO;
  if (std::fwrite(idx->view.table, sizeof(dnaln_bucket), h.n_buckets, f.f) != h.n_buckets) return DNALN_E_IO;
  if (h.n_pos && std::fwrite(idx->view.pos, 4, h.n_pos, f.f) != h.n_pos) return DNALN_E_IO;
  if (std::fflush(f.f) != 0) return DNALN_E_IO;
  return DNALN_OK;
}

int dnaln_index_load(const char* path, const uint8_t* ref, const dnaln_contig* contigs, int32_t n_contigs,
                     dnaln_index* out) {
  return dnaln_amd::c_guard([&]() -> int {
    if (!out || !path || !contigs || n_contigs < 1) return DNALN_E_INVALID;
    std::memset(out, 0, sizeof *out);
    File f{std::fopen(path, "rb")};
    if (!f.f) return DNALN_E_IO;
    IndexFileHeader h{};
    if (std::fread(&h, sizeof h, 1, f.f) != 1 || std::memcmp(h.magic, kIndexMagic, 8) != 0 || h.version != 1)
      return DNALN_E_IO;
    if (h.k < 4 || h.k > 28 || h.w < 1 || h.w > 255 || h.n_buckets < 16 || (h.n_buckets & (h.n_buckets - 1)) ||
        h.n_buckets > (1ull << 40) || h.n_keys * 2 > h.n_buckets || h.n_pos >= (1ull << 32) || h.n_keys > h.n_pos)
      return DNALN_E_IO;
    if (std::fseek(f.f, 0, SEEK_END) != 0) return DNALN_E_IO;
    const long long size = std::ftell(f.f);
    if (size < 0 || (uint64_t)size != sizeof h + h.n_buckets * sizeof(dnaln_bucket) + h.n_pos * 4) return DNALN_E_IO;
    if (h.digest != dnaln_reference_digest(ref, contigs, n_contigs)) return DNALN_E_IO;
    if (std::fseek(f.f, (long)sizeof h, SEEK_SET) != 0) return DNALN_E_IO;
    std::unique_ptr<dnaln_bucket[]> table(new dnaln_bucket[h.n_buckets]);
    std::unique_ptr<uint32_t[]> pos(new uint32_t[h.n_pos + 1]);
    pos[h.n_pos] = 0;
    if (std::fread(table.get(), sizeof(dnaln_bucket), h.n_buckets, f.f) != h.n_buckets) return DNALN_E_IO;
    if (h.n_pos && std::fread(pos.get(), 4, h.n_pos, f.f) != h.n_pos) return DNALN_E_IO;
    // What the device kernels read stays in bounds: every slice inside pos[], every position
    // inside the reference, at least half the buckets empty (so every probe ends); the stated key
    // count and largest count match the table.
    uint64_t total = 0;
    for (int c = 0; c < n_contigs; ++c) total = std::max<uint64_t>(total, contigs[c].offset + contigs[c].len);
    uint64_t keys = 0;
    uint32_t max_count = 0;
    for (uint64_t i = 0; i < h.n_buckets; ++i) {
      const dnaln_bucket& b = table[i];
      if (b.key == ~0ULL) continue;
      if (b.count == 0 || (uint64_t)b.offset + b.count > h.n_pos) return DNALN_E_IO;
      ++keys;
      max_count = std::max(max_count, b.count);
    }
    if (keys != h.n_keys || max_count != h.max_count) return DNALN_E_IO;
    for (uint64_t i = 0; i < h.n_pos; ++i)
      if ((pos[i] >> 1) >= total) return DNALN_E_IO;
    out->view.table = table.release();
    out->view.pos = pos.release();
    out->view.n_buckets = h.n_buckets;
    out->view.n_pos = h.n_pos;
    out->view.k = h.k;
    out->view.w = h.w;
    out->n_keys = h.n_keys;
    out->max_count = h.max_count;
    return DNALN_OK;
  });
}

uint32_t dnaln_fnv1a(const char* s, size_t n) {
  uint32_t h = 2166136261u;
  for (size_t i = 0; i < n; ++i) h = (h ^ (unsigned char)s[i]) * 16777619u;
  return h;
}

void dnaln_encode(const char* s, size_t n, uint8_t* out) {
  for (size_t i = 0; i < n; ++i) out[i] = nt4((unsigned char)s[i]);
}

long dnaln_sam_format(const char* name, size_t name_len, const char* seq, const char* qual, size_t len,
                      const dnaln_record* m, const uint32_t* cigar_pool, const char* contig_name, char* buf,
                      size_t cap) {
  const size_t clen = contig_name ? std::strlen(contig_name) : 0;
  if (dnaln_amd::sam_bound(name_len, len, m->cigar_len, clen) > cap) return -1;
  return (long)(dnaln_amd::sam_record(buf, name, name_len, seq, qual, len, *m, cigar_pool, contig_name, clen) - buf);
}

}  // extern "C"

extern "C" {

// FNV-1a of n names packed back to back (names + offs[i], lens[i]).
void dnaln_fnv1a_many(const char* names, const uint64_t* offs, const int32_t* lens, uint64_t n, uint32_t* out) {
  for (uint64_t i = 0; i < n; ++i) out[i] = dnaln_fnv1a(names + offs[i], (size_t)lens[i]);
}

}  // extern "C"

// The byte model of the seed stage: every read sketched (src/minimizer.cpp:5-68), every minimizer
// looked up by linear probing (src/index.hpp:24-38), the occurrences of keys that pass max_occ
// read as the reference's anchor loop reads them (src/mapper.cpp:24-40).
int dnaln_seed_traffic(const dnaln_index_view* idx, const uint8_t* codes, const uint64_t* offs, const int32_t* lens,
                       uint64_t n, uint32_t max_occ, int32_t threads, dnaln_seed_traffic_counts* out) {
  return dnaln_amd::c_guard([&]() -> int {
    if (!idx || !out || (n && (!codes || !offs || !lens)) || idx->n_buckets == 0 ||
        (idx->n_buckets & (idx->n_buckets - 1)) || idx->w < 1 || idx->w > 255)
      return DNALN_E_INVALID;
    const int nt = std::max(1, std::min(threads > 0 ? threads : 8, 64));
    std::vector<dnaln_seed_traffic_counts> part((size_t)nt);
    std::vector<std::thread> th;
    const uint64_t mask = idx->n_buckets - 1;
    for (int t = 0; t < nt; ++t)
      th.emplace_back([&, t] {
        dnaln_seed_traffic_counts c{};
        std::vector<uint64_t> rh((size_t)idx->w);
        std::vector<uint32_t> rps((size_t)idx->w);
        dn::Ring ring{rh.data(), rps.data(), 1};
        for (uint64_t r = n * t / nt; r < n * (t + 1) / nt; ++r) {
          c.read_bases += (uint64_t)std::max(lens[r], 0);
          dn::sketch_stream(dn::PtrReader{codes + offs[r]}, lens[r], idx->k, idx->w, ring,
                            [&](uint64_t h, uint32_t, uint32_t) {
                              ++c.minimizers;
                              for (uint64_t i = h & mask;; i = (i + 1) & mask) {
                                ++c.buckets;
                                const dnaln_bucket& b = idx->table[i];
                                if (b.key == h) {
                                  if (b.count >= 1 && b.count <= max_occ) c.occurrences += b.count;
                                  break;
                                }
                                if (b.key == ~0ULL) break;
                              }
                            });
        }
        part[(size_t)t] = c;
      });
    for (auto& x : th) x.join();
    std::memset(out, 0, sizeof *out);
    for (const auto& c : part) {
      out->read_bases += c.read_bases;
      out->minimizers += c.minimizers;
      out->buckets += c.buckets;
      out->occurrences += c.occurrences;
    }
    out->bytes = out->read_bases + 16 * out->buckets + 4 * out->occurrences;
    return DNALN_OK;
  });
}
