/* oracle.c — CPU restatement of dnaln's seed-chain-extend path (checker only; see oracle.h).
 *
 * Written for clarity, not speed: the sketch evaluates every window directly, the aligner is
 * the scalar int32 recurrence, and the mapper handles one read at a time (the reference's
 * per-read results do not depend on the batch, src/mapper.cpp:75-223).
 */
#include "oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void or_default_params(or_map_params_t* p) {
  /* src/mapper.hpp:25-35, src/align.hpp:34-40, src/chain.hpp:30-36 */
  p->k = 15;
  p->w = 10;
  p->max_occ = 500;
  p->sc.match = 2;
  p->sc.mismatch = 4;
  p->sc.gap_open = 4;
  p->sc.gap_ext = 2;
  p->sc.ambig = 1;
  p->sc.bw = 16;
  p->ch.max_gap = 5000;
  p->ch.band = 16;
  p->ch.lookback = 50;
  p->ch.min_score = 20;
  p->ch.min_anchors = 2;
  p->max_cand = 10;
  p->pri_ratio = 0.5;
  p->min_aln_score = 40;
}

/* ------------------------------------------------------------------ encoding (common.hpp:23-42) */
static uint8_t nt4(unsigned char c) {
  switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': case 'U': case 'u': return 3;
    default: return 4;
  }
}

void or_encode(const char* s, size_t n, uint8_t* out) {
  for (size_t i = 0; i < n; ++i) out[i] = nt4((unsigned char)s[i]);
}

static uint8_t comp_code(uint8_t c) { return c < 4 ? (uint8_t)(3 - c) : 4; }

void or_revcomp(const uint8_t* in, size_t n, uint8_t* out) {
  for (size_t i = 0; i < n; ++i) out[i] = comp_code(in[n - 1 - i]);
}

/* ------------------------------------------------------------------ hash (minimizer.hpp:21-30) */
uint64_t or_hash64(uint64_t key, uint64_t mask) {
  key = (~key + (key << 21)) & mask;
  key = key ^ key >> 24;
  key = ((key + (key << 3)) + (key << 8)) & mask;
  key = key ^ key >> 14;
  key = ((key + (key << 2)) + (key << 4)) & mask;
  key = key ^ key >> 28;
  key = (key + (key << 31)) & mask;
  return key;
}

/* common.hpp:136-140 */
uint32_t or_fnv1a(const char* s, size_t n) {
  uint32_t h = 2166136261u;
  for (size_t i = 0; i < n; ++i) h = (h ^ (unsigned char)s[i]) * 16777619u;
  return h;
}

/* ------------------------------------------------------------------ sketch (minimizer.cpp:5-68)
 * The reference streams a van Herk/Gil-Werman window minimum; its result is, for every window
 * of w consecutive k-mers inside one N-free run, the leftmost k-mer of smallest canonical hash
 * (palindromes carry hash ~0 and are never chosen), emitted unless its position equals the
 * previous valid window's choice. This is evaluated directly, window by window. */
size_t or_sketch(const uint8_t* seq, uint32_t len, int k, int w, uint32_t offset, or_mini_t* out) {
  const uint64_t mask = (1ULL << (2 * k)) - 1;
  const int shift = 2 * (k - 1);
  const uint64_t kInvalid = ~0ULL;
  if (len == 0) return 0;
  uint64_t* h = (uint64_t*)malloc(sizeof(uint64_t) * len);
  uint32_t* st = (uint32_t*)malloc(sizeof(uint32_t) * len);
  uint32_t* pos = (uint32_t*)malloc(sizeof(uint32_t) * len);
  size_t nout = 0;
  int64_t last = -1;
  uint32_t i = 0;
  while (i < len) {
    if (seq[i] > 3) {
      ++i;
      continue;
    }
    uint32_t j = i;
    while (j < len && seq[j] <= 3) ++j;
    /* k-mers of the run [i, j) */
    uint32_t nk = 0;
    uint64_t fwd = 0, rev = 0;
    for (uint32_t x = i; x < j; ++x) {
      const uint8_t c = seq[x];
      fwd = ((fwd << 2) | c) & mask;
      rev = (rev >> 2) | ((uint64_t)(3 - c) << shift);
      if (x - i + 1 < (uint32_t)k) continue;
      const uint32_t strand = rev < fwd;
      h[nk] = fwd != rev ? or_hash64(strand ? rev : fwd, mask) : kInvalid;
      st[nk] = strand;
      pos[nk] = x - (uint32_t)k + 1;
      ++nk;
    }
    for (uint32_t ws = 0; ws + (uint32_t)w <= nk; ++ws) {
      uint32_t b = ws;
      for (uint32_t x = ws + 1; x < ws + (uint32_t)w; ++x)
        if (h[x] < h[b]) b = x;
      if (h[b] == kInvalid) continue;
      if ((int64_t)pos[b] != last) {
        out[nout].hash = h[b];
        out[nout].pos = pos[b] + offset;
        out[nout].strand = st[b];
        ++nout;
      }
      last = pos[b];
    }
    i = j;
  }
  free(h);
  free(st);
  free(pos);
  return nout;
}

/* ------------------------------------------------------------------ index (index.cpp:10-45) */
static int cmp_mini(const void* x, const void* y) {
  const or_mini_t* a = (const or_mini_t*)x;
  const or_mini_t* b = (const or_mini_t*)y;
  if (a->hash != b->hash) return a->hash < b->hash ? -1 : 1;
  return a->pos < b->pos ? -1 : (a->pos > b->pos);
}

int or_index_build(or_index_t* idx, const uint8_t* seq, const uint64_t* contig_off, const uint32_t* contig_len,
                   int n_contigs, int k, int w) {
  memset(idx, 0, sizeof *idx);
  if (k < 4 || k > 28 || w < 1 || w > 255) return -1;
  idx->k = k;
  idx->w = w;
  uint64_t total = 0;
  for (int c = 0; c < n_contigs; ++c) total += contig_len[c];
  or_mini_t* mins = (or_mini_t*)malloc(sizeof(or_mini_t) * (total + 1));
  size_t n = 0;
  for (int c = 0; c < n_contigs; ++c)
    n += or_sketch(seq + contig_off[c], contig_len[c], k, w, (uint32_t)contig_off[c], mins + n);
  qsort(mins, n, sizeof(or_mini_t), cmp_mini);
  uint64_t n_keys = 0;
  for (size_t i = 0; i < n; ++i)
    if (i == 0 || mins[i].hash != mins[i - 1].hash) ++n_keys;
  uint64_t cap = 16;
  while (cap < n_keys * 2) cap <<= 1;
  idx->table = (or_bucket_t*)malloc(sizeof(or_bucket_t) * cap);
  for (uint64_t s = 0; s < cap; ++s) {
    idx->table[s].key = ~0ULL;
    idx->table[s].offset = idx->table[s].count = 0;
  }
  idx->mask = cap - 1;
  idx->cap = cap;
  idx->pos = (uint32_t*)malloc(sizeof(uint32_t) * (n + 1));
  idx->n_pos = n;
  idx->n_keys = n_keys;
  for (size_t i = 0; i < n;) {
    size_t j = i;
    for (; j < n && mins[j].hash == mins[i].hash; ++j) idx->pos[j] = mins[j].pos << 1 | mins[j].strand;
    uint64_t s = mins[i].hash & idx->mask;
    while (idx->table[s].key != ~0ULL) s = (s + 1) & idx->mask;
    idx->table[s].key = mins[i].hash;
    idx->table[s].offset = (uint32_t)i;
    idx->table[s].count = (uint32_t)(j - i);
    if ((uint32_t)(j - i) > idx->max_count) idx->max_count = (uint32_t)(j - i);
    i = j;
  }
  free(mins);
  return 0;
}

/* index.hpp:24-38 */
const uint32_t* or_index_lookup(const or_index_t* idx, uint64_t h, uint32_t* n) {
  uint64_t i = h & idx->mask;
  for (;;) {
    const or_bucket_t* b = &idx->table[i];
    if (b->key == h) {
      *n = b->count;
      return idx->pos + b->offset;
    }
    if (b->key == ~0ULL) {
      *n = 0;
      return NULL;
    }
    i = (i + 1) & idx->mask;
  }
}

void or_index_free(or_index_t* idx) {
  free(idx->table);
  free(idx->pos);
  memset(idx, 0, sizeof *idx);
}

/* ------------------------------------------------------------------ chaining (chain.cpp:12-71) */
static int cmp_anchor(const void* x, const void* y) {
  const or_anchor_t* a = (const or_anchor_t*)x;
  const or_anchor_t* b = (const or_anchor_t*)y;
  if (a->strand != b->strand) return a->strand < b->strand ? -1 : 1;
  if (a->rpos != b->rpos) return a->rpos < b->rpos ? -1 : 1;
  return a->qpos < b->qpos ? -1 : (a->qpos > b->qpos);
}

static int ilog2_u64(uint64_t x) { return 63 - __builtin_clzll(x); }

typedef struct {
  int32_t f;
  int32_t idx;
} fi_t;

static int cmp_f_desc(const void* x, const void* y) {
  const fi_t* a = (const fi_t*)x;
  const fi_t* b = (const fi_t*)y;
  if (a->f != b->f) return a->f > b->f ? -1 : 1;
  return a->idx < b->idx ? -1 : (a->idx > b->idx);
}

typedef struct {
  or_chain_t c;
  size_t idx;
} chain_ix_t;

static int cmp_chain_stable(const void* x, const void* y) {
  const chain_ix_t* a = (const chain_ix_t*)x;
  const chain_ix_t* b = (const chain_ix_t*)y;
  if (a->c.score != b->c.score) return a->c.score > b->c.score ? -1 : 1;
  return a->idx < b->idx ? -1 : (a->idx > b->idx);
}

size_t or_chain(or_anchor_t* a, size_t n_, int k, const or_chain_params_t* p, or_chain_t* out) {
  if (n_ == 0) return 0;
  const int n = (int)n_;
  qsort(a, n_, sizeof(or_anchor_t), cmp_anchor);
  int32_t* f = (int32_t*)malloc(sizeof(int32_t) * n_);
  int32_t* prev = (int32_t*)malloc(sizeof(int32_t) * n_);
  for (int i = 0; i < n; ++i) {
    int best = k, best_j = -1;
    const int lo = i - p->lookback > 0 ? i - p->lookback : 0;
    for (int j = i - 1; j >= lo; --j) {
      if (a[j].strand != a[i].strand) break;
      const int64_t dr = (int64_t)a[i].rpos - a[j].rpos;
      if (dr > p->max_gap) break;
      const int64_t dq = (int64_t)a[i].qpos - a[j].qpos;
      if (dr <= 0 || dq <= 0) continue;
      const int64_t dd = dr > dq ? dr - dq : dq - dr;
      if (dd > p->band) continue;
      const int64_t m = dq < dr ? dq : dr;
      const int gain = (int)(m < k ? m : k);
      const int cost = dd ? (int)(0.01 * k * (double)dd + 0.5 * ilog2_u64((uint64_t)dd)) : 0;
      const int sc = f[j] + gain - cost;
      if (sc > best) {
        best = sc;
        best_j = j;
      }
    }
    f[i] = best;
    prev[i] = best_j;
  }
  fi_t* order = (fi_t*)malloc(sizeof(fi_t) * n_);
  for (int i = 0; i < n; ++i) {
    order[i].f = f[i];
    order[i].idx = i;
  }
  qsort(order, n_, sizeof(fi_t), cmp_f_desc);
  uint8_t* used = (uint8_t*)calloc(n_, 1);
  chain_ix_t* tmp = (chain_ix_t*)malloc(sizeof(chain_ix_t) * n_);
  size_t m = 0;
  for (int o = 0; o < n; ++o) {
    const int end = order[o].idx;
    if (used[end]) continue;
    int i = end, cnt = 0;
    int64_t dmin = INT64_MAX, dmax = INT64_MIN;
    while (i >= 0 && !used[i]) {
      used[i] = 1;
      ++cnt;
      const int64_t d = (int64_t)a[i].rpos - a[i].qpos;
      if (d < dmin) dmin = d;
      if (d > dmax) dmax = d;
      i = prev[i];
    }
    const int score = f[end] - (i >= 0 ? f[i] : 0);
    if (score < p->min_score || cnt < p->min_anchors) continue;
    tmp[m].c.score = score;
    tmp[m].c.strand = a[end].strand;
    tmp[m].c.n_anchors = cnt;
    tmp[m].c.dmin = dmin;
    tmp[m].c.dmax = dmax;
    tmp[m].c.rpos = a[end].rpos;
    tmp[m].idx = m;
    ++m;
  }
  qsort(tmp, m, sizeof(chain_ix_t), cmp_chain_stable); /* std::stable_sort by score desc */
  for (size_t c = 0; c < m; ++c) out[c] = tmp[c].c;
  free(f);
  free(prev);
  free(order);
  free(used);
  free(tmp);
  return m;
}

/* ------------------------------------------------------------------ banded SW (align.cpp:31-143) */
static int sub_score(const or_score_t* sp, uint8_t a, uint8_t b) {
  return (a > 3 || b > 3) ? -sp->ambig : (a == b ? sp->match : -sp->mismatch);
}

typedef struct {
  uint32_t* c;
  int n, cap, overflow;
} cig_t;

static void cig_push(cig_t* c, uint32_t op, int len) {
  if (c->n > 0 && (c->c[c->n - 1] & 0xf) == op) {
    c->c[c->n - 1] += (uint32_t)len << 4;
    return;
  }
  if (c->n >= c->cap) {
    c->overflow = 1;
    return;
  }
  c->c[c->n++] = (uint32_t)len << 4 | op;
}

/* align.cpp:87-143 — walks the direction bytes back from the best cell. */
static void traceback(const uint8_t* dir, int W, int i_end, int d_end, int score, const uint8_t* q,
                      const uint8_t* t, const or_score_t* sp, or_result_t* r, cig_t* c) {
  const int oe = sp->gap_open + sp->gap_ext, ge = sp->gap_ext;
  int i = i_end, d = d_end, v = score, nm = 0;
  for (;;) {
    int run = 0;
    uint8_t x = 0;
    while (i >= 0 && v > 0 && ((x = dir[(size_t)i * W + d]) & 0x80)) {
      const uint8_t a = q[i], b = t[i + d];
      v -= sub_score(sp, a, b);
      nm += a != b || a > 3;
      ++run;
      --i;
    }
    if (run) cig_push(c, 0 /*M*/, run);
    if (i < 0 || v <= 0) break;
    run = 0;
    if (x & 0x40) { /* insertion: E(i,d) <- (i-1, d+1) */
      int ext;
      do {
        ext = (dir[(size_t)i * W + d] & 0x20) != 0;
        v += ext ? ge : oe;
        ++run;
        --i;
        ++d;
      } while (ext);
      cig_push(c, 1 /*I*/, run);
    } else { /* deletion: F(i,d) <- (i, d-1) */
      int ext;
      do {
        ext = (dir[(size_t)i * W + d] & 0x10) != 0;
        v += ext ? ge : oe;
        ++run;
        --d;
      } while (ext);
      cig_push(c, 2 /*D*/, run);
    }
    nm += run;
  }
  for (int a = 0, b = c->n - 1; a < b; ++a, --b) {
    const uint32_t s = c->c[a];
    c->c[a] = c->c[b];
    c->c[b] = s;
  }
  r->qb = i + 1;
  r->tb = i + 1 + d;
  r->qe = i_end + 1;
  r->te = i_end + d_end + 1;
  r->nm = nm;
}

int or_align(const or_score_t* sp, const uint8_t* q, int ql, const uint8_t* tin, int tlen, int tb, or_result_t* r,
             uint32_t* cigar, int cigar_cap) {
  const int W = 2 * sp->bw + 1, tmax = ql + W - 1;
  const int oe = sp->gap_open + sp->gap_ext, ge = sp->gap_ext;
  uint8_t* t = (uint8_t*)malloc((size_t)(tmax > 0 ? tmax : 1));
  const int tl = tlen < tmax ? tlen : tmax;
  if (tl > 0) memcpy(t, tin, (size_t)tl);
  for (int j = tl < 0 ? 0 : tl; j < tmax; ++j) t[j] = 4;
  int32_t* H = (int32_t*)calloc((size_t)W + 1, sizeof(int32_t));
  int32_t* E = (int32_t*)calloc((size_t)W + 1, sizeof(int32_t));
  uint8_t* D = tb ? (uint8_t*)malloc((size_t)(ql > 0 ? ql : 1) * (size_t)W) : NULL;
  int best = 0, best_i = -1, best_d = -1;
  for (int i = 0; i < ql; ++i) {
    int f = 0, hleft = 0, hdiag = H[0];
    for (int d = 0; d < W; ++d) {
      const int hup = H[d + 1], eup = E[d + 1];
      const int hd = hdiag + sub_score(sp, q[i], t[i + d]);
      const int eo = hup - oe, ee = eup - ge;
      const int e = eo > ee ? eo : ee;
      const int fo = hleft - oe, fe = f - ge;
      f = fo > fe ? fo : fe;
      int h = hd > e ? hd : e;
      h = h > 0 ? h : 0;
      h = h > f ? h : f;
      H[d] = h;
      E[d] = e;
      if (tb) {
        D[(size_t)i * W + d] =
            (uint8_t)((h == hd ? 0x80 : 0) | (h == e ? 0x40 : 0) | (ee > eo ? 0x20 : 0) | (fe > fo ? 0x1F : 0));
        if (h > best) {
          best = h;
          best_i = i;
          best_d = d;
        }
      } else {
        best = h > best ? h : best;
      }
      hleft = h;
      hdiag = hup;
    }
  }
  memset(r, 0, sizeof *r);
  r->score = best;
  int rc = 0;
  if (tb && best > 0) {
    cig_t c = {cigar, 0, cigar_cap, 0};
    traceback(D, W, best_i, best_d, best, q, tin, sp, r, &c);
    r->n_cigar = c.overflow ? -1 : c.n;
    rc = c.overflow ? -1 : 0;
  }
  free(t);
  free(H);
  free(E);
  free(D);
  return rc;
}

/* align.cpp:145-162 */
int or_score_cigar(const or_score_t* sp, const uint8_t* q, const uint8_t* t, int qb, int tb, const uint32_t* cigar,
                   int n_cigar) {
  int s = 0, i = qb, j = tb;
  for (int c = 0; c < n_cigar; ++c) {
    const uint32_t op = cigar[c] & 0xf;
    const int len = (int)(cigar[c] >> 4);
    if (op == 0 || op == 7 || op == 8) {
      for (int x = 0; x < len; ++x) s += sub_score(sp, q[i++], t[j++]);
    } else if (op == 1) {
      s -= sp->gap_open + len * sp->gap_ext;
      i += len;
    } else if (op == 2) {
      s -= sp->gap_open + len * sp->gap_ext;
      j += len;
    }
  }
  return s;
}

/* ------------------------------------------------------------------ mapper (mapper.cpp:8-223) */
typedef struct {
  uint32_t rev;
  int32_t contig;
  int64_t diag, wstart;
  int32_t sw;
  int32_t chain_score; /* score of the chain the candidate came from */
} cand_t;

static int contig_of(const uint64_t* off, int nc, uint64_t gpos) { /* reference.cpp:25-32 */
  size_t lo = 0, hi = (size_t)nc;
  while (hi - lo > 1) {
    const size_t mid = (lo + hi) / 2;
    if (off[mid] <= gpos) lo = mid;
    else hi = mid;
  }
  return (int)lo;
}

/* reference.cpp:40-47: window [gstart, gstart+len), N outside contig ci */
static void fetch_window(const uint8_t* ref, const uint64_t* off, const uint32_t* clen, int ci, int64_t gstart,
                         int len, uint8_t* out) {
  const int64_t b = (int64_t)off[ci], e = b + clen[ci];
  for (int i = 0; i < len; ++i) {
    const int64_t p = gstart + i;
    out[i] = (p >= b && p < e) ? ref[p] : 4;
  }
}

/* seed_and_chain (mapper.cpp:14-63) for one read (forward codes): sketch, lookups, anchors,
 * Chainer::run, candidates. Writes at most p->max_cand candidates; returns their number. The
 * anchor and chain arrays grow as needed (capacity *acap). */
static int seed_and_chain(const or_map_params_t* pp, const or_index_t* idx, const uint64_t* contig_off, int n_contigs,
                          const uint8_t* fwd, int len, or_mini_t* mins, or_anchor_t** anchors_p,
                          or_chain_t** chains_p, size_t* acap_p, cand_t* cands) {
  const or_map_params_t p = *pp;
  const int k = idx->k, bw = p.sc.bw;
  size_t acap = *acap_p;
  or_anchor_t* anchors = *anchors_p;
  or_chain_t* chains = *chains_p;
  const size_t nm = or_sketch(fwd, (uint32_t)len, k, idx->w, 0, mins);
  size_t na = 0;
  for (size_t i = 0; i < nm; ++i) {
    uint32_t cnt;
    const uint32_t* pos = or_index_lookup(idx, mins[i].hash, &cnt);
    if (cnt == 0 || cnt > p.max_occ) continue;
    if (na + cnt > acap) {
      while (na + cnt > acap) acap *= 2;
      anchors = (or_anchor_t*)realloc(anchors, sizeof(or_anchor_t) * acap);
      chains = (or_chain_t*)realloc(chains, sizeof(or_chain_t) * acap);
    }
    for (uint32_t j = 0; j < cnt; ++j) {
      const uint32_t v = pos[j];
      const uint32_t rev = (v & 1) ^ mins[i].strand;
      anchors[na].rpos = v >> 1;
      anchors[na].qpos = rev ? len - (int32_t)mins[i].pos - k : (int32_t)mins[i].pos;
      anchors[na].strand = rev;
      ++na;
    }
  }
  int nc = 0;
  const size_t nch = or_chain(anchors, na, k, &p.ch, chains);
  if (nch > 0) {
    const double min_score = chains[0].score * p.pri_ratio;
    for (size_t c = 0; c < nch; ++c) {
      if (chains[c].score < min_score || nc >= p.max_cand) break;
      const int64_t diag = chains[c].dmin + (chains[c].dmax - chains[c].dmin) / 2;
      int dup = 0;
      for (int i = 0; i < nc && !dup; ++i) {
        const int64_t dd = cands[i].diag - diag;
        dup = cands[i].rev == chains[c].strand && (dd < 0 ? -dd : dd) <= bw;
      }
      if (dup) continue;
      cands[nc].rev = chains[c].strand;
      cands[nc].contig = contig_of(contig_off, n_contigs, chains[c].rpos);
      cands[nc].diag = diag;
      cands[nc].wstart = diag - bw;
      cands[nc].sw = 0;
      cands[nc].chain_score = chains[c].score;
      ++nc;
    }
  }
  *acap_p = acap;
  *anchors_p = anchors;
  *chains_p = chains;
  return nc;
}

int or_map(const or_map_params_t* pin, const uint8_t* ref, const uint64_t* contig_off, const uint32_t* contig_len,
           int n_contigs, const or_index_t* idx, const char* seq, const uint64_t* seq_off, const int32_t* seq_len,
           const char* names, const uint64_t* name_off, const int32_t* name_len, size_t n, or_record_t* out,
           uint32_t* cigar_pool, size_t cap, uint64_t* cells) {
  or_map_params_t p = *pin;
  if (p.ch.band > p.sc.bw) p.ch.band = p.sc.bw; /* mapper.cpp:9 */
  const int bw = p.sc.bw, W = 2 * bw + 1;
  int maxlen = 0;
  for (size_t r = 0; r < n; ++r) maxlen = seq_len[r] > maxlen ? seq_len[r] : maxlen;
  uint8_t* codes = (uint8_t*)malloc(2 * (size_t)maxlen + 1);
  or_mini_t* mins = (or_mini_t*)malloc(sizeof(or_mini_t) * ((size_t)maxlen + 1));
  size_t acap = 1024;
  or_anchor_t* anchors = (or_anchor_t*)malloc(sizeof(or_anchor_t) * acap);
  or_chain_t* chains = (or_chain_t*)malloc(sizeof(or_chain_t) * acap);
  cand_t* cands = (cand_t*)malloc(sizeof(cand_t) * (size_t)(p.max_cand > 0 ? p.max_cand : 1));
  uint8_t* win = (uint8_t*)malloc((size_t)(p.max_cand > 0 ? p.max_cand : 1) * ((size_t)maxlen + 2 * (size_t)bw));
  uint32_t* cig = (uint32_t*)malloc(sizeof(uint32_t) * (2 * (size_t)maxlen + 4));
  size_t used = 0;
  uint64_t ncells = 0;
  int rc = 0;
  for (size_t r = 0; r < n; ++r) {
    const int len = seq_len[r];
    uint8_t* fwd = codes;
    uint8_t* rvc = codes + len;
    or_encode(seq + seq_off[r], (size_t)len, fwd);
    or_revcomp(fwd, (size_t)len, rvc);
    const int nc = seed_and_chain(&p, idx, contig_off, n_contigs, fwd, len, mins, &anchors, &chains, &acap, cands);
    /* windows (mapper.cpp:103-121): the N-padded fetch covers both cases */
    const int wl = len + 2 * bw;
    for (int c = 0; c < nc; ++c)
      fetch_window(ref, contig_off, contig_len, cands[c].contig, cands[c].wstart, wl, win + (size_t)c * wl);
    /* rank pass (mapper.cpp:124-140) */
    or_result_t res;
    if (nc >= 2) {
      for (int c = 0; c < nc; ++c) {
        or_align(&p.sc, cands[c].rev ? rvc : fwd, len, win + (size_t)c * wl, wl, 0, &res, NULL, 0);
        cands[c].sw = res.score;
        ncells += (uint64_t)len * W;
      }
    }
    /* select (mapper.cpp:142-174) */
    int chosen = -1, second = 0;
    if (nc == 1) chosen = 0;
    if (nc >= 2) {
      int s1 = -1, ties = 0;
      for (int c = 0; c < nc; ++c) {
        if (cands[c].sw > s1) {
          s1 = cands[c].sw;
          ties = 1;
        } else if (cands[c].sw == s1) {
          ++ties;
        }
      }
      int s2 = 0;
      if (ties > 1) s2 = s1;
      else
        for (int c = 0; c < nc; ++c)
          if (cands[c].sw != s1 && cands[c].sw > s2) s2 = cands[c].sw;
      uint32_t pick = or_fnv1a(names + name_off[r], (size_t)name_len[r]) % (uint32_t)ties;
      for (int c = 0; c < nc; ++c) {
        if (cands[c].sw == s1 && pick-- == 0) {
          chosen = c;
          break;
        }
      }
      second = s2;
    }
    /* records (mapper.cpp:189-222) */
    or_record_t* m = &out[r];
    memset(m, 0, sizeof *m);
    m->contig = -1;
    m->n_cand = nc;
    m->cigar_off = used;
    if (chosen < 0) continue;
    const cand_t* cd = &cands[chosen];
    const uint8_t* q = cd->rev ? rvc : fwd;
    if (or_align(&p.sc, q, len, win + (size_t)chosen * wl, wl, 1, &res, cig, 2 * maxlen + 4) != 0) {
      rc = -1;
      break;
    }
    ncells += (uint64_t)len * W;
    if (res.score < p.min_aln_score) continue;
    const int nclip = (res.qb > 0) + (len > res.qe);
    if (used + (size_t)res.n_cigar + (size_t)nclip > cap) {
      rc = -1;
      break;
    }
    m->contig = cd->contig;
    m->rev = (uint8_t)cd->rev;
    m->pos = cd->wstart + res.tb - (int64_t)contig_off[cd->contig];
    m->score = res.score;
    m->nm = res.nm;
    uint32_t* o = cigar_pool + used;
    int no = 0;
    if (res.qb > 0) o[no++] = (uint32_t)res.qb << 4 | 4;
    for (int c = 0; c < res.n_cigar; ++c) o[no++] = cig[c];
    if (len > res.qe) o[no++] = (uint32_t)(len - res.qe) << 4 | 4;
    m->n_cigar = no;
    used += (size_t)no;
    if (nc <= 1) {
      m->mapq = 60;
    } else {
      const double qv = second >= res.score ? 0.0 : 6.02 * (res.score - second) / p.sc.match + 0.499;
      m->mapq = (uint8_t)(qv < 60.0 ? qv : 60.0);
    }
  }
  if (cells) *cells = ncells;
  free(codes);
  free(mins);
  free(anchors);
  free(chains);
  free(win);
  free(cig);
  free(cands);
  return rc;
}

/* The candidates of n reads (ASCII) as seed_and_chain leaves them (src/mapper.cpp:14-63):
 * n_cand[r], and for candidate c of read r the four values rev, contig, diag and the score of
 * its chain at cand[4 * (r * max_cand + c)]. */
int or_candidates(const or_map_params_t* pin, const uint64_t* contig_off, int n_contigs, const or_index_t* idx,
                  const char* seq, const uint64_t* seq_off, const int32_t* seq_len, size_t n, int32_t* n_cand,
                  int64_t* cand) {
  or_map_params_t p = *pin;
  if (p.ch.band > p.sc.bw) p.ch.band = p.sc.bw; /* mapper.cpp:9 */
  const int mc = p.max_cand > 0 ? p.max_cand : 1;
  int maxlen = 0;
  for (size_t r = 0; r < n; ++r) maxlen = seq_len[r] > maxlen ? seq_len[r] : maxlen;
  uint8_t* fwd = (uint8_t*)malloc((size_t)maxlen + 1);
  or_mini_t* mins = (or_mini_t*)malloc(sizeof(or_mini_t) * ((size_t)maxlen + 1));
  size_t acap = 1024;
  or_anchor_t* anchors = (or_anchor_t*)malloc(sizeof(or_anchor_t) * acap);
  or_chain_t* chains = (or_chain_t*)malloc(sizeof(or_chain_t) * acap);
  cand_t* cands = (cand_t*)malloc(sizeof(cand_t) * (size_t)mc);
  for (size_t r = 0; r < n; ++r) {
    or_encode(seq + seq_off[r], (size_t)seq_len[r], fwd);
    const int nc = seed_and_chain(&p, idx, contig_off, n_contigs, fwd, seq_len[r], mins, &anchors, &chains, &acap, cands);
    n_cand[r] = nc;
    for (int c = 0; c < nc; ++c) {
      int64_t* o = cand + 4 * ((size_t)r * mc + c);
      o[0] = cands[c].rev;
      o[1] = cands[c].contig;
      o[2] = cands[c].diag;
      o[3] = cands[c].chain_score;
    }
  }
  free(fwd);
  free(mins);
  free(anchors);
  free(chains);
  free(cands);
  return 0;
}

/* ------------------------------------------------------------------ SAM (mapper.cpp:231-270) */
static char comp_char(char c) { /* common.cpp:6-14 */
  switch (c) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
    case 'a': return 't';
    case 'c': return 'g';
    case 'g': return 'c';
    case 't': return 'a';
    case 'n': return 'n';
    default: return 'N';
  }
}

long or_format_sam(const char* name, int name_len, const char* seq, const char* qual, int len,
                   const or_record_t* m, const uint32_t* cigar_pool, const char* contig_name, char* buf, size_t cap) {
  static const char kOps[] = "MIDNSHP=X";
  size_t need = (size_t)name_len + 2 * (size_t)len + 128 + (size_t)m->n_cigar * 12 + (contig_name ? strlen(contig_name) : 0);
  if (need > cap) return -1;
  char* o = buf;
  memcpy(o, name, (size_t)name_len);
  o += name_len;
  if (m->contig < 0) {
    o += sprintf(o, "\t4\t*\t0\t0\t*\t*\t0\t0\t");
    if (len == 0) *o++ = '*';
    else {
      memcpy(o, seq, (size_t)len);
      o += len;
    }
    *o++ = '\t';
    if (len == 0 || !qual) *o++ = '*';
    else {
      memcpy(o, qual, (size_t)len);
      o += len;
    }
    *o++ = '\n';
    return (long)(o - buf);
  }
  o += sprintf(o, "\t%d\t%s\t%lld\t%d\t", m->rev ? 16 : 0, contig_name, (long long)(m->pos + 1), (int)m->mapq);
  for (int c = 0; c < m->n_cigar; ++c) {
    const uint32_t x = cigar_pool[m->cigar_off + (size_t)c];
    o += sprintf(o, "%u%c", x >> 4, kOps[x & 0xf]);
  }
  o += sprintf(o, "\t*\t0\t0\t");
  if (m->rev) {
    for (int i = 0; i < len; ++i) *o++ = comp_char(seq[len - 1 - i]);
    *o++ = '\t';
    if (len == 0 || !qual) *o++ = '*';
    else
      for (int i = 0; i < len; ++i) *o++ = qual[len - 1 - i];
  } else {
    memcpy(o, seq, (size_t)len);
    o += len;
    *o++ = '\t';
    if (len == 0 || !qual) *o++ = '*';
    else {
      memcpy(o, qual, (size_t)len);
      o += len;
    }
  }
  o += sprintf(o, "\tNM:i:%d\tAS:i:%d\n", m->nm, m->score);
  return (long)(o - buf);
}
