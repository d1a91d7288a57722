/* synth.c — deterministic synthetic genomes and reads (see synth.h). */
#include "synth.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

void synth_rng_seed(synth_rng* r, uint64_t seed) {
  /* splitmix64 expands the 64-bit seed into the 256-bit xoshiro state. */
  for (int i = 0; i < 4; ++i) {
    seed += 0x9e3779b97f4a7c15ULL;
    uint64_t z = seed;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    r->s[i] = z ^ (z >> 31);
  }
}

uint64_t synth_rng_next(synth_rng* r) {
  uint64_t* s = r->s;
  const uint64_t out = rotl64(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl64(s[3], 45);
  return out;
}

double synth_rng_uniform(synth_rng* r) { return (double)(synth_rng_next(r) >> 11) * 0x1.0p-53; }
uint64_t synth_rng_below(synth_rng* r, uint64_t n) { return synth_rng_next(r) % n; }

int synth_genome(int kind, uint64_t total_len, int n_contigs, uint64_t seed, double repeat_frac, double div_lo,
                 double div_hi, uint8_t* codes, uint64_t* contig_off, uint32_t* contig_len) {
  if (n_contigs < 1 || total_len < (uint64_t)n_contigs || !codes || !contig_off || !contig_len) return -1;
  const uint64_t per = total_len / (uint64_t)n_contigs;
  for (int c = 0; c < n_contigs; ++c) {
    contig_off[c] = per * (uint64_t)c;
    contig_len[c] = (uint32_t)(c == n_contigs - 1 ? total_len - per * (uint64_t)c : per);
  }
  synth_rng r;
  synth_rng_seed(&r, seed);
  for (uint64_t i = 0; i < total_len; ++i) codes[i] = (uint8_t)(synth_rng_next(&r) >> 62);
  if (kind != SYNTH_GENOME_REPEATS || repeat_frac <= 0) return 0;

  /* Segmental copies: each copy reads the genome as it stands (earlier copies included). */
  uint8_t* tmp = (uint8_t*)malloc(10001);
  if (!tmp) return -1;
  const uint64_t target = (uint64_t)(repeat_frac * (double)total_len);
  uint64_t copied = 0;
  while (copied < target) {
    const uint32_t L = 1000 + (uint32_t)synth_rng_below(&r, 9001);
    const int cs = (int)synth_rng_below(&r, (uint64_t)n_contigs);
    const int cd = (int)synth_rng_below(&r, (uint64_t)n_contigs);
    if (contig_len[cs] < L || contig_len[cd] < L) break;
    const uint64_t ps = contig_off[cs] + synth_rng_below(&r, contig_len[cs] - L + 1);
    const uint64_t pd = contig_off[cd] + synth_rng_below(&r, contig_len[cd] - L + 1);
    const int strand = (int)(synth_rng_next(&r) & 1);
    const double div = div_lo + (div_hi - div_lo) * synth_rng_uniform(&r);
    for (uint32_t i = 0; i < L; ++i) {
      uint8_t b = strand ? (uint8_t)(3 - codes[ps + L - 1 - i]) : codes[ps + i];
      if (synth_rng_uniform(&r) < div) b = (uint8_t)((b + 1 + synth_rng_below(&r, 3)) & 3);
      tmp[i] = b;
    }
    memcpy(codes + pd, tmp, L);
    copied += L;
  }
  free(tmp);
  return 0;
}

static int read_length(synth_rng* r, const synth_read_params* p) {
  switch (p->len_mode) {
    case SYNTH_LEN_FIXED:
      return p->len_a;
    case SYNTH_LEN_LOGUNIFORM: {
      const double lo = log((double)p->len_a), hi = log((double)p->len_b + 1.0);
      int L = (int)floor(exp(lo + (hi - lo) * synth_rng_uniform(r)));
      if (L > p->len_b) L = p->len_b;
      if (L < p->len_a) L = p->len_a;
      return L;
    }
    case SYNTH_LEN_UNIFORM:
      return p->len_a + (int)synth_rng_below(r, (uint64_t)(p->len_b - p->len_a + 1));
    default: { /* edge mix: 5% of reads are 0..19 bp, the rest log-uniform in [a, b] */
      if (synth_rng_uniform(r) < 0.05) return (int)synth_rng_below(r, 20);
      const double lo = log((double)p->len_a), hi = log((double)p->len_b + 1.0);
      int L = (int)floor(exp(lo + (hi - lo) * synth_rng_uniform(r)));
      return L > p->len_b ? p->len_b : L;
    }
  }
}

static int max_read_length(const synth_read_params* p) {
  return p->len_mode == SYNTH_LEN_FIXED ? p->len_a : p->len_b;
}

uint64_t synth_reads_capacity(const synth_read_params* p) {
  return p->n * (uint64_t)(max_read_length(p) > 0 ? max_read_length(p) : 1);
}

static const char kAcgt[4] = {'A', 'C', 'G', 'T'};

/* Walks the genome from `start`, emitting `L` read bases with substitutions and indels. No indel
 * is placed before the first base, and an insertion never reaches the last base, so the true
 * alignment starts and ends with a match. Returns 0, or -1 when the walk runs off the contig. */
/* Appends op (BAM code: 0 M, 1 I, 2 D) of length len to a CIGAR, merging with the last op. */
static void cigar_push(uint32_t* cig, int* ncig, uint32_t op, uint32_t len) {
  if (!cig) return;
  if (*ncig > 0 && (cig[*ncig - 1] & 0xf) == op) cig[*ncig - 1] += len << 4;
  else cig[(*ncig)++] = len << 4 | op;
}

/* Also records the true alignment (genome orientation) in cig when it is not NULL; the random
 * draws are the same either way. */
static int walk_read(synth_rng* r, const uint8_t* codes, uint64_t start, uint64_t limit, int L,
                     const synth_read_params* p, char* out, uint32_t* cig, int* ncig) {
  uint64_t rp = start;
  int n = 0, last_match = 1;
  if (ncig) *ncig = 0;
  while (n < L) {
    if (rp >= limit) return -1;
    if (n > 0 && last_match && p->indel > 0 && synth_rng_uniform(r) < p->indel) {
      int g = 1;
      while (g < p->indel_max && synth_rng_uniform(r) < 0.3) ++g;
      if (synth_rng_uniform(r) < 0.5) {
        if (n + g >= L) continue;
        for (int i = 0; i < g; ++i) out[n++] = kAcgt[synth_rng_below(r, 4)];
        cigar_push(cig, ncig, 1, (uint32_t)g);
      } else {
        rp += (uint64_t)g;
        cigar_push(cig, ncig, 2, (uint32_t)g);
      }
      last_match = 0;
      continue;
    }
    uint8_t b = codes[rp];
    if (synth_rng_uniform(r) < p->sub) b = (uint8_t)((b + 1 + synth_rng_below(r, 3)) & 3);
    out[n++] = kAcgt[b & 3];
    cigar_push(cig, ncig, 0, 1);
    last_match = 1;
    ++rp;
  }
  return 0;
}

static char comp_char(char c) {
  switch (c) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
    default: return 'N';
  }
}

int64_t synth_reads(const uint8_t* codes, const uint64_t* contig_off, const uint32_t* contig_len, int n_contigs,
                    const synth_read_params* p, char* seq, uint64_t* offs, int32_t* lens, int64_t* truth_pos,
                    uint8_t* truth_rev) {
  return synth_reads_truth(codes, contig_off, contig_len, n_contigs, p, seq, offs, lens, truth_pos, truth_rev, NULL,
                           NULL, 0);
}

int64_t synth_reads_truth(const uint8_t* codes, const uint64_t* contig_off, const uint32_t* contig_len, int n_contigs,
                          const synth_read_params* p, char* seq, uint64_t* offs, int32_t* lens, int64_t* truth_pos,
                          uint8_t* truth_rev, uint32_t* truth_cigar, uint64_t* truth_cigar_off, uint64_t cigar_cap) {
  if (!codes || n_contigs < 1 || !p || !seq || !offs || !lens) return -1;
  int32_t maxlen = 1;
  for (int j = 0; j < 2; ++j) maxlen = (j ? p->len_b : p->len_a) > maxlen ? (j ? p->len_b : p->len_a) : maxlen;
  /* scratch for one read's ops (at most 2L: a deletion may follow every base), copied out */
  uint32_t* one = truth_cigar ? (uint32_t*)malloc((2 * (size_t)maxlen + 2) * sizeof(uint32_t)) : NULL;
  if (truth_cigar && !one) return -1;
  if ((truth_cigar != NULL) != (truth_cigar_off != NULL)) return -1;
  uint64_t co = 0;
  if (truth_cigar_off) truth_cigar_off[0] = 0;
  uint64_t total = 0;
  for (int c = 0; c < n_contigs; ++c) total += contig_len[c];
  synth_rng r;
  synth_rng_seed(&r, p->seed);
  uint64_t o = 0;
  for (uint64_t i = 0; i < p->n; ++i) {
    const int L = read_length(&r, p);
    char* out = seq + o;
    int64_t pos = -1;
    if (L > 0 && p->random_prob > 0 && synth_rng_uniform(&r) < p->random_prob) {
      for (int j = 0; j < L; ++j) out[j] = kAcgt[synth_rng_below(&r, 4)];
    } else if (L > 0) {
      for (int attempt = 0;; ++attempt) {
        if (attempt > 1000) {
          free(one);
          return -1;
        }
        int c;
        uint64_t start;
        if (p->end_prob > 0 && synth_rng_uniform(&r) < p->end_prob) {
          c = (int)synth_rng_below(&r, (uint64_t)n_contigs);
          const uint64_t slack = synth_rng_below(&r, 40);
          if (synth_rng_next(&r) & 1) {
            start = contig_off[c] + slack;
          } else {
            const uint64_t room = contig_len[c] > (uint64_t)L + slack ? contig_len[c] - (uint64_t)L - slack : 0;
            start = contig_off[c] + room;
          }
        } else {
          const uint64_t g = synth_rng_below(&r, total);
          c = 0;
          while (c + 1 < n_contigs && contig_off[c + 1] <= g) ++c;
          start = g;
        }
        int ncig = 0;
        if (walk_read(&r, codes, start, contig_off[c] + contig_len[c], L, p, out, one, &ncig) == 0) {
          pos = (int64_t)start;
          if (truth_cigar) {
            if (co + (uint64_t)ncig > cigar_cap) {
              free(one);
              return -2;
            }
            memcpy(truth_cigar + co, one, (size_t)ncig * sizeof(uint32_t));
            co += (uint64_t)ncig;
          }
          break;
        }
      }
    }
    uint8_t rev = 0;
    if (synth_rng_uniform(&r) < p->rev_prob) {
      rev = 1;
      for (int a = 0, b = L - 1; a < b; ++a, --b) {
        const char x = out[a];
        out[a] = comp_char(out[b]);
        out[b] = comp_char(x);
      }
      if (L & 1) out[L / 2] = comp_char(out[L / 2]);
    }
    if (p->n_rate > 0)
      for (int j = 0; j < L; ++j)
        if (synth_rng_uniform(&r) < p->n_rate) out[j] = 'N';
    offs[i] = o;
    lens[i] = L;
    if (truth_pos) truth_pos[i] = pos;
    if (truth_rev) truth_rev[i] = rev;
    if (truth_cigar_off) truth_cigar_off[i + 1] = co;
    o += (uint64_t)L;
  }
  free(one);
  return (int64_t)o;
}

int synth_write_fasta(const char* path, const uint8_t* codes, const uint64_t* contig_off, const uint32_t* contig_len,
                      int n_contigs) {
  FILE* fp = fopen(path, "w");
  if (!fp) return -1;
  char line[81];
  for (int c = 0; c < n_contigs; ++c) {
    fprintf(fp, ">chr%d\n", c + 1);
    for (uint64_t i = 0; i < contig_len[c]; i += 80) {
      const uint64_t m = contig_len[c] - i < 80 ? contig_len[c] - i : 80;
      for (uint64_t j = 0; j < m; ++j) line[j] = "ACGTN"[codes[contig_off[c] + i + j] > 3 ? 4 : codes[contig_off[c] + i + j]];
      line[m] = '\n';
      fwrite(line, 1, m + 1, fp);
    }
  }
  return fclose(fp) == 0 ? 0 : -1;
}

int synth_write_fastq(const char* path, const char* seq, const uint64_t* offs, const int32_t* lens, uint64_t n,
                      uint64_t first_index) {
  FILE* fp = fopen(path, "w");
  if (!fp) return -1;
  int32_t maxlen = 0;
  for (uint64_t i = 0; i < n; ++i) maxlen = lens[i] > maxlen ? lens[i] : maxlen;
  char* qual = (char*)malloc((size_t)maxlen + 1);
  if (!qual) {
    fclose(fp);
    return -1;
  }
  memset(qual, 'I', (size_t)maxlen);
  for (uint64_t i = 0; i < n; ++i) {
    fprintf(fp, "@r%llu\n", (unsigned long long)(first_index + i));
    fwrite(seq + offs[i], 1, (size_t)lens[i], fp);
    fputs("\n+\n", fp);
    fwrite(qual, 1, (size_t)lens[i], fp);
    fputc('\n', fp);
  }
  free(qual);
  return fclose(fp) == 0 ? 0 : -1;
}
