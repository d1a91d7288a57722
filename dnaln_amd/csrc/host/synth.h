/* synth.h — deterministic synthetic genomes and reads for the seed-chain-extend workloads.
 *
 * Every workload (uniform 4.64 Mbp genome, 120 Mbp repeat-rich genome, fixed / log-uniform /
 * uniform read lengths, substitution and geometric-indel error model) is regenerated bit for bit
 * from a seed, so the GPU path, the CPU oracle and the reference binary all read the same input.
 * The random stream is xoshiro256** seeded by splitmix64 (Blackman & Vigna, 2018).
 */
#ifndef DNALN_SYNTH_H
#define DNALN_SYNTH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  uint64_t s[4];
} synth_rng;

void synth_rng_seed(synth_rng* r, uint64_t seed);
uint64_t synth_rng_next(synth_rng* r);
double synth_rng_uniform(synth_rng* r);            /* [0, 1) with 53 random bits */
uint64_t synth_rng_below(synth_rng* r, uint64_t n); /* next() % n */

/* Genome kinds. */
enum { SYNTH_GENOME_UNIFORM = 0, SYNTH_GENOME_REPEATS = 1 };

/* Fills `codes` (total_len bytes, 0..3) and the contig table. Contigs are equal-sized
 * (the last one takes the remainder). For SYNTH_GENOME_REPEATS, `repeat_frac` of the bases are
 * then overwritten by copies of 1–10 kb segments (uniform length, random source, random strand)
 * at a per-copy substitution divergence drawn uniformly from [div_lo, div_hi].
 * Returns 0 on success, -1 on bad arguments. */
int synth_genome(int kind, uint64_t total_len, int n_contigs, uint64_t seed, double repeat_frac, double div_lo,
                 double div_hi, uint8_t* codes, uint64_t* contig_off, uint32_t* contig_len);

/* Read length models. */
enum { SYNTH_LEN_FIXED = 0, SYNTH_LEN_LOGUNIFORM = 1, SYNTH_LEN_UNIFORM = 2, SYNTH_LEN_EDGE = 3 };

typedef struct {
  uint64_t n;          /* number of reads */
  int len_mode;        /* SYNTH_LEN_* */
  int len_a, len_b;    /* fixed: a; log-uniform: floor(exp(U[ln a, ln(b+1)))); uniform: [a, b] */
  double sub;          /* per-base substitution rate */
  double indel;        /* per-base indel event rate (half insertions, half deletions) */
  int indel_max;       /* indel length = 1 + Geometric(0.3), truncated at indel_max */
  double rev_prob;     /* probability of reverse-complementing a read */
  double n_rate;       /* per-base probability of replacing a read base by 'N' (edge sets) */
  double end_prob;     /* probability a read starts within 40 bp of a contig end (edge sets) */
  double random_prob;  /* probability a read is random sequence (unmapped; edge sets) */
  uint64_t seed;
} synth_read_params;

/* Upper bound of the total read bases for `p` (to size `seq`). */
uint64_t synth_reads_capacity(const synth_read_params* p);

/* Generates reads as ASCII (A/C/G/T/N). seq receives the concatenated bases; offs[i], lens[i]
 * give read i. truth_pos/truth_rev (may be NULL) receive the global start and strand.
 * Returns the total number of bases written, or -1 on error. */
int64_t synth_reads(const uint8_t* codes, const uint64_t* contig_off, const uint32_t* contig_len, int n_contigs,
                    const synth_read_params* p, char* seq, uint64_t* offs, int32_t* lens, int64_t* truth_pos,
                    uint8_t* truth_rev);
/* synth_reads plus each read's true alignment: truth_cigar_off[i] .. [i + 1] delimit read i's
 * CIGAR (BAM encoding, M/I/D, genome orientation, starting at truth_pos[i]) in truth_cigar of
 * cigar_cap words; reads with truth_pos -1 have none. Returns -2 when cigar_cap is short. The
 * reads are the same as synth_reads gives. */
int64_t synth_reads_truth(const uint8_t* codes, const uint64_t* contig_off, const uint32_t* contig_len, int n_contigs,
                          const synth_read_params* p, char* seq, uint64_t* offs, int32_t* lens, int64_t* truth_pos,
                          uint8_t* truth_rev, uint32_t* truth_cigar, uint64_t* truth_cigar_off, uint64_t cigar_cap);

/* Writes a FASTA (80 columns) for the genome, contigs named chr1..chrN. Returns 0 or -1. */
int synth_write_fasta(const char* path, const uint8_t* codes, const uint64_t* contig_off, const uint32_t* contig_len,
                      int n_contigs);

/* Writes a FASTQ with names r<first_index + i>, quality 'I'. Returns 0 or -1. */
int synth_write_fastq(const char* path, const char* seq, const uint64_t* offs, const int32_t* lens, uint64_t n,
                      uint64_t first_index);

#ifdef __cplusplus
}
#endif

#endif
