"""Deterministic synthetic genomes, reads and the named workloads of the seed-chain-extend path.

The generator is native (csrc/host/synth.c); this module wraps it with numpy arrays.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib


class _ReadParams(C.Structure):
    _fields_ = [
        ("n", C.c_uint64), ("len_mode", C.c_int), ("len_a", C.c_int), ("len_b", C.c_int),
        ("sub", C.c_double), ("indel", C.c_double), ("indel_max", C.c_int), ("rev_prob", C.c_double),
        ("n_rate", C.c_double), ("end_prob", C.c_double), ("random_prob", C.c_double), ("seed", C.c_uint64),
    ]


LEN_FIXED, LEN_LOGUNIFORM, LEN_UNIFORM, LEN_EDGE = 0, 1, 2, 3


@dataclass
class Genome:
    codes: np.ndarray          # uint8 0..3 (4 = N)
    contig_off: np.ndarray     # uint64
    contig_len: np.ndarray     # uint32
    names: list[str]

    @property
    def n_contigs(self) -> int:
        return len(self.contig_len)


@dataclass
class Reads:
    seq: np.ndarray            # uint8 ASCII, concatenated
    offs: np.ndarray           # uint64
    lens: np.ndarray           # int32
    truth_pos: np.ndarray      # int64 global start (-1 for random reads)
    truth_rev: np.ndarray      # uint8
    first_index: int = 0       # read i is named r<first_index + i>
    truth_cigar: np.ndarray | None = None    # uint32 BAM CIGAR words, genome orientation (make_reads(truth=True))
    truth_cig_off: np.ndarray | None = None  # uint64, n + 1: read i's words are [off[i], off[i + 1])

    def __len__(self) -> int:
        return len(self.lens)

    def names(self) -> list[str]:
        return [f"r{self.first_index + i}" for i in range(len(self))]

    def name_arrays(self) -> tuple[bytes, np.ndarray, np.ndarray]:
        names = self.names()
        lens = np.fromiter((len(s) for s in names), dtype=np.int32, count=len(names))
        offs = np.zeros(len(names), dtype=np.uint64)
        if len(names):
            offs[1:] = np.cumsum(lens[:-1], dtype=np.uint64)
        return "".join(names).encode(), offs, lens

    def subset(self, lo: int, hi: int) -> "Reads":
        base = int(self.offs[lo]) if hi > lo else 0
        end = int(self.offs[hi - 1] + self.lens[hi - 1]) if hi > lo else 0
        tc = to = None
        if self.truth_cigar is not None:
            c0, c1 = int(self.truth_cig_off[lo]), int(self.truth_cig_off[hi])
            tc, to = self.truth_cigar[c0:c1].copy(), (self.truth_cig_off[lo:hi + 1] - c0).astype(np.uint64)
        return Reads(self.seq[base:end].copy(), (self.offs[lo:hi] - base).astype(np.uint64), self.lens[lo:hi].copy(),
                     self.truth_pos[lo:hi].copy(), self.truth_rev[lo:hi].copy(), self.first_index + lo, tc, to)

    def read(self, i: int) -> bytes:
        o = int(self.offs[i])
        return self.seq[o:o + int(self.lens[i])].tobytes()


def make_genome(kind: int, total_len: int, n_contigs: int, seed: int, repeat_frac: float = 0.0,
                div_lo: float = 0.005, div_hi: float = 0.02) -> Genome:
    lib = _lib.host()
    codes = np.empty(total_len, dtype=np.uint8)
    off = np.empty(n_contigs, dtype=np.uint64)
    ln = np.empty(n_contigs, dtype=np.uint32)
    rc = lib.synth_genome(kind, total_len, n_contigs, seed, repeat_frac, div_lo, div_hi,
                          codes.ctypes.data, off.ctypes.data, ln.ctypes.data)
    if rc != 0:
        raise ValueError("synth_genome rejected its arguments")
    return Genome(codes, off, ln, [f"chr{i + 1}" for i in range(n_contigs)])


def make_reads(g: Genome, n: int, len_mode: int, len_a: int, len_b: int = 0, sub: float = 0.01,
               indel: float = 0.001, indel_max: int = 3, rev_prob: float = 0.5, seed: int = 11,
               n_rate: float = 0.0, end_prob: float = 0.0, random_prob: float = 0.0, truth: bool = False) -> Reads:
    """Reads drawn from g; truth=True also records each read's true CIGAR (the reads are the same)."""
    lib = _lib.host()
    p = _ReadParams(n, len_mode, len_a, len_b or len_a, sub, indel, indel_max, rev_prob, n_rate, end_prob,
                    random_prob, seed)
    cap = int(lib.synth_reads_capacity(C.byref(p)))
    seq = np.empty(max(cap, 1), dtype=np.uint8)
    offs = np.empty(n, dtype=np.uint64)
    lens = np.empty(n, dtype=np.int32)
    tpos = np.empty(n, dtype=np.int64)
    trev = np.empty(n, dtype=np.uint8)
    if not truth:
        total = lib.synth_reads(g.codes.ctypes.data, g.contig_off.ctypes.data, g.contig_len.ctypes.data, g.n_contigs,
                                C.byref(p), seq.ctypes.data, offs.ctypes.data, lens.ctypes.data, tpos.ctypes.data,
                                trev.ctypes.data)
        if total < 0:
            raise RuntimeError("synth_reads failed")
        return Reads(seq[:total].copy(), offs, lens, tpos, trev)
    ccap = 8 * n + 1024
    while True:
        cig = np.empty(ccap, dtype=np.uint32)
        coff = np.empty(n + 1, dtype=np.uint64)
        total = lib.synth_reads_truth(g.codes.ctypes.data, g.contig_off.ctypes.data, g.contig_len.ctypes.data,
                                      g.n_contigs, C.byref(p), seq.ctypes.data, offs.ctypes.data, lens.ctypes.data,
                                      tpos.ctypes.data, trev.ctypes.data, cig.ctypes.data, coff.ctypes.data, ccap)
        if total != -2:
            break
        ccap = 2 * cap + 2 * n + 2
    if total < 0:
        raise RuntimeError("synth_reads_truth failed")
    return Reads(seq[:total].copy(), offs, lens, tpos, trev, 0, cig[:int(coff[n])].copy(), coff)


def write_fasta(g: Genome, path: str) -> None:
    if _lib.host().synth_write_fasta(path.encode(), g.codes.ctypes.data, g.contig_off.ctypes.data,
                                     g.contig_len.ctypes.data, g.n_contigs) != 0:
        raise OSError(f"cannot write {path}")


def write_fastq(r: Reads, path: str) -> None:
    if _lib.host().synth_write_fastq(path.encode(), r.seq.ctypes.data, r.offs.ctypes.data, r.lens.ctypes.data,
                                     len(r), r.first_index) != 0:
        raise OSError(f"cannot write {path}")


# ------------------------------------------------------------------------------------------ workloads
@dataclass(frozen=True)
class Workload:
    """A named genome + read set + parameters. Read i is named r<i>."""
    name: str
    genome: tuple            # (kind, total_len, n_contigs, seed, repeat_frac)
    n_reads: int
    len_mode: int
    len_a: int
    len_b: int
    sub: float
    indel: float
    indel_max: int
    read_seed: int
    bw: int
    extra: dict = field(default_factory=dict)


UNIFORM_4M = (0, 4_641_652, 1, 1, 0.0)
REPEATS_120M = (1, 120_000_000, 8, 4, 0.10)

WORKLOADS = {
    # 4.64 Mbp uniform genome, 150 bp reads, 1% sub / 0.1% indel (1-3 bp), bw 16.
    "uniform4m_150bp_10k": Workload("uniform4m_150bp_10k", UNIFORM_4M, 10_000, LEN_FIXED, 150, 150, 0.01, 0.001, 3, 11, 16),
    "uniform4m_150bp_2m": Workload("uniform4m_150bp_2m", UNIFORM_4M, 2_000_000, LEN_FIXED, 150, 150, 0.01, 0.001, 3, 11, 16),
    # 1M reads log-uniform in 50..1000 bp, 2% sub / 0.5% indel (1-12 bp), bw 16 or 64.
    "uniform4m_uneven_1m_bw16": Workload("uniform4m_uneven_1m_bw16", UNIFORM_4M, 1_000_000, LEN_LOGUNIFORM, 50, 1000, 0.02, 0.005, 12, 11, 16),
    "uniform4m_uneven_1m_bw64": Workload("uniform4m_uneven_1m_bw64", UNIFORM_4M, 1_000_000, LEN_LOGUNIFORM, 50, 1000, 0.02, 0.005, 12, 11, 64),
    # 120 Mbp / 8 contigs, 10% segmental copies, 250 bp reads, bw 32.
    "repeats120m_250bp_4m": Workload("repeats120m_250bp_4m", REPEATS_120M, 4_000_000, LEN_FIXED, 250, 250, 0.01, 0.001, 3, 14, 32),
    # Serving: the same genome; read pool of lengths uniform in 100..300 bp, 1% sub / 0.1% indel, bw 32.
    "serve_repeats120m": Workload("serve_repeats120m", REPEATS_120M, 1_000_000, LEN_UNIFORM, 100, 300, 0.01, 0.001, 3, 15, 32),
}


def request_arrivals(rate_per_s: float, duration_s: float, seed: int, max_size: int = 4096):
    """Poisson arrivals over duration_s with log-uniform request sizes floor(exp(U[0, ln(max_size + 1))))
    (1..max_size reads). Returns (sizes, arrival_s)."""
    rng = np.random.default_rng(seed)
    n = max(1, int(rate_per_s * duration_s * 1.2) + 16)
    gaps = rng.exponential(1.0 / rate_per_s, n)
    arrive = np.cumsum(gaps)
    keep = arrive <= duration_s
    sizes = np.floor(np.exp(rng.uniform(0.0, np.log(max_size + 1), n))).astype(np.int64)
    sizes = np.clip(sizes, 1, max_size)
    return sizes[keep], arrive[keep]


def request_trace(n_reads: int, rate_per_s: float, seed: int, max_size: int = 4096, duration_s: float | None = None):
    """Poisson arrivals of requests with floor(exp(U[0, ln(max_size + 1)))) reads each (1..max_size),
    cut from a read pool of n_reads in order. Returns (req_first, arrival_s): request r holds reads
    [req_first[r], req_first[r+1]). Stops when the pool or duration_s runs out."""
    rng = np.random.default_rng(seed)
    first, arrive = [0], []
    t = 0.0
    while True:
        size = int(np.floor(np.exp(rng.uniform(0.0, np.log(max_size + 1)))))
        size = max(1, min(size, max_size))
        t += rng.exponential(1.0 / rate_per_s)
        if first[-1] + size > n_reads or (duration_s is not None and t > duration_s):
            break
        first.append(first[-1] + size)
        arrive.append(t)
    return np.array(first, np.uint64), np.array(arrive, np.float64)

_GENOME_CACHE: dict = {}


def workload_genome(w: Workload) -> Genome:
    if w.genome not in _GENOME_CACHE:
        kind, total, nc, seed, frac = w.genome
        _GENOME_CACHE[w.genome] = make_genome(kind, total, nc, seed, frac)
    return _GENOME_CACHE[w.genome]


def workload_reads(w: Workload, n: int | None = None, truth: bool = False) -> Reads:
    g = workload_genome(w)
    return make_reads(g, w.n_reads if n is None else n, w.len_mode, w.len_a, w.len_b, w.sub, w.indel,
                      w.indel_max, 0.5, w.read_seed, truth=truth)
