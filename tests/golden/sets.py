"""Definitions shared by the fixture writer and the tests: edge read sets, scoring schemes and
random banded-alignment tasks. Every input is regenerated from the seeds below."""
from __future__ import annotations

import numpy as np

from dnaln_amd import synth

UNIFORM_4M = {"kind": 0, "total": 4_641_652, "contigs": 1, "seed": 1, "frac": 0.0}
REPEATS_8M = {"kind": 1, "total": 8_000_000, "contigs": 8, "seed": 7, "frac": 0.3, "div": [0.005, 0.02]}
# Near-identical copies: equal SW scores across candidates, so the fnv1a tie-break is exercised.
TWINS_4M = {"kind": 1, "total": 4_000_000, "contigs": 4, "seed": 9, "frac": 0.3, "div": [0.0, 0.002]}

SETS = {
    # 4.64 Mbp uniform genome, 10,000 x 150 bp, 1% sub, 0.1% indel (1-3), bw 16.
    "uniform4m_150bp_10k": {"genome": UNIFORM_4M, "n": 10_000, "len": ["fixed", 150, 150], "sub": 0.01,
                            "indel": 0.001, "indel_max": 3, "seed": 11, "flags": {"--bw": 16}},
    # Edge mix: 5% of reads 0-19 bp, the rest 50-1000 bp; Ns, contig ends, random reads.
    "edge_uniform_bw64": {"genome": UNIFORM_4M, "n": 4000, "len": ["edge", 50, 1000], "sub": 0.02, "indel": 0.005,
                          "indel_max": 12, "n_rate": 0.003, "end_prob": 0.1, "random_prob": 0.05, "seed": 101,
                          "flags": {"--bw": 64}},
    "edge_uniform_bw16": {"genome": UNIFORM_4M, "n": 4000, "len": ["edge", 50, 1000], "sub": 0.02, "indel": 0.005,
                          "indel_max": 12, "n_rate": 0.003, "end_prob": 0.1, "random_prob": 0.05, "seed": 102,
                          "flags": {"--bw": 16}},
    # Repeat-rich genome: rank pass, MAPQ below 60, many anchors per read.
    "edge_repeats_bw32": {"genome": REPEATS_8M, "n": 6000, "len": ["fixed", 250, 250], "sub": 0.01, "indel": 0.001,
                          "indel_max": 3, "end_prob": 0.05, "seed": 103, "flags": {"--bw": 32}},
    "edge_twins_bw16": {"genome": TWINS_4M, "n": 6000, "len": ["uniform", 100, 300], "sub": 0.01, "indel": 0.001,
                        "indel_max": 3, "n_rate": 0.001, "seed": 104, "flags": {"--bw": 16}},
    # Other scoring schemes and band widths (BWA-like, short-read-like, linear gaps).
    "edge_repeats_bwa_bw8": {"genome": REPEATS_8M, "n": 3000, "len": ["uniform", 100, 300], "sub": 0.02,
                             "indel": 0.003, "indel_max": 6, "n_rate": 0.002, "end_prob": 0.05, "seed": 105,
                             "flags": {"-A": 1, "-B": 4, "-O": 6, "-E": 1, "--bw": 8}},
    "edge_twins_sr_bw5": {"genome": TWINS_4M, "n": 3000, "len": ["uniform", 100, 300], "sub": 0.02, "indel": 0.003,
                          "indel_max": 6, "n_rate": 0.002, "seed": 106,
                          "flags": {"-A": 2, "-B": 8, "-O": 12, "-E": 2, "--bw": 5}},
    "edge_uniform_linear_bw21": {"genome": UNIFORM_4M, "n": 3000, "len": ["edge", 50, 600], "sub": 0.03,
                                 "indel": 0.005, "indel_max": 8, "n_rate": 0.002, "end_prob": 0.1,
                                 "random_prob": 0.05, "seed": 107,
                                 "flags": {"-A": 3, "-B": 2, "-O": 0, "-E": 3, "--bw": 21}},
    # Non-default seeding / selection parameters.
    "edge_repeats_params": {"genome": REPEATS_8M, "n": 3000, "len": ["uniform", 60, 400], "sub": 0.015,
                            "indel": 0.002, "indel_max": 4, "seed": 108,
                            "flags": {"-k": 13, "-w": 7, "--max-occ": 40, "--max-cand": 3, "--min-chain": 30,
                                      "--min-score": 60, "--bw": 24}},
}

# Full-size workloads whose records are pinned by per-chunk digests.
DIGEST_WORKLOADS = ["uniform4m_150bp_2m", "uniform4m_uneven_1m_bw16", "uniform4m_uneven_1m_bw64",
                    "repeats120m_250bp_4m"]

SCHEMES = [
    {"match": 2, "mismatch": 4, "gap_open": 4, "gap_ext": 2, "ambig": 1, "bw": 16},   # default
    {"match": 1, "mismatch": 4, "gap_open": 6, "gap_ext": 1, "ambig": 1, "bw": 8},    # BWA-like
    {"match": 2, "mismatch": 8, "gap_open": 12, "gap_ext": 2, "ambig": 1, "bw": 5},   # short-read-like
    {"match": 3, "mismatch": 2, "gap_open": 0, "gap_ext": 3, "ambig": 1, "bw": 21},   # linear gaps
    {"match": 2, "mismatch": 4, "gap_open": 4, "gap_ext": 2, "ambig": 1, "bw": 32},
    {"match": 2, "mismatch": 4, "gap_open": 4, "gap_ext": 2, "ambig": 1, "bw": 64},
]

_LEN = {"fixed": synth.LEN_FIXED, "loguniform": synth.LEN_LOGUNIFORM, "uniform": synth.LEN_UNIFORM,
        "edge": synth.LEN_EDGE}
_GENOMES: dict = {}


def genome(spec: dict) -> synth.Genome:
    g = spec["genome"]
    key = tuple(sorted((k, tuple(v) if isinstance(v, list) else v) for k, v in g.items()))
    if key not in _GENOMES:
        div = g.get("div", [0.005, 0.02])
        _GENOMES[key] = synth.make_genome(g["kind"], g["total"], g["contigs"], g["seed"], g["frac"], div[0], div[1])
    return _GENOMES[key]


def reads(spec: dict, g: synth.Genome | None = None) -> synth.Reads:
    g = g or genome(spec)
    mode, a, b = spec["len"]
    return synth.make_reads(g, spec["n"], _LEN[mode], a, b, spec["sub"], spec["indel"], spec["indel_max"], 0.5,
                            spec["seed"], spec.get("n_rate", 0.0), spec.get("end_prob", 0.0),
                            spec.get("random_prob", 0.0))


def cli_flags(spec: dict) -> list[str]:
    out = []
    for k, v in spec["flags"].items():
        out += [k, str(v)]
    return out


def map_params(spec: dict) -> dict:
    """Mapping parameters of a set as keyword arguments (names of the C-ABI parameter block)."""
    f = spec["flags"]
    return {"k": f.get("-k", 15), "w": f.get("-w", 10), "max_occ": f.get("--max-occ", 500),
            "match": f.get("-A", 2), "mismatch": f.get("-B", 4), "gap_open": f.get("-O", 4),
            "gap_ext": f.get("-E", 2), "bw": f.get("--bw", 16), "max_cand": f.get("--max-cand", 10),
            "min_chain": f.get("--min-chain", 20), "min_aln_score": f.get("--min-score", 40)}


def random_tasks(rng: np.random.Generator, n: int, bw: int, max_len: int = 300) -> list[tuple[np.ndarray, np.ndarray]]:
    """Related and unrelated (query, target) pairs with Ns, ragged lengths and short targets."""
    out = []
    for _ in range(n):
        ql = int(rng.integers(1, max_len + 1))
        nrate = 0.02 if rng.integers(4) == 0 else 0.0
        q = rng.integers(0, 4, ql, dtype=np.uint8)
        if nrate:
            q[rng.random(ql) < nrate] = 4
        tl = ql + 2 * bw
        if rng.integers(5) == 0:
            t = rng.integers(0, 4, tl, dtype=np.uint8)
        else:
            shift = int(rng.integers(0, 2 * bw + 1))
            sub, indel = 0.03 * rng.random(), 0.02 * rng.random()
            m = []
            for c in q:
                if rng.random() < indel:
                    if rng.integers(2):
                        continue
                    m.append(int(rng.integers(4)))
                m.append(int(rng.integers(4)) if rng.random() < sub else int(c))
            t = np.concatenate([rng.integers(0, 4, shift, dtype=np.uint8), np.array(m, np.uint8),
                                rng.integers(0, 4, max(0, tl - shift - len(m)), dtype=np.uint8)])
            if nrate:
                t[rng.random(len(t)) < nrate] = 4
        if rng.integers(10) == 0:
            t = t[: int(rng.integers(0, len(t) + 1))]
        out.append((q, t))
    return out


def pack_tasks(tasks):
    """Concatenates tasks into one byte pool: (pool, q_off, q_len, t_off, t_len)."""
    parts, qo, ql, to, tl = [], [], [], [], []
    o = 0
    for q, t in tasks:
        qo.append(o)
        ql.append(len(q))
        parts.append(q)
        o += len(q)
        to.append(o)
        tl.append(len(t))
        parts.append(t)
        o += len(t)
    pool = np.concatenate(parts) if parts else np.zeros(1, np.uint8)
    return (pool.astype(np.uint8), np.array(qo, np.int64), np.array(ql, np.int32), np.array(to, np.int64),
            np.array(tl, np.int32))
