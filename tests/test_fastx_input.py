"""The input stage of dnaln_map_file (csrc/host/fastx_input.hpp) against the sequential reader
with the reference's rules (src/seqio.cpp:59-87), record for record, through tests/cpp/fastx_check.

Plain 4-line FASTQ must take the parallel parser at every batch size and thread count; anything
else (FASTA, wrapped FASTQ, blank lines, CRLF, a header-like quality line, a truncated record,
records longer than a batch) must give exactly the sequential reader's records or its error."""
from __future__ import annotations

import gzip
import subprocess
from pathlib import Path

import numpy as np
import pytest

from dnaln_amd import synth

CHECK = Path(__file__).resolve().parent / "cpp" / "bin" / "fastx_check"


def check(path: Path, batch: int, threads: int) -> str:
    if not CHECK.exists():
        pytest.fail("tests/cpp/bin/fastx_check is not built (python -m dnaln_amd.build host)")
    r = subprocess.run([str(CHECK), str(path), str(batch), str(threads)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout.strip()


def fastq_text(n: int, rng: np.random.Generator, lo: int = 50, hi: int = 300, at_qual: float = 0.0,
               names: str = "r{i} extra\tcomment") -> str:
    out = []
    for i in range(n):
        L = int(rng.integers(lo, hi + 1))
        seq = "".join(rng.choice(list("ACGTN"), L, p=[0.24, 0.24, 0.24, 0.24, 0.04]))
        q = rng.integers(33, 75, L)
        if L and rng.random() < at_qual:
            q[0] = ord("@")  # quality line that looks like a header
        out.append(f"@{names.format(i=i)}\n{seq}\n+\n{''.join(map(chr, q))}\n")
    return "".join(out)


@pytest.fixture(scope="module")
def synth_fastq(tmp_path_factory):
    d = tmp_path_factory.mktemp("fq")
    g = synth.make_genome(0, 200_000, 2, 3)
    r = synth.make_reads(g, 20_000, synth.LEN_UNIFORM, 100, 300, seed=4)
    p = d / "reads.fq"
    synth.write_fastq(r, str(p))
    with open(p, "rb") as fi, gzip.open(d / "reads.fq.gz", "wb", compresslevel=1) as fo:
        fo.write(fi.read())
    return d


@pytest.mark.parametrize("batch", [1 << 16, 1 << 20, 256 << 20])
@pytest.mark.parametrize("threads", [1, 3, 16])
def test_plain_fastq_parallel_parse_equals_sequential(synth_fastq, batch, threads):
    assert check(synth_fastq / "reads.fq", batch, threads) == "ok 20000 parallel=1"


def test_gzip_fastq(synth_fastq):
    assert check(synth_fastq / "reads.fq.gz", 1 << 20, 8) == "ok 20000 parallel=1"


@pytest.mark.parametrize("threads", [2, 8])
def test_quality_lines_starting_with_at_sign(tmp_path, threads):
    p = tmp_path / "at.fq"
    p.write_text(fastq_text(3000, np.random.default_rng(1), at_qual=0.5))
    assert check(p, 1 << 16, threads) == "ok 3000 parallel=1"


def test_crlf_and_no_final_newline(tmp_path):
    text = fastq_text(2000, np.random.default_rng(2)).replace("\n", "\r\n").rstrip("\r\n")
    p = tmp_path / "crlf.fq"
    p.write_bytes(text.encode())
    assert check(p, 1 << 16, 4) == "ok 2000 parallel=1"


def test_empty_sequences_and_empty_file(tmp_path):
    rng = np.random.default_rng(3)
    text = fastq_text(500, rng) + "@empty1\n\n+\n\n" + fastq_text(500, rng, names="s{i}") + "@empty2\n\n+\n\n"
    p = tmp_path / "empty_seq.fq"
    p.write_text(text)
    assert check(p, 1 << 16, 4).startswith("ok 1002 ")
    e = tmp_path / "empty.fq"
    e.write_text("")
    assert check(e, 1 << 16, 4) == "ok 0 parallel=1"


@pytest.mark.parametrize("kind", ["fasta", "wrapped", "blank_line", "stray_line", "long_reads"])
def test_other_forms_fall_back_to_the_sequential_rules(tmp_path, kind):
    rng = np.random.default_rng(4)
    if kind == "fasta":
        recs = []
        for i in range(1500):
            s = "".join(rng.choice(list("ACGT"), int(rng.integers(1, 400))))
            recs.append(f">f{i} d\n" + "\n".join(s[j:j + 60] for j in range(0, len(s), 60)) + "\n")
        text, n = "".join(recs), 1500
    elif kind == "wrapped":
        recs = []
        for i in range(1500):
            L = int(rng.integers(1, 300))
            s = "".join(rng.choice(list("ACGT"), L))
            q = "".join(map(chr, rng.integers(33, 74, L)))
            recs.append(f"@w{i}\n" + "\n".join(s[j:j + 70] for j in range(0, L, 70)) + "\n+\n"
                        + "\n".join(q[j:j + 70] for j in range(0, L, 70)) + "\n")
        text, n = "".join(recs), 1500
    elif kind == "blank_line":
        text, n = fastq_text(1000, rng) + "\n" + fastq_text(1000, rng, names="b{i}"), 2000
    elif kind == "stray_line":
        text, n = fastq_text(1000, rng) + "junk line\n" + fastq_text(1000, rng, names="b{i}"), 2000
    else:
        text, n = fastq_text(40, rng, lo=40_000, hi=90_000), 40
    p = tmp_path / f"{kind}.txt"
    p.write_text(text)
    out = check(p, 1 << 16, 8)
    assert out.startswith(f"ok {n} "), out


@pytest.mark.parametrize("tail", ["@t1\nACGT\n", "@t1\nACGT\n+\nII\n", "@t1\nACGT\n+\nIIIIII\n"])
def test_malformed_records_give_the_sequential_readers_error(tmp_path, tail):
    p = tmp_path / "bad.fq"
    p.write_text(fastq_text(1500, np.random.default_rng(5)) + tail)
    assert check(p, 1 << 16, 4).startswith("ok-error")
