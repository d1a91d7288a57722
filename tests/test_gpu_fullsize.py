"""Full-size workloads on the GPU against the reference's per-chunk digests.

tests/golden/chunk_digests.json holds, per workload, the SHA-256 of every 65,536-record chunk of
the SAM lines (QNAME FLAG RNAME POS MAPQ CIGAR NM AS) that the reference binary wrote for the
whole workload (src/mapper.cpp:75-223, src/sam.cpp). Here the device maps the first and the last
chunk of each workload at its own size: the 120 Mbp repeat genome with its 2^26-bucket index built
on the host and on the device, both forms; chunk 0 of the 2M x 150 bp set and both uneven
workloads; and the first repeat-genome chunk served as Poisson-sized requests on four slots,
each request's records the same alone and in the mix."""
from __future__ import annotations

import hashlib

import numpy as np
import pytest

from dnaln_amd import abi, api, synth
from tests import golden_io

pytestmark = pytest.mark.gpu

REPEATS = "repeats120m_250bp_4m"
_CACHE: dict = {}


def _log(msg: str) -> None:
    print(msg, flush=True)


def _workload(name: str, n: int | None = None):
    """Genome and reads [0, n) of a workload (all of them when n is None), generated once."""
    key = (name, n)
    if key not in _CACHE:
        w = synth.WORKLOADS[name]
        g = synth.workload_genome(w)
        _log(f"{name}: generating {n or w.n_reads} reads")
        _CACHE[key] = (w, g, synth.workload_reads(w, n))
    return _CACHE[key]


def _context(name: str, index: str):
    key = ("ctx", name, index)
    if key not in _CACHE:
        w, g, _ = _workload(name, 1)
        _log(f"{name}: context with the {index} index")
        _CACHE[key] = api.Context(g, "gpu" if index == "gpu" else api.Index(g), bw=w.bw)
    return _CACHE[key]


def _digest(recs, pool, names, contig_names) -> str:
    lines = api.record_lines(names, recs, pool, contig_names)
    return hashlib.sha256(("\n".join(lines) + "\n").encode()).hexdigest()


def _golden(name: str) -> dict:
    d = golden_io.chunk_digests()[name]
    assert d["chunk"] == 65536
    return d


def _map(ctx, reads, mode: int):
    rb = api.prepare(reads)
    if mode == abi.MODE_PLAIN:
        recs, pool, _ = ctx.map(rb)
        return recs, pool
    b = ctx.batch(mode=abi.MODE_BATCHED)
    b.upload(rb)
    b.run()
    recs, pool, _ = b.download()
    b.close()
    return recs, pool


@pytest.mark.parametrize("mode", [abi.MODE_BATCHED, abi.MODE_PLAIN], ids=["batched", "plain"])
@pytest.mark.parametrize("index", ["host", "gpu"])
def test_repeat_genome_first_and_last_chunk_match_reference_digests(index, mode):
    d = _golden(REPEATS)
    w, g, reads = _workload(REPEATS)
    assert len(reads) == d["n"] == 4_000_000
    ctx = _context(REPEATS, index)
    info = ctx.index_info()
    _log(f"index {index}: {info}")
    last = (d["n"] - 1) // d["chunk"]
    for c in (0, last):
        lo, hi = c * d["chunk"], min(d["n"], (c + 1) * d["chunk"])
        sub = reads.subset(lo, hi)
        _log(f"{REPEATS} index {index} mode {mode} chunk {c} reads [{lo}, {hi})")
        recs, pool = _map(ctx, sub, mode)
        assert _digest(recs, pool, sub.names(), g.names) == d["sha256"][c], f"chunk {c}"


@pytest.mark.parametrize("mode", [abi.MODE_BATCHED, abi.MODE_PLAIN], ids=["batched", "plain"])
@pytest.mark.parametrize("name", ["uniform4m_150bp_2m", "uniform4m_uneven_1m_bw16", "uniform4m_uneven_1m_bw64"])
def test_first_chunk_matches_reference_digest(name, mode):
    d = _golden(name)
    w, g, reads = _workload(name, d["chunk"])
    ctx = _context(name, "host")
    _log(f"{name} mode {mode} chunk 0")
    recs, pool = _map(ctx, reads, mode)
    assert _digest(recs, pool, reads.names(), g.names) == d["sha256"][0]


def test_serving_on_repeat_genome_matches_reference_digest():
    """The first repeat-genome chunk as Poisson-sized requests (1..4096 reads, log-uniform) on four
    slots of the 120 Mbp context: reassembled, its records give the reference's chunk digest, and
    every request's records equal those it gets alone on one slot."""
    d = _golden(REPEATS)
    w, g, reads = _workload(REPEATS)
    sub = reads.subset(0, d["chunk"])
    rb = api.prepare(sub)
    ctx = _context(REPEATS, "host")
    first, _ = synth.request_trace(len(rb), 1.0, 17, max_size=4096)
    if first[-1] < len(rb):
        first = np.append(first, np.uint64(len(rb)))
    n_req = len(first) - 1
    _log(f"serving {n_req} requests of the first chunk on 4 slots")
    mix = api.Server(ctx, slots=4)
    lat, wall, recs, pool = mix.run_trace(rb, first, np.zeros(n_req), want_records=True)
    st = mix.stats()
    mix.close()
    _log(f"mix: {st}, wall {wall:.3f} s")
    assert st["completed"] == n_req
    assert _digest(recs, pool, sub.names(), g.names) == d["sha256"][0]
    # alone: a sample of requests (the largest and a spread of others), one at a time on one slot
    sizes = np.diff(first.astype(np.int64))
    pick = sorted(set([int(np.argmax(sizes))] + list(range(0, n_req, max(1, n_req // 12)))))
    alone = api.Server(ctx, slots=1)
    for q in pick:
        lo, hi = int(first[q]), int(first[q + 1])
        base = int(rb.offs[lo])
        alone.submit(q, api.ReadBatch(rb.codes[base:], rb.offs[lo:hi] - np.uint64(base), rb.lens[lo:hi], rb.fnv[lo:hi]))
        got = None
        while got is None:
            alone.progress()
            got = alone.take()
        _, arec, apool = got
        for j in range(hi - lo):
            assert api.record_fields(arec[j], apool, g.names) == api.record_fields(recs[lo + j], pool, g.names), (q, lo + j)
    alone.close()
    _log(f"{len(pick)} requests equal alone and in the mix")
