"""dnaln_map_file on the GPU: FASTQ (plain or gzip) in, SAM records out through the overlapped
input / device / output pipeline, byte-identical to the reference binary's SAM body on the golden
sets, at batch sizes that make many batches (so both device batches alternate) and at several
thread counts, in both forms; and dnaln_map_encoded (encoded reads through several device batches at
once) against the reference's records."""
from __future__ import annotations

import gzip
import hashlib

import numpy as np
import pytest

from dnaln_amd import abi, api, synth
from tests import golden_io
from tests.golden import sets

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,batch,threads,mode,gz", [
    ("uniform4m_150bp_10k", 1 << 16, 4, abi.MODE_BATCHED, False),
    ("uniform4m_150bp_10k", 1 << 20, 16, abi.MODE_PLAIN, True),
    ("edge_repeats_bw32", 300_000, 16, abi.MODE_BATCHED, False),
    ("edge_uniform_bw64", 1 << 18, 3, abi.MODE_BATCHED, True),
    ("edge_repeats_params", 0, 1, abi.MODE_BATCHED, False),
    ("edge_twins_sr_bw5", 1 << 17, 8, abi.MODE_PLAIN, False),
])
def test_map_file_sam_equals_reference(tmp_path, name, batch, threads, mode, gz):
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    fq = tmp_path / "reads.fq"
    synth.write_fastq(r, str(fq))
    if gz:
        with open(fq, "rb") as fi, gzip.open(tmp_path / "reads.fq.gz", "wb", compresslevel=1) as fo:
            fo.write(fi.read())
        fq = tmp_path / "reads.fq.gz"
    print(f"set {name} batch {batch} threads {threads} mode {mode} gz {gz}", flush=True)
    ctx = api.Context(g, **sets.map_params(spec))
    st = ctx.map_file(fq, tmp_path / "out.sam", mode=mode, batch_bytes=batch, threads=threads)
    ctx.close()
    assert st["reads"] == len(r) and st["parallel_fastq"] == 1
    if batch and batch < 1 << 20:
        assert st["batches"] >= 2
    body = (tmp_path / "out.sam").read_bytes()
    assert st["sam_bytes"] == len(body)
    assert hashlib.sha256(body).hexdigest() == golden_io.manifest()["sets"][name]["sam_body_sha256"]


def test_map_file_reports_a_bad_input_file(tmp_path):
    spec = sets.SETS["uniform4m_150bp_10k"]
    g = sets.genome(spec)
    bad = tmp_path / "bad.fq"
    bad.write_text("@r1\nACGTACGT\n+\nIIII\n")
    ctx = api.Context(g)
    with pytest.raises(api.DnalnError) as e:
        ctx.map_file(bad)
    assert e.value.code == abi.E_IO and "quality length mismatch" in str(e.value)
    with pytest.raises(api.DnalnError):
        ctx.map_file(tmp_path / "missing.fq")
    # the context still maps afterwards
    r = sets.reads(spec, g)
    fq = tmp_path / "ok.fq"
    synth.write_fastq(r, str(fq))
    assert ctx.map_file(fq)["reads"] == len(r)
    ctx.close()


def _golden_bad(name, recs, pool, g):
    gold = golden_io.records(name)
    return [i for i in range(len(recs)) if api.record_fields(recs[i], pool, g.names) != gold[i][1:]]


@pytest.mark.parametrize("name,n_batches,chunk,pinned", [
    ("uniform4m_150bp_10k", 2, 777, True),
    ("uniform4m_150bp_10k", 1, 0, False),
    ("edge_uniform_bw64", 3, 500, True),
    ("edge_repeats_bw32", 4, 1000, False),
    ("edge_twins_bw16", 2, 333, True),
])
def test_map_encoded_gives_reference_records(name, n_batches, chunk, pinned):
    """dnaln_map_encoded: chunks on several device batches at once (threads finishing out of order),
    CIGAR words appended in read order; every record is the reference's, from page-locked or
    ordinary host memory, and a second (warm) call gives the same."""
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    rb = api.prepare(r)
    own = None
    if pinned:
        rb, own = api.pinned_reads(rb)
    print(f"set {name} batches {n_batches} chunk {chunk} pinned {pinned}", flush=True)
    ctx = api.Context(g, **sets.map_params(spec))
    for _ in range(2):
        recs, pool, st = ctx.map_encoded(rb, n_batches=n_batches, chunk_reads=chunk)
        assert st["reads"] == len(r)
        assert st["batches"] == (len(r) + (chunk or 1 << 18) - 1) // (chunk or 1 << 18)
        assert _golden_bad(name, recs, pool, g) == []
    ctx.close()
    del own


def test_map_encoded_reports_a_short_cigar_pool():
    spec = sets.SETS["uniform4m_150bp_10k"]
    g = sets.genome(spec)
    rb = api.prepare(sets.reads(spec, g).subset(0, 2000))
    ctx = api.Context(g, **sets.map_params(spec))
    with pytest.raises(api.DnalnError) as e:
        ctx.map_encoded(rb, n_batches=2, chunk_reads=300, pool=np.zeros(10, np.uint32))
    assert e.value.code == abi.E_CAPACITY
    recs, pool, _ = ctx.map_encoded(rb, n_batches=2, chunk_reads=300)  # the batches still work
    assert (recs["contig"] >= 0).sum() > 1900
    ctx.close()
