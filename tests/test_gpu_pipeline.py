"""dnaln_map_file on the GPU: FASTQ (plain or gzip) in, SAM records out through the overlapped
input / device / output pipeline, byte-identical to the reference binary's SAM body on the golden
sets, at batch sizes that make many batches (so both device batches alternate) and at several
thread counts, in both forms."""
from __future__ import annotations

import gzip
import hashlib

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
