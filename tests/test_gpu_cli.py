"""`dnaln-amd map` end to end on the GPU: FASTA + FASTQ in, SAM out, the SAM body byte-identical
to the reference binary's on the golden sets (src/main.cpp:12-75, src/mapper.cpp:225-270)."""
from __future__ import annotations

import hashlib
import subprocess
from pathlib import Path

import pytest

from dnaln_amd import synth
from tests import golden_io
from tests.golden import sets

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parent.parent / "dnaln_amd" / "bin" / "dnaln-amd"


@pytest.mark.parametrize("name,extra", [("uniform4m_150bp_10k", []), ("uniform4m_150bp_10k", ["--plain"]),
                                        ("edge_repeats_bw32", ["--tune"]), ("edge_uniform_bw64", ["--batch", "777"]),
                                        ("edge_repeats_params", [])])
def test_cli_sam_body_equals_reference(tmp_path, name, extra):
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    synth.write_fasta(g, str(tmp_path / "ref.fa"))
    synth.write_fastq(r, str(tmp_path / "reads.fq"))
    flags = [str(x) for kv in spec["flags"].items() for x in kv]
    cmd = [str(EXE), "map", *flags, *extra, "-o", str(tmp_path / "out.sam"), str(tmp_path / "ref.fa"),
           str(tmp_path / "reads.fq")]
    print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=100)
    assert res.returncode == 0, res.stderr
    body = b"".join(l for l in (tmp_path / "out.sam").read_bytes().splitlines(keepends=True) if not l.startswith(b"@"))
    assert hashlib.sha256(body).hexdigest() == golden_io.manifest()["sets"][name]["sam_body_sha256"]
    head = [l for l in (tmp_path / "out.sam").read_text().splitlines() if l.startswith("@")]
    assert head[0] == "@HD\tVN:1.6\tSO:unsorted" and len(head) == 2 + g.n_contigs


ALIGNER = Path(__file__).resolve().parent / "cpp" / "bin" / "hip_aligner_check"


@pytest.mark.parametrize("si", range(len(sets.SCHEMES)))
def test_hip_aligner_matches_reference_kernel_vectors(tmp_path, si):
    """dnaln_amd::HipAligner (the reference's Aligner interface over dnaln_hip_align) gives the
    reference kernel's scores, coordinates and CIGARs, with and without traceback."""
    import numpy as np
    kv = golden_io.kernel_vectors()
    sp = kv[f"s{si}_sp"].astype(np.int32)
    pool, qo, ql, to, tl = (kv[f"s{si}_{k}"] for k in ("pool", "qoff", "qlen", "toff", "tlen"))
    n = len(ql)
    tasks = np.zeros(n, np.dtype([("q", "<u8"), ("t", "<u8"), ("ql", "<i4"), ("tl", "<i4")]))
    tasks["q"], tasks["t"], tasks["ql"], tasks["tl"] = qo, to, ql, tl
    f = tmp_path / "tasks.bin"
    f.write_bytes(sp.tobytes() + np.array([n, len(pool)], np.uint64).tobytes() + pool.astype(np.uint8).tobytes()
                  + tasks.tobytes())
    print(f"scheme {si} n {n}", flush=True)
    res = subprocess.run([str(ALIGNER), str(f)], capture_output=True, text=True, timeout=100)
    assert res.returncode == 0, res.stderr
    lines = res.stdout.splitlines()
    assert lines[-1] == "name hip-gfx950"
    for tb, block in ((1, lines[:n]), (0, lines[n:2 * n])):
        want, cn, cig = kv[f"s{si}_tb{tb}_res"], kv[f"s{si}_tb{tb}_cn"], kv[f"s{si}_tb{tb}_cig"]
        offs = np.concatenate([[0], np.cumsum(cn)])
        for i, line in enumerate(block):
            v = [int(x) for x in line.split()]
            assert v[:6] == want[i].tolist(), (tb, i)
            assert v[6:] == cig[offs[i]:offs[i + 1]].tolist(), (tb, i)


@pytest.mark.parametrize("how", ["gpu-build", "file"])
def test_cli_with_gpu_built_or_loaded_index_gives_reference_sam(tmp_path, how):
    """`map --gpu-index` and `index --gpu` + `map --index`: the same SAM body as the reference."""
    name = "edge_repeats_params"
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    synth.write_fasta(g, str(tmp_path / "ref.fa"))
    synth.write_fastq(r, str(tmp_path / "reads.fq"))
    flags = [str(x) for kv in spec["flags"].items() for x in kv]
    extra = ["--gpu-index"]
    if how == "file":
        kw = [str(x) for kv in spec["flags"].items() if kv[0] in ("-k", "-w") for x in kv]
        cmd = [str(EXE), "index", "--gpu", *kw, str(tmp_path / "ref.fa"), "-o", str(tmp_path / "ref.dnx")]
        print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True, timeout=100)
        assert res.returncode == 0 and "built on the GPU" in res.stderr, res.stderr
        extra = ["--index", str(tmp_path / "ref.dnx")]
    cmd = [str(EXE), "map", *flags, *extra, "-o", str(tmp_path / "out.sam"), str(tmp_path / "ref.fa"),
           str(tmp_path / "reads.fq")]
    print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=100)
    assert res.returncode == 0, res.stderr
    assert ("loaded in" if how == "file" else "built on the GPU in") in res.stderr
    body = b"".join(l for l in (tmp_path / "out.sam").read_bytes().splitlines(keepends=True) if not l.startswith(b"@"))
    assert hashlib.sha256(body).hexdigest() == golden_io.manifest()["sets"][name]["sam_body_sha256"]
