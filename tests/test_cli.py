"""`dnaln-amd map` argument handling (src/main.cpp:12-30, src/cli.hpp:76-96) and its refusal to
run without a GPU (there is no CPU fallback). The SAM it writes is checked on the GPU
(test_gpu_cli.py)."""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

import pytest

from dnaln_amd import synth

EXE = Path(__file__).resolve().parent.parent / "dnaln_amd" / "bin" / "dnaln-amd"


def _run(*args, timeout=120):
    if not EXE.exists():
        pytest.fail("dnaln-amd is not built (python -m dnaln_amd.build host)")
    return subprocess.run([str(EXE), *map(str, args)], capture_output=True, text=True, timeout=timeout)


def test_usage_and_argument_errors():
    assert _run("map", "--help").returncode == 0
    r = _run("map", "only_one.fa")
    assert r.returncode == 1 and "Usage: dnaln-amd map" in r.stderr
    r = _run("map", "--frobnicate", "a.fa", "b.fq")
    assert r.returncode == 1 and "unknown option --frobnicate" in r.stderr
    r = _run("map", "-k", "x", "a.fa", "b.fq")
    assert r.returncode == 1 and "expects an integer" in r.stderr
    r = _run("simulate")
    assert r.returncode == 1 and "unknown command" in r.stderr


@pytest.mark.parametrize("args,msg", [(["--bw", "0"], "--bw must be in [1, 1000]"),
                                      (["--bw", "1001"], "--bw must be in [1, 1000]"),
                                      (["-A", "0"], "scores must satisfy"),
                                      (["-E", "0"], "scores must satisfy"),
                                      (["-B", "-1"], "scores must satisfy")])
def test_parameter_validation_matches_reference(args, msg):
    r = _run("map", *args, "a.fa", "b.fq")
    assert r.returncode == 1 and msg in r.stderr


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present: the GPU tests cover this path")
def test_refuses_to_map_without_a_gpu(tmp_path):
    g = synth.make_genome(0, 20000, 1, 5)
    r = synth.make_reads(g, 20, 0, 100)
    synth.write_fasta(g, str(tmp_path / "ref.fa"))
    synth.write_fastq(r, str(tmp_path / "reads.fq"))
    res = _run("map", tmp_path / "ref.fa", tmp_path / "reads.fq", "-o", tmp_path / "out.sam")
    assert res.returncode == 1
    assert "dnaln_hip_create" in res.stderr
    assert "index: k=15 w=10" in res.stderr  # the host part ran; the device part refused


def test_index_command_writes_a_loadable_file(tmp_path):
    """`dnaln-amd index` (host build) writes the index file; it loads against its reference and
    equals the in-process host build; a different reference refuses it."""
    from dnaln_amd import api
    import numpy as np

    g = synth.make_genome(0, 60_000, 3, 5)
    synth.write_fasta(g, str(tmp_path / "ref.fa"))
    r = _run("index", "-k", "13", "-w", "7", tmp_path / "ref.fa", "-o", tmp_path / "ref.dnx")
    assert r.returncode == 0, r.stderr
    assert "built on the host" in r.stderr
    loaded = api.Index.load(tmp_path / "ref.dnx", g)
    mine = api.Index(g, 13, 7)
    assert (loaded.k, loaded.w) == (13, 7)
    assert np.array_equal(loaded.table(), mine.table()) and np.array_equal(loaded.positions(), mine.positions())
    r = _run("index", tmp_path / "ref.fa")
    assert r.returncode == 1 and "Usage: dnaln-amd index" in r.stderr
    other = synth.make_genome(0, 60_000, 3, 6)
    with pytest.raises(api.DnalnError):
        api.Index.load(tmp_path / "ref.dnx", other)
