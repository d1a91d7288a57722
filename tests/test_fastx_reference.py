"""The FASTA/FASTQ record rules (csrc/host/fastx_record.hpp) against the reference's own reader.

The reference binary, built from its sources into oracle/_ref, maps each input against a poly-A
contig with `--batch 1`. Reads here hold no A, so none maps, and the reference writes every record
it parsed as an unmapped SAM line whose columns 1, 10 and 11 are the record's name, sequence and
quality verbatim (src/mapper.cpp:231-239); a malformed record ends the run with its die() message
(src/seqio.cpp:76-83, src/common.cpp:55-64), after the records before it. dnaln_amd::SeqReader
(through tests/cpp/fastx_check --dump) must give the same records and the same message, on fixed
awkward inputs and on random texts built from header, separator, sequence and junk lines."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = ROOT / "oracle" / "_ref" / "dnaln"
CHECK = ROOT / "tests" / "cpp" / "bin" / "fastx_check"

pytestmark = pytest.mark.skipif(not REF.exists(), reason="the reference is not built here (oracle/ref.mk)")


@pytest.fixture(scope="module")
def poly_a(tmp_path_factory) -> Path:
    p = tmp_path_factory.mktemp("fx") / "polyA.fa"
    p.write_text(">chrA\n" + "A" * 2000 + "\n")
    return p


def reference_records(ref_fa: Path, reads: Path) -> list[str]:
    r = subprocess.run([str(REF), "map", "--batch", "1", str(ref_fa), str(reads)], capture_output=True, timeout=60)
    lines = r.stdout.decode("latin-1").split("\n")
    assert lines[0].startswith("@HD") and lines[1].startswith("@SQ") and lines[2].startswith("@PG"), lines[:3]
    out = []
    for ln in lines[3:]:
        if not ln:
            continue
        f = ln.split("\t")
        assert len(f) == 11 and f[1] == "4", ln  # unmapped
        out.append("\t".join((f[0], f[9], f[10])))
    if r.returncode != 0:
        err = r.stderr.decode("latin-1").strip().split("\n")[-1]
        assert err.startswith("[dnaln] error: "), err
        out.append("ERROR " + err[len("[dnaln] error: "):])
    return out


def our_records(reads: Path) -> list[str]:
    if not CHECK.exists():
        pytest.fail("tests/cpp/bin/fastx_check is not built (python -m dnaln_amd.build host)")
    r = subprocess.run([str(CHECK), "--dump", str(reads)], capture_output=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return [x for x in r.stdout.decode("latin-1").split("\n") if x]


FIXED = {
    "fasta_wrapped": ">a x\nCGT\nGGC\n>b\n\nTT\n",
    "fastq_plain": "@r1 c\nCGTC\n+\nIIII\n@r2\nGG\n+r2\n#!\n",
    "crlf_no_final_newline": "@r1\r\nCGT\r\n+\r\nIII\r\n>f\r\nGC",
    "junk_before_and_between": "junk\n\n  \n@r1\nC\n+\nI\nstray\n+\n>f\nT\n",
    "fasta_plus_line": ">f\nCG\n+notsep\nTT\n>g\nC\n",
    "empty_fastq_sequence": "@e\n+\n@f\nC\n+\nI\n",
    "empty_fastq_then_blank": "@e\n\n+\n\n@f\nCC\n+\nII\n",
    "quality_starting_with_at": "@r\nCGT\n+\n@II\n@s\nC\n+\n@\n",
    "quality_over_lines": "@r\nCGTCGT\n+\nII\nII\nII\n",
    "bare_markers": "@\nC\n+\nI\n>\nG\n",
    "name_with_tab_first": "@\tx\nC\n+\nI\n>\t\nG\n",
    "header_in_fastq_sequence": "@r\nCG\n@s\nC\n+\nI\n",
    "fastq_at_eof": "@r\nCG\n",
    "quality_too_short_at_eof": "@r\nCGT\n+\nII\n",
    "quality_too_long": "@r\nCGT\n+\nIIII\n@s\nC\n+\nI\n",
    "cr_only_lines": ">f\n\r\nCG\r\n\r\n>g\rx\nT\n",
    "cr_then_marker": ">f\n\r>notheader\nC\n",
    "empty_file": "",
    "only_junk": "xyz\n+\n\n",
    "mixed_formats": ">f1\nCG\n@q1\nTT\n+\nII\n>f2\nGGG\nC\n",
}


@pytest.mark.parametrize("name", sorted(FIXED))
def test_fixed_inputs_match_the_reference_reader(tmp_path, poly_a, name):
    p = tmp_path / f"{name}.txt"
    p.write_bytes(FIXED[name].encode())
    assert our_records(p) == reference_records(poly_a, p)


_HEADERS = [">", "@", ">n1 desc", "@n2\tx", "@@z", ">>y", "@r id", ">f"]
_PLUS = ["+", "+n1", "+ "]
_BODY = ["", "CGT", "CGTN", "cg t", "xyz", "II#", "!!!!", "GGGGGGGG", "C", "TT", "@I", ">x", "\r"]


def _random_text(rng: np.random.Generator) -> bytes:
    lines = []
    for _ in range(int(rng.integers(1, 40))):
        u = rng.random()
        if u < 0.25:
            lines.append(_HEADERS[int(rng.integers(len(_HEADERS)))])
        elif u < 0.4:
            lines.append(_PLUS[int(rng.integers(len(_PLUS)))])
        else:
            lines.append(_BODY[int(rng.integers(len(_BODY)))])
    eol = "\r\n" if rng.random() < 0.2 else "\n"
    text = eol.join(lines)
    if rng.random() < 0.7:
        text += eol
    return text.encode()


def _valid_fastq(rng: np.random.Generator) -> bytes:
    out = []
    for i in range(int(rng.integers(1, 12))):
        L = int(rng.integers(0, 30))
        s = "".join(rng.choice(list("CGTN"), L))
        q = "".join(rng.choice(list("I#!@+>"), L))
        wrap = int(rng.integers(1, 8))
        sl = [s[j:j + wrap] for j in range(0, L, wrap)] or [""]
        ql = [q[j:j + wrap] for j in range(0, L, wrap)]
        out.append(f"@v{i}\n" + "\n".join(sl) + "\n+\n" + "".join(x + "\n" for x in ql))
    return "".join(out).encode()


@pytest.mark.parametrize("seed", range(12))
def test_random_texts_match_the_reference_reader(tmp_path, poly_a, seed):
    rng = np.random.default_rng(1000 + seed)
    for k in range(25):
        text = _random_text(rng) if k % 3 else _valid_fastq(rng)
        p = tmp_path / f"t{seed}_{k}.txt"
        p.write_bytes(text)
        assert our_records(p) == reference_records(poly_a, p), text
