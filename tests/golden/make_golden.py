"""Writes the golden fixtures of tests/golden from the reference dnaln (built from its sources by
oracle/ref.mk into oracle/_ref/). Runs only where the reference is mounted.

    python tests/golden/make_golden.py [--skip-digests]

Fixtures (all inputs are regenerated from the seeds recorded in manifest.json):
  <set>.records.tsv.gz   QNAME FLAG RNAME POS MAPQ CIGAR NM AS for every read of the set
  kernel_vectors.npz     random banded-alignment tasks x 4 scoring schemes -> score, qb, qe,
                         tb, te, nm, CIGAR (reference SIMD kernel, both passes)
  chunk_digests.json     SHA-256 of the record lines per 65,536-read chunk of the full-size
                         workloads
  manifest.json          generator parameters, mapping flags and whole-file digests
"""
from __future__ import annotations

import argparse
import gzip
import hashlib
import json
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from dnaln_amd import synth  # noqa: E402
from tests.golden import sets  # noqa: E402

OUT = Path(__file__).resolve().parent
CHUNK = 65536


def record_line(f: list[str]) -> str:
    """QNAME FLAG RNAME POS MAPQ CIGAR NM AS of one SAM line ('*' tags when unmapped)."""
    nm = next((x[5:] for x in f[11:] if x.startswith("NM:i:")), "*")
    a = next((x[5:] for x in f[11:] if x.startswith("AS:i:")), "*")
    return "\t".join([f[0], f[1], f[2], f[3], f[4], f[5], nm, a])


def run_reference(fa: str, fq: str, flags: list[str], out: str) -> None:
    exe = oracle.ref_binary()
    subprocess.run([str(exe), "map", *flags, "-o", out, fa, fq], check=True, stderr=subprocess.DEVNULL)


def sam_records(path: str) -> list[str]:
    with open(path) as fp:
        return [record_line(l.rstrip("\n").split("\t")) for l in fp if not l.startswith("@")]


def sam_body_digest(path: str) -> str:
    h = hashlib.sha256()
    with open(path, "rb") as fp:
        for l in fp:
            if not l.startswith(b"@"):
                h.update(l)
    return h.hexdigest()


def make_set(name: str, spec: dict, tmp: Path) -> dict:
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    fa, fq, sam = tmp / f"{name}.fa", tmp / f"{name}.fq", tmp / f"{name}.sam"
    synth.write_fasta(g, str(fa))
    synth.write_fastq(r, str(fq))
    run_reference(str(fa), str(fq), sets.cli_flags(spec), str(sam))
    recs = sam_records(str(sam))
    assert len(recs) == len(r), (name, len(recs), len(r))
    with gzip.open(OUT / f"{name}.records.tsv.gz", "wt", compresslevel=9) as fp:
        fp.write("\n".join(recs) + "\n")
    mapped = sum(1 for x in recs if x.split("\t")[1] != "4")
    multi = sum(1 for x in recs if x.split("\t")[4] not in ("60", "0") or x.split("\t")[1] == "4")
    return {"spec": spec, "n": len(recs), "mapped": mapped, "sam_body_sha256": sam_body_digest(str(sam)),
            "records_sha256": hashlib.sha256(("\n".join(recs) + "\n").encode()).hexdigest(), "non60": multi}


def make_digests(name: str, w: synth.Workload, tmp: Path, procs: int = 8) -> dict:
    """Per-chunk SHA-256 of record lines, running the reference on shards in parallel."""
    g = synth.workload_genome(w)
    r = synth.workload_reads(w)
    fa = tmp / f"{name}.fa"
    synth.write_fasta(g, str(fa))
    flags = ["--bw", str(w.bw)]
    n = len(r)
    bounds = [(lo, min(n, lo + CHUNK)) for lo in range(0, n, CHUNK)]

    def one(b):
        lo, hi = b
        fq, sam = tmp / f"{name}.{lo}.fq", tmp / f"{name}.{lo}.sam"
        synth.write_fastq(r.subset(lo, hi), str(fq))
        run_reference(str(fa), str(fq), flags, str(sam))
        recs = sam_records(str(sam))
        os.unlink(fq)
        os.unlink(sam)
        assert len(recs) == hi - lo
        mapped = sum(1 for x in recs if x.split("\t")[1] != "4")
        return hashlib.sha256(("\n".join(recs) + "\n").encode()).hexdigest(), mapped

    with ThreadPoolExecutor(procs) as ex:
        res = list(ex.map(one, bounds))
    return {"workload": name, "bw": w.bw, "chunk": CHUNK, "n": n, "sha256": [d for d, _ in res],
            "mapped": int(sum(m for _, m in res))}


def make_kernel_vectors() -> None:
    ref = oracle.ref_lib()
    out = {}
    for si, sc in enumerate(sets.SCHEMES):
        tasks = sets.random_tasks(np.random.default_rng(1000 + si), 1500, sc["bw"], max_len=300)
        pool, qo, ql, to, tl = sets.pack_tasks(tasks)
        n = len(tasks)
        sp = np.array([sc["match"], sc["mismatch"], sc["gap_open"], sc["gap_ext"], sc["ambig"], sc["bw"]], np.int32)
        for tb in (0, 1):
            res = np.zeros((n, 6), np.int32)
            cap = int(ql.sum()) * 2 + 8 * n
            cig = np.zeros(cap, np.uint32)
            coff = np.zeros(n, np.int64)
            cn = np.zeros(n, np.int32)
            rc = ref.ref_align(sp.ctypes.data, 1, pool.ctypes.data, qo.ctypes.data, ql.ctypes.data, to.ctypes.data,
                               tl.ctypes.data, n, tb, res.ctypes.data, cig.ctypes.data, coff.ctypes.data,
                               cn.ctypes.data, cap)
            assert rc == 0
            out[f"s{si}_tb{tb}_res"] = res
            out[f"s{si}_tb{tb}_cig"] = cig[:int(cn.sum())]
            out[f"s{si}_tb{tb}_cn"] = cn
        out[f"s{si}_pool"] = pool
        out[f"s{si}_qoff"] = qo
        out[f"s{si}_qlen"] = ql
        out[f"s{si}_toff"] = to
        out[f"s{si}_tlen"] = tl
        out[f"s{si}_sp"] = sp
    np.savez_compressed(OUT / "kernel_vectors.npz", **out)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--skip-digests", action="store_true")
    ap.add_argument("--only", default="")
    a = ap.parse_args()
    if oracle.ref_binary() is None or oracle.ref_lib() is None:
        print("reference build missing: make -f oracle/ref.mk", file=sys.stderr)
        return 1
    manifest_path = OUT / "manifest.json"
    manifest = json.loads(manifest_path.read_text()) if manifest_path.exists() else {"sets": {}, "digests": {}}
    with tempfile.TemporaryDirectory() as td:
        tmp = Path(td)
        for name, spec in sets.SETS.items():
            if a.only and name != a.only:
                continue
            print("set", name, flush=True)
            manifest["sets"][name] = make_set(name, spec, tmp)
        if not a.only:
            print("kernel vectors", flush=True)
            make_kernel_vectors()
        if not a.skip_digests:
            for name in sets.DIGEST_WORKLOADS:
                print("digests", name, flush=True)
                manifest["digests"][name] = make_digests(name, synth.WORKLOADS[name], tmp)
                digest_path = OUT / "chunk_digests.json"
                digests = json.loads(digest_path.read_text()) if digest_path.exists() else {}
                digests[name] = manifest["digests"].pop(name)
                digest_path.write_text(json.dumps(digests, indent=1) + "\n")
    manifest_path.write_text(json.dumps(manifest, indent=1, sort_keys=True) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
