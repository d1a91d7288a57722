"""Readers for the committed golden fixtures."""
from __future__ import annotations

import gzip
import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


def manifest() -> dict:
    return json.loads((GOLDEN / "manifest.json").read_text())


def chunk_digests() -> dict:
    p = GOLDEN / "chunk_digests.json"
    return json.loads(p.read_text()) if p.exists() else {}


def records(name: str) -> list[tuple]:
    """Per read: (QNAME, FLAG, RNAME, POS, MAPQ, CIGAR, NM, AS) with ints where numeric, None for '*'."""
    out = []
    with gzip.open(GOLDEN / f"{name}.records.tsv.gz", "rt") as fp:
        for line in fp:
            f = line.rstrip("\n").split("\t")
            nm = None if f[6] == "*" else int(f[6])
            a = None if f[7] == "*" else int(f[7])
            out.append((f[0], int(f[1]), f[2], int(f[3]), int(f[4]), f[5], nm, a))
    return out


def record_line(qname: str, fields: tuple) -> str:
    flag, rname, pos, mapq, cigar, nm, a = fields
    return "\t".join([qname, str(flag), rname, str(pos), str(mapq), cigar, "*" if nm is None else str(nm),
                      "*" if a is None else str(a)])


def kernel_vectors() -> dict:
    with np.load(GOLDEN / "kernel_vectors.npz") as z:
        return {k: z[k] for k in z.files}
