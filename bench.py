"""Benchmark of the seed-chain-extend path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu] [--no-parity]

A step is one pass of the whole hot path (sketch, probes, anchors, chaining, rank fill, select,
traceback fill + walk, records, CIGAR gather) over one batch of synthetic reads already resident
in HBM. The default workload is 2,000,000 reads of 150 bp on the 4.64 Mbp uniform genome at
bw 16. `value` is the batched form's reads/s; the plain form (one launch sequence per
65,536-read chunk) is timed beside it. Records of the last step are checked against the
reference's per-chunk SHA-256 digests (tests/golden/chunk_digests.json). Prints one JSON line.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from dnaln_amd import abi, api, synth  # noqa: E402

METRIC = "reads/s + GCUPS, batched vs plain GPU form, bit-exact SAM; p50/p99 latency"
P16_PEAK = 256 * 4 * 16 * 2 * 2.4e9     # packed int16 VALU ops/s (256 CU x 4 SIMD x 16 lanes x 2 x 2.4 GHz)
HBM_PEAK = 8.0e12
OPS_PER_CELL = {"traceback": 15, "score": 11}


def cells_from_records(recs: np.ndarray, lens: np.ndarray, W: int) -> tuple[int, int]:
    """Band cells of the rank and traceback passes, as the reference counts them (src/mapper.cpp:139, 186)."""
    nc = recs["n_cand"].astype(np.int64)
    L = lens.astype(np.int64)
    rank = int((np.where(nc >= 2, nc, 0) * L).sum()) * W
    tb = int((np.where(nc >= 1, 1, 0) * L).sum()) * W
    return rank, tb


def chunk_digests(recs, pool, names: list[str], contig_names, chunk: int) -> list[str]:
    lines = api.record_lines(names, recs, pool, contig_names)
    return [hashlib.sha256(("\n".join(lines[i:i + chunk]) + "\n").encode()).hexdigest()
            for i in range(0, len(lines), chunk)]


def time_form(ctx, rb, mode: int, steps: int, warmup: int, tune: bool = False, **kw) -> dict:
    b = ctx.batch(mode=mode, chunk_reads=65536, **kw)
    b.upload(rb)
    report = b.tune() if tune else None
    b.run()
    b.poll()
    for _ in range(warmup):
        b.run()
        b.poll()
    fam = np.zeros(5)
    launches = np.zeros(5)
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run()
        b.poll()
        ms, ln = b.timings()
        fam += np.array(ms)
        launches += np.array(ln)
    dt = time.perf_counter() - t0
    recs, pool, stats = b.download()
    b.close()
    return {"ms_per_step": 1e3 * dt / steps, "reads_per_s": len(rb) * steps / dt, "family_ms": (fam / steps).tolist(),
            "launches": (launches / steps).tolist(), "recs": recs, "pool": pool, "stats": stats, "tune": report}


def cpu_baseline(w: synth.Workload, g: synth.Genome, reads: synth.Reads, per_proc: int) -> dict | None:
    """The reference binary (built from its sources into oracle/_ref) as independent single-threaded
    processes on this host's cores, each mapping its own shard of the workload's first reads."""
    import oracle
    exe = oracle.ref_binary()
    if exe is None:
        return None
    procs = max(1, min(16, os.cpu_count() or 1))
    n = min(len(reads), procs * per_proc)
    with tempfile.TemporaryDirectory() as td:
        fa = os.path.join(td, "ref.fa")
        synth.write_fasta(g, fa)
        jobs = []
        for p in range(procs):
            lo, hi = p * n // procs, (p + 1) * n // procs
            fq = os.path.join(td, f"s{p}.fq")
            synth.write_fastq(reads.subset(lo, hi), fq)
            jobs.append(subprocess.Popen([str(exe), "map", "--bw", str(w.bw), "-o", os.devnull, fa, fq],
                                         stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True))
        secs, total = [], 0
        for j in jobs:
            _, err = j.communicate()
            m = re.search(r"mapped (\d+) / (\d+) reads .* in ([0-9.]+) s", err)
            if j.returncode != 0 or not m:
                return None
            total += int(m.group(2))
            secs.append(float(m.group(3)))
    return {"value": total / max(secs), "unit": "reads/s", "cores": procs, "kind": "reference",
            "sample": f"first {n} reads of {w.name}, {procs} independent single-threaded `dnaln map` processes "
                      f"(reference built from source with its AVX2 x16 kernel), index build excluded",
            "per_process_reads_per_s": total / procs / (sum(secs) / len(secs))}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--workload", default="uniform4m_150bp_2m")
    ap.add_argument("--reads", type=int, default=0, help="override the read count (smaller runs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--cpu-reads-per-proc", type=int, default=60000)
    ap.add_argument("--len-split", default="0,0", help="batched form: read-length bucket limits (0 default, -1 off)")
    ap.add_argument("--dir-budget", type=int, default=0, help="bytes of traceback directions per SW launch")
    ap.add_argument("--no-tune", action="store_true", help="batched form: skip the autotuner (it runs before timing)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported (reads shard "
                                   "trivially across GPUs, but no multi-GPU path is built)"}))
        return 2

    w = synth.WORKLOADS[a.workload]
    t = time.time()
    g = synth.workload_genome(w)
    reads = synth.workload_reads(w, a.reads or None)
    rb = api.prepare(reads)
    print(f"[bench] {w.name}: {len(reads)} reads generated in {time.time() - t:.1f}s", file=sys.stderr, flush=True)
    ctx = api.Context(g, bw=w.bw)
    W = 2 * w.bw + 1

    plain = time_form(ctx, rb, abi.MODE_PLAIN, a.plain_steps, 1) if a.plain_steps > 0 else None
    if plain:
        print(f"[bench] plain: {plain['reads_per_s']:.0f} reads/s", file=sys.stderr, flush=True)
    split = tuple(int(x) for x in a.len_split.split(","))
    batched = time_form(ctx, rb, abi.MODE_BATCHED, a.steps, a.warmup, tune=not a.no_tune and split == (0, 0)
                        and not a.dir_budget, len_split=split, dir_budget=a.dir_budget)
    print(f"[bench] batched: {batched['reads_per_s']:.0f} reads/s", file=sys.stderr, flush=True)

    recs, pool = batched["recs"], batched["pool"]
    rank_cells, tb_cells = cells_from_records(recs, reads.lens, W)
    fam = batched["family_ms"]
    tb_ms, rank_ms = fam[3], fam[2]
    tb_launches = max(batched["launches"][3], 1)

    parity = {"records": len(recs), "mapped": int((recs["contig"] >= 0).sum())}
    if plain:
        same = all(np.array_equal(plain["recs"][k], recs[k]) for k in ("contig", "pos", "rev", "mapq", "score", "nm"))
        parity["plain_equals_batched"] = bool(same)
    if not a.no_parity:
        dg = json.loads((ROOT / "tests" / "golden" / "chunk_digests.json").read_text()).get(w.name) \
            if (ROOT / "tests" / "golden" / "chunk_digests.json").exists() else None
        if dg and dg["n"] == len(recs):
            got = chunk_digests(recs, pool, reads.names(), g.names, dg["chunk"])
            parity["reference_chunk_digests"] = f"{sum(x == y for x, y in zip(got, dg['sha256']))}/{len(dg['sha256'])}"
        else:
            parity["reference_chunk_digests"] = "not available for this read count"

    # Roofline of the dominant kernel: the traceback fill (packed int16 VALU), time from HIP events.
    tb_ops = tb_cells * OPS_PER_CELL["traceback"]
    achieved = tb_ops / (tb_ms * 1e-3) if tb_ms > 0 else 0.0
    roofline = {"bound": "valu-int16", "kernel": f"k_sw_packed<{W},traceback> + k_walk<{W}>",
                "achieved": achieved / 1e12, "peak": P16_PEAK / 1e12, "unit": "Tops/s (packed int16)",
                "frac": achieved / P16_PEAK, "traffic": None,
                "ops_per_launch": tb_ops / tb_launches, "ms_per_launch": tb_ms / tb_launches,
                "score_fill_frac": (rank_cells * OPS_PER_CELL["score"] / (rank_ms * 1e-3) / P16_PEAK) if rank_ms > 0 else None}
    tpath = ROOT / "profiles" / "traffic.json"
    if tpath.exists():
        roofline["traffic"] = json.loads(tpath.read_text()).get("k_sw_packed_tb_bytes_per_launch")

    out = {
        "metric": METRIC, "value": batched["reads_per_s"], "unit": "reads/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": batched["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int16", "data": "synthetic (seeded generator, reads resident in HBM)",
        "config": {"workload": w.name, "reads": len(reads), "read_len": f"{w.len_a}-{w.len_b}", "genome_bp": int(len(g.codes)),
                   "bw": w.bw, "form": "batched", "parallelism": "1 GPU"},
        "plain": {"reads_per_s": plain["reads_per_s"], "ms_per_step": plain["ms_per_step"], "chunk_reads": 65536}
        if plain else None,
        "batched_over_plain": batched["reads_per_s"] / plain["reads_per_s"] if plain else None,
        "gcups": (rank_cells + tb_cells) / ((rank_ms + tb_ms) * 1e-3) / 1e9 if rank_ms + tb_ms > 0 else None,
        "gcups_per_step_wall": (rank_cells + tb_cells) / (batched["ms_per_step"] * 1e-3) / 1e9,
        "cells": {"rank": rank_cells, "traceback": tb_cells},
        "stage_ms": dict(zip(["seed", "chain", "rank_fill", "traceback_fill_walk", "run"], fam)),
        "plain_stage_ms": dict(zip(["seed", "chain", "rank_fill", "traceback_fill_walk", "run"], plain["family_ms"]))
        if plain else None,
        "autotune": {"len_split": batched["tune"]["len_split"], "dir_budget": batched["tune"]["dir_budget"],
                     "tried": [(t["len_split"], t["dir_budget"] >> 20, round(t["ms"], 2)) for t in batched["tune"]["tried"]]}
        if batched["tune"] else None,
        "batched_stats": {k: int(getattr(batched["stats"], k)) for k in
                          ("anchors", "candidates", "rank_tasks", "tb_tasks", "global_path_reads")},
        "parity": parity, "roofline": roofline,
    }
    if not a.no_cpu:
        cb = cpu_baseline(w, g, reads, a.cpu_reads_per_proc)
        out["cpu_baseline"] = cb
        if cb:
            out["gpu_over_cpu"] = batched["reads_per_s"] / cb["value"]
    ctx.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
