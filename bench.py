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
DEFAULT_WORKLOAD = "uniform4m_150bp_2m"
# Lanes per pair of the packed SW fill the batched form uses by default (the band widths with a
# split fill); the autotuner's report names what it chose.
DEFAULT_SW_LANES = {32: 2, 64: 4}


def fill_kernels(W: int, lanes: int) -> str:
    """The traceback / score-only fill kernels a batched run at band W with `lanes` lanes per pair
    launches (the score-only fill splits only past W = 65)."""
    tb = f"k_sw_split<{W}, {lanes}, true>" if lanes > 1 else f"k_sw_packed<{W}, true>"
    sc = f"k_sw_split<{W}, {lanes}, false>" if lanes > 1 and W > 65 else f"k_sw_packed<{W}, false>"
    return f"{tb} / {sc}"


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
    fam = np.zeros(len(api.Batch.FAMILIES))
    launches = np.zeros(len(api.Batch.FAMILIES))
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


def measured_peak() -> dict:
    """Packed-int16 VALU peak measured on this device (dnaln_hip_peak_int16: 16 independent chains
    of v_pk_add_u16 + v_pk_max_i16 per lane, 8 waves per SIMD), beside the nominal figure."""
    tops, ms = api.peak_int16(0)
    print(f"[bench] packed int16 peak: measured {tops:.1f} T ops/s ({ms:.1f} ms), nominal {P16_PEAK / 1e12:.1f}",
          file=sys.stderr, flush=True)
    return {"measured_tops": tops, "nominal_tops": P16_PEAK / 1e12, "kernel_ms": ms,
            "kernel": "k_peak_pk16: 16 chains x (v_pk_add_u16 + v_pk_max_i16) per lane, 8 waves/SIMD on every CU"}


def config_block(name: str, peak_ops: float, steps: int, warmup: int, plain_steps: int, parity: bool) -> dict:
    """One more workload at its full size, compactly: batched reads/s after the autotuner, the plain
    form, the fill fractions of peak, the seed stage against its HBM byte model, chunk-digest
    parity with the reference."""
    w = synth.WORKLOADS[name]
    t = time.time()
    g = synth.workload_genome(w)
    reads = synth.workload_reads(w)
    rb = api.prepare(reads)
    ix = api.Index(g, build="gpu")
    ctx = api.Context(g, ix, bw=w.bw)
    gen_s = time.time() - t
    W = 2 * w.bw + 1
    plain = time_form(ctx, rb, abi.MODE_PLAIN, plain_steps, 0) if plain_steps > 0 else None
    b = time_form(ctx, rb, abi.MODE_BATCHED, steps, warmup, tune=True)
    recs, pool = b["recs"], b["pool"]
    rank_cells, tb_cells = cells_from_records(recs, reads.lens, W)
    fam = b["family_ms"]
    seed = ix.seed_traffic(rb, int(ctx.params.max_occ))
    out = {"workload": name, "reads": len(reads), "read_len": f"{w.len_a}-{w.len_b}", "genome_bp": int(len(g.codes)),
           "bw": w.bw, "reads_per_s": b["reads_per_s"], "ms_per_step": b["ms_per_step"],
           "plain_reads_per_s": plain["reads_per_s"] if plain else None,
           "batched_over_plain": b["reads_per_s"] / plain["reads_per_s"] if plain else None,
           "stage_ms": dict(zip(api.Batch.FAMILIES, fam)),
           "traceback_fill_frac": tb_cells * OPS_PER_CELL["traceback"] / (fam[3] * 1e-3) / peak_ops if fam[3] > 0 else None,
           "score_fill_frac": rank_cells * OPS_PER_CELL["score"] / (fam[2] * 1e-3) / peak_ops
           if fam[2] > 0 and rank_cells >= 10**8 else None,
           "rank_cells": rank_cells, "traceback_cells": tb_cells,
           "kernels": fill_kernels(W, b["tune"]["sw_lanes"] if b["tune"] else 1),
           "seed_model": seed,
           "seed_hbm_frac": seed["bytes"] / (fam[0] * 1e-3) / HBM_PEAK if fam[0] > 0 else None,
           "index_bytes": int(ix.n_buckets) * 16 + int(ix.n_pos) * 4,
           "autotune": {"len_split": b["tune"]["len_split"], "walk_kind": b["tune"]["walk_kind"],
                        "sw_lanes": b["tune"]["sw_lanes"], "tried": len(b["tune"]["tried"])} if b["tune"] else None}
    if plain:
        out["plain_equals_batched"] = bool(all(np.array_equal(plain["recs"][k], recs[k])
                                               for k in ("contig", "pos", "rev", "mapq", "score", "nm")))
    if parity:
        dg = json.loads((ROOT / "tests" / "golden" / "chunk_digests.json").read_text()).get(name)
        if dg and dg["n"] == len(recs):
            got = chunk_digests(recs, pool, reads.names(), g.names, dg["chunk"])
            out["reference_chunk_digests"] = f"{sum(x == y for x, y in zip(got, dg['sha256']))}/{len(dg['sha256'])}"
    ctx.close()
    print(f"[bench] {name}: {out['reads_per_s']:.0f} reads/s (plain {out['plain_reads_per_s'] or 0:.0f}), tb fill "
          f"{out['traceback_fill_frac'] or 0:.3f}, seed HBM {out['seed_hbm_frac'] or 0:.3f}, digests "
          f"{out.get('reference_chunk_digests')} (workload ready in {gen_s:.1f} s)", file=sys.stderr, flush=True)
    return out


def copy_calibration(nbytes: int) -> None:
    """One device-to-device hipMemcpy of nbytes: reads and writes exactly nbytes of HBM."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(src, 1, C.c_size_t(nbytes)) != 0 or hip.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipMemset failed")
    if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) != 0 or hip.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipMemcpy failed")
    hip.hipFree(src)
    hip.hipFree(dst)


def _pmc_rows(path: str) -> list[dict]:
    import csv
    with open(path, newline="") as fp:
        return list(csv.DictReader(fp))


def summarize_pmc(fetch_csv: str, write_csv: str, calib: int, workload: str) -> int:
    """HBM bytes per launch of the traceback fill kernel (k_sw_packed or k_sw_split) from
    FETCH_SIZE / WRITE_SIZE (KiB per dispatch), each scaled by what the same counter reads for the
    calibration copy. For the default workload the summary also goes to profiles/traffic.json,
    which the benchmark's roofline block reads."""
    out = {}
    for counter, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        rows = [r for r in _pmc_rows(path) if r.get("Counter_Name") == counter]
        # per kernel, only the launches of its largest grid: the batch-wide launch of the step (the
        # fused kernel's second pass over the few reads its first shape declined is left out)
        per, top = {}, {}
        for r in rows:
            g = int(r.get("Grid_Size") or 0)
            k = r["Kernel_Name"]
            if g > top.get(k, -1):
                top[k], per[k] = g, []
            if g == top[k]:
                per[k].append(float(r["Counter_Value"]))
        # the calibration copy: the copy launch that counted the most bytes (other copies may have
        # larger grids)
        copy = [float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"]]
        fill = [v for k, vs in per.items() if re.search(r"::k_sw_(packed|split)(_g)?<", k) and ", true>" in k for v in vs]
        walk = [v for k, vs in per.items() if re.search(r"k_walk(_g)?<", k) for v in vs]
        fused = [v for k, vs in per.items() if re.search(r"k_seed_chain_fused<|k_fused_g<", k) for v in vs]
        out[counter] = {"copy_kib": max(copy) if copy else None, "fill_kib_per_launch": sum(fill) / len(fill) if fill else None,
                        "walk_kib_per_launch": sum(walk) / len(walk) if walk else None, "fill_launches": len(fill),
                        "fused_kib_per_launch": sum(fused) / len(fused) if fused else None}
    res = {"counters": out}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        c = out[counter]
        if c["copy_kib"] and calib:
            c["scale"] = calib / (c["copy_kib"] * 1024.0)
    f, w = out["FETCH_SIZE"], out["WRITE_SIZE"]
    if f.get("scale") and w.get("scale") and f["fill_kib_per_launch"] is not None and w["fill_kib_per_launch"] is not None:
        res["k_sw_packed_tb_bytes_per_launch"] = f["fill_kib_per_launch"] * 1024 * f["scale"] + \
            w["fill_kib_per_launch"] * 1024 * w["scale"]
        res["k_sw_packed_tb_read_bytes"] = f["fill_kib_per_launch"] * 1024 * f["scale"]
        res["k_sw_packed_tb_write_bytes"] = w["fill_kib_per_launch"] * 1024 * w["scale"]
        for key, name in (("walk_kib_per_launch", "k_walk_bytes_per_launch"),
                          ("fused_kib_per_launch", "k_seed_chain_fused_bytes_per_launch")):
            if f[key] is not None and w[key] is not None:
                res[name] = f[key] * 1024 * f["scale"] + w[key] * 1024 * w["scale"]
    res["copy_bytes"] = calib
    res["workload"] = workload
    if workload == DEFAULT_WORKLOAD:
        (ROOT / "profiles").mkdir(exist_ok=True)
        (ROOT / "profiles" / "traffic.json").write_text(json.dumps(res, indent=1) + "\n")
    print(json.dumps(res, indent=1))
    return 0


def serve_sweep(seconds: float, levels: list[float], slots: int) -> dict:
    """Config 5: the 120 Mbp repeat-genome index resident; Poisson arrivals of requests of 1..4096
    reads (log-uniform); reads of 100..300 bp drawn from a pool; offered load as a fraction of the
    batched form's peak reads/s on the same pool. Latency is arrival to completion of a request."""
    w = synth.WORKLOADS["serve_repeats120m"]
    t = time.time()
    g = synth.workload_genome(w)
    pool = synth.workload_reads(w)
    rb = api.prepare(pool)
    tc = time.perf_counter()
    ctx = api.Context(g, "gpu", bw=w.bw)  # the index built on the device from the uploaded codes
    start = {"context_s": time.perf_counter() - tc, "index_build_ms": ctx.index_info()["build_ms"]}
    print(f"[bench] serving pool of {len(rb)} reads ready in {time.time() - t:.1f}s (context with a GPU-built "
          f"index in {start['context_s']:.2f} s)", file=sys.stderr, flush=True)
    b = ctx.batch(mode=abi.MODE_BATCHED)
    b.upload(rb)
    b.run()
    b.poll()
    t = time.perf_counter()
    b.run()
    b.poll()
    peak = len(rb) / (time.perf_counter() - t)
    b.close()
    mean_size = 4096 / np.log(4097.0)
    srv = api.Server(ctx, slots=slots)
    rng = np.random.default_rng(123)
    out = []
    for i, load in enumerate(levels):
        sizes, arrive = synth.request_arrivals(load * peak / mean_size, seconds, 1000 + i)
        starts = np.array([rng.integers(0, len(rb) - s + 1) for s in sizes], np.uint64)
        st0 = srv.stats()
        al0 = api.device_allocs()
        lat, wall, _, _ = srv.run_trace(rb, starts, arrive, req_count=sizes.astype(np.uint64))
        al1 = api.device_allocs()
        st = srv.stats()
        batches = st["batches"] - st0["batches"]
        row = {"load": load, "requests": int(len(sizes)), "offered_reads_per_s": float(sizes.sum() / seconds),
               "achieved_reads_per_s": float(sizes.sum() / wall) if wall > 0 else None,
               "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
               "batches": int(batches), "mean_reads_per_batch": float((st["reads"] - st0["reads"]) / max(batches, 1)),
               "device_allocs": al1[0] - al0[0], "device_alloc_bytes": al1[1] - al0[1]}
        out.append(row)
        print(f"[bench] serve load {load:.1f}: {row['requests']} requests, p50 {row['p50_ms']:.2f} ms "
              f"p99 {row['p99_ms']:.2f} ms, {row['achieved_reads_per_s']:.0f} reads/s", file=sys.stderr, flush=True)
    st = srv.stats()
    srv.close()
    ctx.close()
    return {"workload": w.name, "genome_bp": int(len(g.codes)), "read_len": "100-300", "bw": w.bw,
            "request_reads": "1-4096 log-uniform", "arrivals": "Poisson", "seconds_per_level": seconds,
            "slots": slots, "peak_batched_reads_per_s": peak, "start": start, "levels": out,
            "server": {k: st[k] for k in ("batches", "reads", "max_inflight_batches", "max_inflight_bytes",
                                          "budget_bytes", "deferred_by_budget")}}


def accuracy(ctx: api.Context, w: synth.Workload, g: synth.Genome, n: int, recs: np.ndarray, pool: np.ndarray,
             sample: int) -> dict:
    """The batched step's records scored against the generator's truth (src/evaluate.cpp:40-84),
    and the true-locus kernel checks (src/bench.cpp:272-314) on a sample of the reads."""
    t = time.perf_counter()
    tr = synth.workload_reads(w, n, truth=True)  # the same reads, with their true alignments
    e = api.evaluate(recs, pool, tr, g)
    tc = api.truth_check(ctx, tr.subset(0, min(sample, len(tr))), g)
    out = {"accuracy": e["accuracy"], "base_concordance": e["base_concordance"],
           "mapped": e["mapped"] / max(e["n_truth"], 1),
           "mapq60_error_rate": e["wrong_mapq60"] / max(e["mapped_mapq60"], 1),
           "truth_check": tc}
    print(f"[bench] accuracy {e['accuracy']:.4f}, base concordance {e['base_concordance']:.4f}, true-locus "
          f"checks {tc} ({time.perf_counter() - t:.1f} s)", file=sys.stderr, flush=True)
    return out


def end_to_end(ctx: api.Context, reads: synth.Reads, threads: int, batch_bytes: int) -> dict:
    """The workload as a FASTQ file (in the page cache) mapped to SAM text by dnaln_map_file: the
    input stage (parse, encode, hash names), the device (upload from page-locked staging, run,
    download records) and the output stage (SAM formatting, written to /dev/null) overlapped. The
    second of two runs is reported. device_pipeline_reads_per_s is reads over the time a device
    thread is busy: encoded reads in page-locked memory to records on the host, copies included;
    batches reach the device as fast as the input stage parses them."""
    with tempfile.TemporaryDirectory() as d:
        fq = os.path.join(d, "reads.fq")
        synth.write_fastq(reads, fq)
        size = os.path.getsize(fq)
        for _ in range(2):
            st = ctx.map_file(fq, None, mode=abi.MODE_BATCHED, batch_bytes=batch_bytes, threads=threads)
    n = st["reads"]
    out = {"input": f"4-line FASTQ file in the page cache, {size} bytes", "output": "SAM text to /dev/null",
           "threads": threads, "batch_bytes": batch_bytes, "reads": n, "batches": st["batches"],
           "parallel_fastq": bool(st["parallel_fastq"]), "sam_bytes": st["sam_bytes"],
           "reads_per_s": n / st["s_wall"], "device_pipeline_reads_per_s": n / st["s_device"],
           "wall_s": st["s_wall"], "busy_s": {k: st["s_" + k] for k in ("input", "device", "format", "write")}}
    print(f"[bench] end to end: {out['reads_per_s']:.0f} reads/s (device pipeline "
          f"{out['device_pipeline_reads_per_s']:.0f} reads/s), busy s {out['busy_s']}", file=sys.stderr, flush=True)
    return out


def device_pipeline(ctx: api.Context, rb: api.ReadBatch, recs: np.ndarray, pool: np.ndarray) -> dict:
    """The device pipeline: encoded reads in page-locked memory to records in page-locked memory,
    uploads and downloads included, closed loop through dnaln_map_encoded (several device batches
    at once, each on its own stream and host thread, so copies overlap kernels). A few (batches,
    chunk) settings are timed, best of three runs each; the records of the fastest setting are
    compared with the resident batched run's."""
    prb, own = api.pinned_reads(rb)
    out = api.PinnedArray(len(rb), abi.RECORD_DTYPE)
    opool = api.PinnedArray(8 * len(rb) + 64, np.uint32)
    tried = []
    for nb, chunk in ((2, 1 << 18), (3, 1 << 18), (4, 1 << 17), (6, 1 << 17), (8, 1 << 16)):
        ctx.map_encoded(prb, nb, chunk, out=out.a, pool=opool.a)  # warm: the batches' buffers
        best = min(ctx.map_encoded(prb, nb, chunk, out=out.a, pool=opool.a)[2]["s_wall"] for _ in range(3))
        tried.append({"batches": nb, "chunk_reads": chunk, "reads_per_s": len(rb) / best})
    top = max(tried, key=lambda t: t["reads_per_s"])
    r2, p2, _ = ctx.map_encoded(prb, top["batches"], top["chunk_reads"], out=out.a, pool=opool.a)
    # records field by field, and the CIGAR words they point at (the batched pool is a buffer of
    # its bound, so only its used prefix counts)
    used = int((recs["cigar_off"] + recs["cigar_len"]).max()) if len(recs) else 0
    same = bool(all(np.array_equal(r2[k], recs[k]) for k in recs.dtype.names if k != "reserved")
                and len(p2) == used and np.array_equal(p2, pool[:used]))
    res = {"reads_per_s": top["reads_per_s"], "batches": top["batches"], "chunk_reads": top["chunk_reads"],
           "tried": tried, "records_equal_batched": same,
           "input": "encoded reads (codes, offsets, lengths, name hashes) in page-locked host memory",
           "output": "records and CIGAR words in page-locked host memory"}
    del own
    print(f"[bench] device pipeline: {top['reads_per_s']:.0f} reads/s ({top['batches']} batches of "
          f"{top['chunk_reads']} reads), records equal: {same}", file=sys.stderr, flush=True)
    return res


def index_block(g: synth.Genome, k: int = 15, w: int = 10) -> tuple[api.Index, dict]:
    """MinimizerIndex (src/index.cpp:10-45) built on the host and on the GPU, checked equal
    (bucket count, key count, pos[], the set of buckets), and its file written and read back.
    Returns the host index (the context uploads it) and the timings."""
    t = time.perf_counter()
    host = api.Index(g, k, w)
    host_s = time.perf_counter() - t
    t = time.perf_counter()
    gpu = api.Index(g, k, w, build="gpu")
    gpu_s = time.perf_counter() - t
    empty = np.uint64(0xFFFFFFFFFFFFFFFF)
    th, tg = host.table(), gpu.table()
    same = (gpu.n_buckets, gpu.n_keys, gpu.max_count) == (host.n_buckets, host.n_keys, host.max_count) \
        and np.array_equal(gpu.positions(), host.positions()) \
        and np.array_equal(np.sort(th[th["key"] != empty], order="key"), np.sort(tg[tg["key"] != empty], order="key"))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "ref.dnx")
        t = time.perf_counter()
        host.save(path)
        save_s = time.perf_counter() - t
        t = time.perf_counter()
        loaded = api.Index.load(path, g)
        load_s = time.perf_counter() - t
        size = os.path.getsize(path)
    same_file = np.array_equal(loaded.table(), th) and np.array_equal(loaded.positions(), host.positions())
    blk = {"genome_bp": int(len(g.codes)), "keys": int(host.n_keys), "positions": int(host.n_pos),
           "host_build_s": host_s, "gpu_build_ms": gpu.build_ms, "gpu_build_and_download_s": gpu_s,
           "gpu_equals_host": bool(same), "file_bytes": size, "file_save_s": save_s, "file_load_s": load_s,
           "file_round_trip_equal": bool(same_file)}
    print(f"[bench] index: host {host_s:.2f} s, GPU {gpu.build_ms:.1f} ms (equal: {same}), file load {load_s:.2f} s",
          file=sys.stderr, flush=True)
    return host, blk


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD)
    ap.add_argument("--reads", type=int, default=0, help="override the read count (smaller runs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--cpu-reads-per-proc", type=int, default=60000)
    ap.add_argument("--len-split", default="0,0", help="batched form: read-length bucket limits (0 default, -1 off)")
    ap.add_argument("--dir-budget", type=int, default=0, help="bytes of traceback directions per SW launch")
    ap.add_argument("--no-tune", action="store_true", help="batched form: skip the autotuner (it runs before timing)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving sweep (config 5)")
    ap.add_argument("--serve-seconds", type=float, default=1.0, help="seconds of arrivals per load level")
    ap.add_argument("--serve-slots", type=int, default=6)
    ap.add_argument("--no-e2e", action="store_true", help="skip the FASTQ-to-SAM end-to-end run")
    ap.add_argument("--configs", default="uniform4m_uneven_1m_bw16,uniform4m_uneven_1m_bw64,repeats120m_250bp_4m",
                    help="further workloads timed at full size in compact sub-blocks ('' for none)")
    ap.add_argument("--config-steps", type=int, default=2)
    ap.add_argument("--no-accuracy", action="store_true", help="skip scoring the records against the truth")
    ap.add_argument("--e2e-threads", type=int, default=16, help="host threads of the end-to-end run")
    ap.add_argument("--e2e-batch-bytes", type=int, default=64 << 20, help="input bytes per batch of that run")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="after the run, copy BYTES device-to-device once (a copy kernel of known traffic for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="read two rocprofv3 --pmc counter_collection.csv files (FETCH_SIZE, WRITE_SIZE) of a "
                         "--calibrate-copy BYTES run (pass the same BYTES), write profiles/traffic.json and print a "
                         "summary; no GPU needed")
    a = ap.parse_args()
    if a.summarize_pmc:
        if not a.calibrate_copy:
            ap.error("--summarize-pmc needs --calibrate-copy BYTES (the size the profiled run copied)")
        return summarize_pmc(*a.summarize_pmc, a.calibrate_copy, a.workload)
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
    host_index, ix_blk = index_block(g)
    ctx = api.Context(g, host_index, bw=w.bw)
    W = 2 * w.bw + 1
    peak = measured_peak()
    peak_ops = peak["measured_tops"] * 1e12

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
    tb_ms, walk_ms, rank_ms = fam[3], fam[4], fam[2]
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

    # Roofline of the dominant kernel: the traceback fill (packed int16 VALU), timed with HIP
    # events around its launches on the batch stream.
    tb_ops = tb_cells * OPS_PER_CELL["traceback"]
    achieved = tb_ops / (tb_ms * 1e-3) if tb_ms > 0 else 0.0
    lanes = batched["tune"]["sw_lanes"] if batched["tune"] else DEFAULT_SW_LANES.get(w.bw, 1)
    roofline = {"bound": "valu-int16", "kernel": fill_kernels(W, lanes).split(" / ")[0],
                "achieved": achieved / 1e12, "peak": peak_ops / 1e12, "unit": "Tops/s (packed int16)",
                "frac": achieved / peak_ops, "traffic": None, "peak_source": peak,
                "frac_of_nominal_peak": achieved / P16_PEAK,
                "ops_per_launch": tb_ops / tb_launches, "ms_per_launch": tb_ms / tb_launches,
                "launches": int(tb_launches), "walk_ms_per_launch": walk_ms / max(batched["launches"][4], 1),
                "fill_plus_walk_frac": (tb_ops / ((tb_ms + walk_ms) * 1e-3) / peak_ops) if tb_ms + walk_ms > 0 else None,
                "score_fill_frac": (rank_cells * OPS_PER_CELL["score"] / (rank_ms * 1e-3) / peak_ops) if rank_ms > 0 else None}
    seed_model = host_index.seed_traffic(rb, int(ctx.params.max_occ))
    seed_ms = fam[0]
    hbm = {"peak_gbps": HBM_PEAK / 1e9,
           "seed_stage_model": {"bytes": seed_model["bytes"], "ms": seed_ms, "counts": seed_model,
                                "frac": seed_model["bytes"] / (seed_ms * 1e-3) / HBM_PEAK if seed_ms > 0 else None,
                                "note": "read bases + 16 B per bucket visited + 4 B per occurrence read, over the "
                                        "seed stage's event time (fused seed + chain kernels)"}}
    tpath = ROOT / "profiles" / "traffic.json"
    tr = json.loads(tpath.read_text()) if tpath.exists() else None
    if tr and tr.get("workload", DEFAULT_WORKLOAD) == w.name:
        roofline["traffic"] = tr.get("k_sw_packed_tb_bytes_per_launch")
        # HBM-side view of the two memory-bound kernels: counted bytes per launch (from the
        # committed counter runs) over this run's measured time per launch.
        n_seed = max(batched["launches"][0], 1)
        for name, key, ms in (("k_seed_chain_fused", "k_seed_chain_fused_bytes_per_launch", seed_ms / n_seed),
                              ("k_walk", "k_walk_bytes_per_launch", walk_ms / max(batched["launches"][4], 1))):
            if tr.get(key) and ms > 0:
                gbps = tr[key] / (ms * 1e-3) / 1e9
                hbm[name] = {"bytes_per_launch": tr[key], "ms_per_launch": ms, "gbps": gbps, "frac": gbps * 1e9 / HBM_PEAK}

    out = {
        "metric": METRIC, "value": batched["reads_per_s"], "unit": "reads/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": batched["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int16", "data": "synthetic (seeded generator, reads resident in HBM)",
        "config": {"workload": w.name, "reads": len(reads), "read_len": f"{w.len_a}-{w.len_b}", "genome_bp": int(len(g.codes)),
                   "bw": w.bw, "form": "batched", "parallelism": "1 GPU"},
        "plain": {"reads_per_s": plain["reads_per_s"], "ms_per_step": plain["ms_per_step"], "chunk_reads": 65536}
        if plain else None,
        "batched_over_plain": batched["reads_per_s"] / plain["reads_per_s"] if plain else None,
        "gcups": (rank_cells + tb_cells) / ((rank_ms + tb_ms + walk_ms) * 1e-3) / 1e9 if rank_ms + tb_ms > 0 else None,
        "gcups_per_step_wall": (rank_cells + tb_cells) / (batched["ms_per_step"] * 1e-3) / 1e9,
        "cells": {"rank": rank_cells, "traceback": tb_cells},
        "stage_ms": dict(zip(api.Batch.FAMILIES, fam)),
        "hbm": hbm,
        "plain_stage_ms": dict(zip(api.Batch.FAMILIES, plain["family_ms"]))
        if plain else None,
        "autotune": {"len_split": batched["tune"]["len_split"], "dir_budget": batched["tune"]["dir_budget"],
                     "walk_kind": batched["tune"]["walk_kind"], "sw_lanes": batched["tune"]["sw_lanes"],
                     "tried": [(t["len_split"], t["dir_budget"] >> 20, t["walk_kind"], t["sw_lanes"], round(t["ms"], 2))
                               for t in batched["tune"]["tried"]]}
        if batched["tune"] else None,
        "batched_stats": {k: int(getattr(batched["stats"], k)) for k in
                          ("anchors", "candidates", "rank_tasks", "tb_tasks", "global_path_reads", "fused_retry_reads")},
        "parity": parity, "roofline": roofline, "index": ix_blk,
    }
    if not a.no_accuracy:
        out["accuracy"] = accuracy(ctx, w, g, len(reads), recs, pool, 20000)
    out["device_pipeline"] = device_pipeline(ctx, rb, recs, pool)
    out["device_pipeline_reads_per_s"] = out["device_pipeline"]["reads_per_s"]
    if not a.no_e2e:
        out["end_to_end"] = end_to_end(ctx, reads, a.e2e_threads, a.e2e_batch_bytes)
    if not a.no_cpu:
        cb = cpu_baseline(w, g, reads, a.cpu_reads_per_proc)
        out["cpu_baseline"] = cb
        if cb:
            out["gpu_over_cpu"] = batched["reads_per_s"] / cb["value"]
            if "device_pipeline_reads_per_s" in out:
                out["device_pipeline_over_cpu"] = out["device_pipeline_reads_per_s"] / cb["value"]
    ctx.close()
    if a.configs:
        out["configs"] = [config_block(n, peak_ops, a.config_steps, 1, 1, not a.no_parity)
                          for n in a.configs.split(",") if n]
    if not a.no_serve:
        out["serving"] = serve_sweep(a.serve_seconds, [round(0.1 * x, 1) for x in range(1, 11)], a.serve_slots)
    if a.calibrate_copy:
        copy_calibration(a.calibrate_copy)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
