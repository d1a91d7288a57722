"""Python front end of the HIP path (through the C ABI only): index, context, align, map, batches.

Mirrors the reference's C++ surface: Context(...) plays Mapper(ref, idx, params)
(src/mapper.hpp:53), Context.align plays Aligner::align (src/align.hpp:66) and Context.map
plays Mapper::map (src/mapper.hpp:54) + append_sam. Nothing here computes alignments itself.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _lib, abi
from .synth import Genome, Reads

_ENCODE = np.full(256, 4, dtype=np.uint8)
for _c, _v in zip(b"ACGTacgtUu", (0, 1, 2, 3, 0, 1, 2, 3, 3, 3)):
    _ENCODE[_c] = _v

CIGAR_OPS = "MIDNSHP=X"


class DnalnError(RuntimeError):
    def __init__(self, code: int, where: str):
        msg = _lib.hip().dnaln_hip_last_error().decode(errors="replace")
        super().__init__(f"{where}: status {code}: {msg}")
        self.code = code


def _check(rc: int, where: str) -> None:
    if rc != abi.OK:
        raise DnalnError(rc, where)


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def encode(seq_ascii: np.ndarray) -> np.ndarray:
    """ASCII -> codes 0..4 (src/common.hpp:23-38)."""
    return _ENCODE[seq_ascii]


def device_allocs() -> tuple[int, int]:
    """Device allocations the library has made in this process: (count, bytes)."""
    ev, by = C.c_uint64(), C.c_uint64()
    _lib.hip().dnaln_hip_alloc_counters(C.byref(ev), C.byref(by))
    return int(ev.value), int(by.value)


def peak_int16(device: int = 0, iters: int = 200_000) -> tuple[float, float]:
    """Measured packed-int16 VALU peak of the device: (T int16 ops/s, ms of the timed launch)."""
    tops, ms = C.c_double(), C.c_double()
    _check(_lib.hip().dnaln_hip_peak_int16(device, iters, C.byref(tops), C.byref(ms)), "dnaln_hip_peak_int16")
    return float(tops.value), float(ms.value)


def map_params(**kw) -> abi.MapParams:
    p = abi.MapParams()
    _lib.hip().dnaln_hip_default_params(C.byref(p))
    for k, v in kw.items():
        if k in ("match", "mismatch", "gap_open", "gap_ext", "ambig", "bw"):
            setattr(p.sc, k, v)
        elif k == "min_chain":
            p.ch.min_score = v
        elif k in ("max_gap", "band", "lookback", "min_anchors"):
            setattr(p.ch, k, v)
        else:
            setattr(p, k, v)
    return p


class Index:
    """MinimizerIndex (src/index.cpp:10-45) in host memory, uploaded by Context.

    build="host" sketches and sorts on the CPU (dnaln_index_build); build="gpu" builds the same
    index on GPU `device` and downloads it (dnaln_index_build_gpu; k <= 16), with its device wall
    time in `build_ms`. Index.load reads a file written by save(), checked against the genome.
    """

    def __init__(self, genome: Genome, k: int = 15, w: int = 10, build: str = "host", device: int = 0,
                 _empty: bool = False):
        self._contigs = contig_array(genome)
        self.genome = genome
        self.idx = abi.Index()
        self.k, self.w = k, w
        self.build_ms = 0.0
        if _empty:
            return
        if build == "host":
            _check(_lib.host().dnaln_index_build(_p(genome.codes), C.cast(self._contigs, C.c_void_p),
                                                 len(self._contigs), k, w, C.byref(self.idx)), "dnaln_index_build")
        elif build == "gpu":
            ms = C.c_double(0)
            _check(_lib.host().dnaln_index_build_gpu(_p(genome.codes), C.cast(self._contigs, C.c_void_p),
                                                     len(self._contigs), k, w, device, C.byref(self.idx),
                                                     C.byref(ms)), "dnaln_index_build_gpu")
            self.build_ms = ms.value
        else:
            raise ValueError(f"build must be 'host' or 'gpu', not {build!r}")

    def save(self, path: str) -> None:
        _check(_lib.host().dnaln_index_save(C.byref(self.idx), _p(self.genome.codes),
                                            C.cast(self._contigs, C.c_void_p), len(self._contigs),
                                            str(path).encode()), f"dnaln_index_save({path})")

    @classmethod
    def load(cls, path: str, genome: Genome) -> "Index":
        ix = cls(genome, _empty=True)
        _check(_lib.host().dnaln_index_load(str(path).encode(), _p(genome.codes), C.cast(ix._contigs, C.c_void_p),
                                            len(ix._contigs), C.byref(ix.idx)), f"dnaln_index_load({path})")
        ix.k, ix.w = ix.idx.view.k, ix.idx.view.w
        return ix

    @property
    def max_count(self) -> int:
        return self.idx.max_count

    @property
    def n_buckets(self) -> int:
        return self.idx.view.n_buckets

    @property
    def n_keys(self) -> int:
        return self.idx.n_keys

    @property
    def n_pos(self) -> int:
        return self.idx.view.n_pos

    def table(self) -> np.ndarray:
        n = self.idx.view.n_buckets
        buf = (C.c_uint8 * (n * 16)).from_address(self.idx.view.table)
        return np.frombuffer(buf, dtype=np.dtype([("key", "<u8"), ("offset", "<u4"), ("count", "<u4")])).copy()

    def positions(self) -> np.ndarray:
        n = self.idx.view.n_pos
        if n == 0:
            return np.zeros(0, np.uint32)
        return np.frombuffer((C.c_uint32 * n).from_address(self.idx.view.pos), dtype=np.uint32).copy()

    def seed_traffic(self, rb: "ReadBatch", max_occ: int = 500, threads: int = 16) -> dict:
        """Algorithmic bytes of the seed stage over a prepared batch (dnaln_seed_traffic): read
        bases, minimizers, buckets visited, occurrences read, and bytes = bases + 16 B per bucket +
        4 B per occurrence."""
        t = abi.SeedTraffic()
        _check(_lib.host().dnaln_seed_traffic(C.byref(self.idx.view), _p(rb.codes), _p(rb.offs), _p(rb.lens),
                                              len(rb), max_occ, threads, C.byref(t)), "dnaln_seed_traffic")
        return {k: int(getattr(t, k)) for k, _ in t._fields_}

    def __del__(self):
        if getattr(self, "idx", None) is not None and self.idx.view.table:
            _lib.host().dnaln_index_free(C.byref(self.idx))


def contig_array(genome: Genome):
    arr = (abi.Contig * genome.n_contigs)()
    for i in range(genome.n_contigs):
        arr[i].offset = int(genome.contig_off[i])
        arr[i].len = int(genome.contig_len[i])
    return arr


@dataclass
class ReadBatch:
    """Reads encoded for the device: codes, offsets, lengths, FNV-1a of the names."""
    codes: np.ndarray
    offs: np.ndarray
    lens: np.ndarray
    fnv: np.ndarray

    def __len__(self) -> int:
        return len(self.lens)


def prepare(reads: Reads) -> ReadBatch:
    names, noff, nlen = reads.name_arrays()
    fnv = np.empty(len(reads), np.uint32)
    nb = np.frombuffer(names, dtype=np.uint8) if names else np.zeros(1, np.uint8)
    _lib.host().dnaln_fnv1a_many(_p(nb), _p(noff), _p(nlen), len(reads), _p(fnv))
    return ReadBatch(encode(reads.seq), np.ascontiguousarray(reads.offs, np.uint64),
                     np.ascontiguousarray(reads.lens, np.int32), fnv)


class PinnedArray:
    """A numpy array in page-locked host memory (dnaln_hip_host_alloc), freed with the object."""

    def __init__(self, n: int, dtype):
        dt = np.dtype(dtype)
        self.ptr = C.c_void_p()
        _check(_lib.hip().dnaln_hip_host_alloc(max(n, 1) * dt.itemsize, C.byref(self.ptr)), "dnaln_hip_host_alloc")
        buf = (C.c_uint8 * (max(n, 1) * dt.itemsize)).from_address(self.ptr.value)
        self.a = np.frombuffer(buf, dtype=dt, count=n)

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            _lib.hip().dnaln_hip_host_free(self.ptr)
            self.ptr = C.c_void_p()


def pinned_reads(rb: ReadBatch) -> tuple[ReadBatch, list]:
    """A copy of rb in page-locked memory, and the owners that keep it alive."""
    own = [PinnedArray(len(x), x.dtype) for x in (rb.codes, rb.offs, rb.lens, rb.fnv)]
    for o, x in zip(own, (rb.codes, rb.offs, rb.lens, rb.fnv)):
        o.a[:] = x
    return ReadBatch(*(o.a for o in own)), own


class Context:
    """Resident reference + index + parameters on one GPU (dnaln_hip_create)."""

    def __init__(self, genome: Genome, index: Index | str | None = None, device: int = 0, **params):
        """index: an Index (uploaded), None (built on the host, then uploaded), or "gpu" (built
        on the device from the uploaded codes; nothing crosses PCIe but the codes)."""
        self.params = map_params(**params)
        self.genome = genome
        gpu = isinstance(index, str)
        if gpu and index != "gpu":
            raise ValueError(f"index must be an Index, None or 'gpu', not {index!r}")
        self.index = None if gpu else (index or Index(genome, self.params.k, self.params.w))
        self._contigs = contig_array(genome)
        self.ctx = C.c_void_p()
        lib = _lib.hip()
        view = None if gpu else C.byref(self.index.idx.view)
        _check(lib.dnaln_hip_create(C.byref(self.params), _p(genome.codes), len(genome.codes),
                                    C.cast(self._contigs, C.c_void_p), genome.n_contigs, view,
                                    device, C.byref(self.ctx)), "dnaln_hip_create")

    def index_info(self) -> dict:
        """Bucket and position counts of the resident index, and its GPU build time (0 if uploaded)."""
        nb, npos, ms = C.c_uint64(0), C.c_uint64(0), C.c_double(0)
        _check(_lib.hip().dnaln_hip_context_index(self.ctx, C.byref(nb), C.byref(npos), C.byref(ms)),
               "dnaln_hip_context_index")
        return {"n_buckets": nb.value, "n_pos": npos.value, "build_ms": ms.value}

    def close(self) -> None:
        for group in list(getattr(self, "_file_batches", {}).values()) + list(getattr(self, "_enc_batches", {}).values()):
            for b in group:
                b.close()
        self._file_batches = {}
        self._enc_batches = {}
        if self.ctx:
            _lib.hip().dnaln_hip_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------------------- Aligner::align
    def align(self, pool: np.ndarray, tasks: np.ndarray, traceback: bool, score: dict | None = None):
        """tasks: TASK_DTYPE array over the byte pool. Returns (RESULT_DTYPE array, cigar pool)."""
        pool = np.ascontiguousarray(pool, np.uint8)
        tasks = np.ascontiguousarray(tasks, abi.TASK_DTYPE)
        n = len(tasks)
        out = np.zeros(n, abi.RESULT_DTYPE)
        cap = int(tasks["qlen"].astype(np.int64).sum()) * 2 + 4 * n + 4
        cig = np.zeros(cap, np.uint32)
        sp = None
        if score is not None:
            sp = abi.ScoreParams(score["match"], score["mismatch"], score["gap_open"], score["gap_ext"],
                                 score.get("ambig", 1), score["bw"])
        _check(_lib.hip().dnaln_hip_align(self.ctx, C.byref(sp) if sp is not None else None, _p(tasks), n, _p(pool),
                                          len(pool), int(traceback), _p(out), _p(cig), cap), "dnaln_hip_align")
        return out, cig

    # ---------------------------------------------------------------------------- Mapper::map
    def map(self, reads: Reads | ReadBatch):
        rb = reads if isinstance(reads, ReadBatch) else prepare(reads)
        n = len(rb)
        out = np.zeros(n, abi.RECORD_DTYPE)
        cap = int(rb.lens.astype(np.int64).sum()) * 2 + 4 * n + 4
        cig = np.zeros(cap, np.uint32)
        stats = abi.MapStats()
        _check(_lib.hip().dnaln_hip_map_batch(self.ctx, _p(rb.codes), _p(rb.offs), _p(rb.lens), _p(rb.fnv), n,
                                              _p(out), _p(cig), cap, C.byref(stats)), "dnaln_hip_map_batch")
        return out, cig, stats

    def map_file(self, reads_path: str, out_path: str | None = None, mode: int = abi.MODE_BATCHED,
                 batch_bytes: int = 0, threads: int = 16, tune: bool = False) -> dict:
        """FASTA/FASTQ file (plain or gzip) -> SAM records (no header) at out_path (None: discarded),
        through the overlapped input / device / output pipeline (dnaln_map_file). Returns its stats."""
        # two device batches per form, kept on the context: later calls run warm
        if not hasattr(self, "_file_batches"):
            self._file_batches = {}
        key = (mode, bool(tune))
        if key not in self._file_batches:
            self._file_batches[key] = (self.batch(mode=mode, tune=tune), self.batch(mode=mode, tune=tune))
        b0, b1 = self._file_batches[key]
        names = (C.c_char_p * self.genome.n_contigs)(*[n.encode() for n in self.genome.names])
        st = abi.PipelineStats()
        fd = os.open(out_path or os.devnull, os.O_WRONLY | os.O_CREAT | os.O_TRUNC, 0o644)
        try:
            rc = _lib.host().dnaln_map_file_on(b0.b, b1.b, str(reads_path).encode(), fd, names,
                                               self.genome.n_contigs, batch_bytes, threads, C.byref(st))
        finally:
            os.close(fd)
        if rc != abi.OK:
            msg = _lib.host().dnaln_host_last_error().decode(errors="replace")
            raise DnalnError(rc, f"dnaln_map_file({reads_path}): {msg}")
        return {f: getattr(st, f) for f, _ in abi.PipelineStats._fields_ if f != "reserved"}

    def map_encoded(self, rb: ReadBatch, n_batches: int = 2, chunk_reads: int = 0, mode: int = abi.MODE_BATCHED,
                    out: np.ndarray | None = None, pool: np.ndarray | None = None):
        """Encoded reads -> (records, CIGAR pool, stats) through n_batches device batches at once
        (dnaln_map_encoded): chunks of chunk_reads, one batch's copies overlapping another's
        kernels. Page-locked rb / out / pool (pinned_reads, PinnedArray) make the copies direct."""
        if not hasattr(self, "_enc_batches"):
            self._enc_batches = {}
        key = (mode, n_batches)
        if key not in self._enc_batches:
            self._enc_batches[key] = [self.batch(mode=mode) for _ in range(n_batches)]
        bs = self._enc_batches[key]
        arr = (C.c_void_p * n_batches)(*[b.b for b in bs])
        if out is None:
            out = np.zeros(len(rb), abi.RECORD_DTYPE)
        if pool is None:
            pool = np.zeros(8 * len(rb) + 64, np.uint32)
        used = C.c_uint64()
        st = abi.PipelineStats()
        rc = _lib.host().dnaln_map_encoded(arr, n_batches, _p(rb.codes), _p(rb.offs), _p(rb.lens), _p(rb.fnv), len(rb),
                                           chunk_reads, _p(out), _p(pool), len(pool), C.byref(used), C.byref(st))
        if rc != abi.OK:
            msg = _lib.host().dnaln_host_last_error().decode(errors="replace")
            raise DnalnError(rc, f"dnaln_map_encoded: {msg}")
        return out, pool[:used.value], {f: getattr(st, f) for f, _ in abi.PipelineStats._fields_ if f != "reserved"}

    def batch(self, mode: int = abi.MODE_PLAIN, chunk_reads: int = 65536, max_reads: int = 0, max_bases: int = 0,
              tune: bool = False, len_split: tuple[int, int] = (0, 0), dir_budget: int = 0,
              walk_kind: int = 0, graph: int = 0, anchors_per_read: int = 0, sw_lanes: int = 0) -> "Batch":
        return Batch(self, mode, chunk_reads, max_reads, max_bases, tune, len_split, dir_budget, walk_kind, graph,
                     anchors_per_read, sw_lanes)


class Batch:
    """A device-resident batch (dnaln_hip_batch_*): upload once, run many times, download."""

    def __init__(self, ctx: Context, mode: int, chunk_reads: int, max_reads: int, max_bases: int, tune: bool,
                 len_split: tuple[int, int] = (0, 0), dir_budget: int = 0, walk_kind: int = 0, graph: int = 0,
                 anchors_per_read: int = 0, sw_lanes: int = 0):
        """graph (batched form): 0 the host-synchronized form (default); 2 the device-count kernel
        sequence launched one by one, 1 it captured as a HIP graph per plan (both refused by the
        device build; the CPU emulation runs them).
        anchors_per_read (batched form): first anchor capacity of the global path per read (0:
        estimated; a run that needs more grows it and runs again).
        sw_lanes: lanes per pair of alignments in the packed SW fill (1; 2 for bw 32 or 4 for
        bw 64 splits the band over that many lanes; 0 the default: the split for the batched
        form where the band width has one)."""
        self.ctx = ctx
        cfg = abi.BatchConfig()
        _lib.hip().dnaln_hip_default_batch_config(C.byref(cfg))
        cfg.mode, cfg.chunk_reads, cfg.max_reads, cfg.max_bases, cfg.tune = mode, chunk_reads, max_reads, max_bases, int(tune)
        cfg.len_split[0], cfg.len_split[1] = len_split
        cfg.dir_budget = dir_budget
        cfg.walk_kind = walk_kind
        cfg.graph = graph
        cfg.sw_lanes = sw_lanes
        if anchors_per_read:
            cfg.anchors_per_read = anchors_per_read
        self.cfg = cfg
        self.b = C.c_void_p()
        _check(_lib.hip().dnaln_hip_batch_create(ctx.ctx, C.byref(cfg), C.byref(self.b)), "dnaln_hip_batch_create")
        self._keep = None
        self.n = 0
        self.bases = 0

    def upload(self, rb: ReadBatch) -> None:
        self._keep = rb
        self.n = len(rb)
        self.bases = int(rb.lens.astype(np.int64).sum())
        _check(_lib.hip().dnaln_hip_batch_upload(self.b, _p(rb.codes), _p(rb.offs), _p(rb.lens), _p(rb.fnv), self.n),
               "dnaln_hip_batch_upload")

    def run(self) -> None:
        _check(_lib.hip().dnaln_hip_batch_run(self.b), "dnaln_hip_batch_run")

    def poll(self, wait: bool = True) -> bool:
        rc = _lib.hip().dnaln_hip_batch_poll(self.b, int(wait))
        if rc == abi.E_PENDING:
            return False
        _check(rc, "dnaln_hip_batch_poll")
        return True

    def download(self):
        out = np.zeros(self.n, abi.RECORD_DTYPE)
        cap = self.bases * 2 + 4 * self.n + 4
        cig = np.zeros(cap, np.uint32)
        stats = abi.MapStats()
        _check(_lib.hip().dnaln_hip_batch_download(self.b, _p(out), _p(cig), cap, C.byref(stats)),
               "dnaln_hip_batch_download")
        return out, cig, stats

    def candidates(self):
        """The last run's candidates after seeding and chaining: (n_cand[n], cand[n, max_cand] of
        abi.CAND_DTYPE), in the reference's order (dnaln_hip_batch_candidates)."""
        mc = max(int(self.ctx.params.max_cand), 1)
        ncand = np.zeros(max(self.n, 1), np.int32)
        cand = np.zeros((max(self.n, 1), mc), abi.CAND_DTYPE)
        _check(_lib.hip().dnaln_hip_batch_candidates(self.b, _p(ncand), _p(cand), mc), "dnaln_hip_batch_candidates")
        return ncand[:self.n], cand[:self.n]

    FAMILIES = ("seed", "chain", "rank_fill", "traceback_fill", "traceback_walk", "run")

    def timings(self) -> tuple[list[float], list[int]]:
        """(ms, launches) per family: seed, chain, rank fill, traceback fill, traceback walk, whole run."""
        ms = (C.c_double * len(self.FAMILIES))()
        ln = (C.c_uint64 * len(self.FAMILIES))()
        _check(_lib.hip().dnaln_hip_batch_timings(self.b, ms, ln), "dnaln_hip_batch_timings")
        return list(ms), list(ln)

    def tune(self) -> dict:
        """Runs the autotuner on the resident batch (dnaln_hip_batch_tune) and returns its report."""
        r = abi.TuneReport()
        _check(_lib.hip().dnaln_hip_batch_tune(self.b, C.byref(r)), "dnaln_hip_batch_tune")
        tried = [{"len_split": (r.cand_split[i][0], r.cand_split[i][1]), "dir_budget": int(r.cand_budget[i]),
                  "walk_kind": r.cand_walk[i], "sw_lanes": r.cand_lanes[i], "ms": r.ms[i]} for i in range(r.n_tried)]
        return {"len_split": (r.len_split[0], r.len_split[1]), "dir_budget": int(r.dir_budget),
                "walk_kind": r.walk_kind, "sw_lanes": r.sw_lanes, "chosen": r.chosen, "tried": tried}

    def close(self) -> None:
        if self.b:
            _lib.hip().dnaln_hip_batch_destroy(self.b)
            self.b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cigar_str(ops) -> str:
    return "".join(f"{int(c) >> 4}{CIGAR_OPS[int(c) & 15]}" for c in ops) or "*"


def record_fields(rec, pool: np.ndarray, contig_names: list[str]) -> tuple:
    """(FLAG, RNAME, POS, MAPQ, CIGAR, NM, AS) as SAM prints them (src/mapper.cpp:231-270)."""
    if rec["contig"] < 0:
        return (4, "*", 0, 0, "*", None, None)
    o = int(rec["cigar_off"])
    return (16 if rec["rev"] else 0, contig_names[int(rec["contig"])], int(rec["pos"]) + 1, int(rec["mapq"]),
            cigar_str(pool[o:o + int(rec["cigar_len"])]), int(rec["nm"]), int(rec["score"]))


def record_lines(names: list[str], recs: np.ndarray, pool: np.ndarray, contig_names: list[str]) -> list[str]:
    out = []
    for i, r in enumerate(recs):
        f = record_fields(r, pool, contig_names)
        out.append("\t".join([names[i], str(f[0]), f[1], str(f[2]), str(f[3]), f[4],
                              "*" if f[5] is None else str(f[5]), "*" if f[6] is None else str(f[6])]))
    return out


class Server:
    """Continuous batching on the resident index (dnaln_serve_*): requests are queued, packed into
    batches as slots and the HBM budget allow, and run concurrently on the slots' streams."""

    def __init__(self, ctx: Context, slots: int = 6, mode: int = abi.MODE_BATCHED, hbm_budget: int = 0,
                 max_batch_reads: int = 1 << 18, max_batch_bases: int = 1 << 27, dir_budget: int = 0):
        self.ctx = ctx
        cfg = abi.ServeConfig()
        _lib.host().dnaln_serve_default_config(C.byref(cfg))
        cfg.slots, cfg.mode, cfg.hbm_budget = slots, mode, hbm_budget
        cfg.max_batch_reads, cfg.max_batch_bases, cfg.dir_budget = max_batch_reads, max_batch_bases, dir_budget
        self.s = C.c_void_p()
        _check(_lib.host().dnaln_serve_create(ctx.ctx, C.byref(cfg), C.byref(self.s)), "dnaln_serve_create")

    def submit(self, rid: int, rb: ReadBatch) -> None:
        _check(_lib.host().dnaln_serve_submit(self.s, rid, _p(rb.codes), _p(rb.offs), _p(rb.lens), _p(rb.fnv), len(rb)),
               "dnaln_serve_submit")

    def progress(self) -> int:
        rc = _lib.host().dnaln_serve_progress(self.s)
        if rc < 0:
            raise DnalnError(rc, "dnaln_serve_progress: " + _lib.host().dnaln_serve_last_error(self.s).decode())
        return rc

    def take(self):
        """(id, records, cigar pool) of one finished request, or None."""
        lib = _lib.host()
        rid, n, nc = C.c_uint64(), C.c_uint64(), C.c_uint64()
        rc = lib.dnaln_serve_take(self.s, C.byref(rid), None, 0, None, 0, C.byref(n), C.byref(nc))
        if rc == abi.E_PENDING:
            return None
        _check(rc, "dnaln_serve_take")
        recs = np.zeros(n.value, abi.RECORD_DTYPE)
        pool = np.zeros(max(nc.value, 1), np.uint32)
        _check(lib.dnaln_serve_take(self.s, C.byref(rid), _p(recs), len(recs), _p(pool), len(pool), C.byref(n),
                                    C.byref(nc)), "dnaln_serve_take")
        return rid.value, recs, pool

    def stats(self) -> dict:
        st = abi.ServeStats()
        _check(_lib.host().dnaln_serve_stats_get(self.s, C.byref(st)), "dnaln_serve_stats_get")
        return {k: int(getattr(st, k)) for k, _ in abi.ServeStats._fields_}

    def batch_bytes(self, n: int, bases: int, maxlen: int) -> int:
        return int(_lib.host().dnaln_serve_batch_bytes(self.s, n, bases, maxlen))

    def run_trace(self, rb: ReadBatch, req_first: np.ndarray, arrival_s: np.ndarray, want_records: bool = False,
                  req_count: np.ndarray | None = None):
        """Replays arrivals in real time; returns (latency_s per request, wall_s, records, cigar pool).
        Request r holds reads [req_first[r], req_first[r] + req_count[r]); without req_count, req_first
        has one more entry and requests are consecutive ranges."""
        n_req = len(arrival_s)
        req_first = np.ascontiguousarray(req_first, np.uint64)
        if req_count is None:
            req_count = np.diff(req_first)
            req_first = req_first[:-1].copy()
        req_count = np.ascontiguousarray(req_count, np.uint64)
        arrival_s = np.ascontiguousarray(arrival_s, np.float64)
        lat = np.zeros(n_req, np.float64)
        wall = C.c_double()
        recs = np.zeros(len(rb), abi.RECORD_DTYPE) if want_records else None
        pool = None
        if want_records:  # CIGAR slots: 2 * len + 4 words per read of every request
            csum = np.concatenate([[0], np.cumsum(2 * rb.lens.astype(np.int64) + 4)])
            a, c = req_first.astype(np.int64), req_count.astype(np.int64)
            pool = np.zeros(int((csum[a + c] - csum[a]).sum()) + 1, np.uint32)
        _check(_lib.host().dnaln_serve_run_trace(self.s, _p(rb.codes), _p(rb.offs), _p(rb.lens), _p(rb.fnv), len(rb),
                                                  _p(req_first), _p(req_count), _p(arrival_s), n_req, _p(lat),
                                                  _p(recs) if want_records else None,
                                                  _p(pool) if want_records else None,
                                                  len(pool) if want_records else 0, C.byref(wall)),
               "dnaln_serve_run_trace")
        return lat, wall.value, recs, pool

    def close(self) -> None:
        if self.s:
            _lib.host().dnaln_serve_destroy(self.s)
            self.s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ accuracy against simulated truth
def _truth_arrays(genome: Genome, reads: Reads):
    """(contig, start within contig, strand) of each read's true locus; start -1 for random reads."""
    pos = np.asarray(reads.truth_pos, np.int64)
    c = np.searchsorted(genome.contig_off.astype(np.int64), np.maximum(pos, 0), side="right").astype(np.int32) - 1
    start = np.where(pos >= 0, pos - genome.contig_off.astype(np.int64)[np.maximum(c, 0)], -1).astype(np.int64)
    return np.ascontiguousarray(c, np.int32), np.ascontiguousarray(start), np.ascontiguousarray(reads.truth_rev,
                                                                                                np.uint8)


def evaluate(recs: np.ndarray, cigars: np.ndarray, reads: Reads, genome: Genome, tol: int = 10) -> dict:
    """Evaluator (src/evaluate.hpp:27-52) over device records of `reads` (made with truth=True
    for base concordance): accuracy = correct / reads in the truth set."""
    c, start, rev = _truth_arrays(genome, reads)
    recs = np.ascontiguousarray(recs, abi.RECORD_DTYPE)
    cig = np.ascontiguousarray(cigars, np.uint32) if len(cigars) else np.zeros(1, np.uint32)
    tc = reads.truth_cigar
    st = abi.EvalStats()
    _check(_lib.host().dnaln_evaluate(_p(recs), _p(cig), len(recs), _p(c), _p(start), _p(rev),
                                      _p(tc) if tc is not None and len(tc) else None,
                                      _p(reads.truth_cig_off) if tc is not None else None, tol, C.byref(st)),
           "dnaln_evaluate")
    mqm, mqw = np.array(st.mq_mapped[:], np.int64), np.array(st.mq_wrong[:], np.int64)
    out = {k: int(getattr(st, k)) for k in ("n_truth", "records", "mapped", "correct", "bases", "bases_ok")}
    out["accuracy"] = out["correct"] / out["n_truth"] if out["n_truth"] else 0.0
    out["base_concordance"] = out["bases_ok"] / out["bases"] if out["bases"] else 0.0
    for q in (60, 30, 10, 1):
        out[f"mapped_mapq{q}"] = int(mqm[q:].sum())
        out[f"wrong_mapq{q}"] = int(mqw[q:].sum())
    return out


def truth_tasks(genome: Genome, reads: Reads, bw: int, limit: int | None = None):
    """Alignment tasks at each read's true locus (src/groundtruth.cpp:11-44): the read in
    reference orientation against [true start - bw, true start + len + bw) of its contig, N
    outside the contig. Returns (pool, tasks, read indices)."""
    c, start, rev = _truth_arrays(genome, reads)
    idx = np.flatnonzero(start >= 0)
    if limit is not None:
        idx = idx[:limit]
    codes = encode(reads.seq)
    comp = np.array([3, 2, 1, 0, 4], np.uint8)
    parts, tasks = [], np.zeros(len(idx), abi.TASK_DTYPE)
    off = 0
    for k, i in enumerate(idx):
        L = int(reads.lens[i])
        q = codes[int(reads.offs[i]):int(reads.offs[i]) + L]
        if rev[i]:
            q = comp[q[::-1]]
        co, cl = int(genome.contig_off[c[i]]), int(genome.contig_len[c[i]])
        lo, wl = int(start[i]) - bw, L + 2 * bw
        t = np.full(wl, 4, np.uint8)
        a, b = max(lo, 0), min(lo + wl, cl)
        if b > a:
            t[a - lo:b - lo] = genome.codes[co + a:co + b]
        tasks[k] = (off, off + L, L, wl)
        parts += [q, t]
        off += L + wl
    pool = np.concatenate(parts) if parts else np.zeros(1, np.uint8)
    return pool, tasks, idx


def truth_check(ctx: "Context", reads: Reads, genome: Genome, limit: int | None = None) -> dict:
    """The true-locus kernel checks of src/bench.cpp:272-314 on the device aligner's traceback
    results: re-scoring, optimality against the true path, base placement."""
    sp = ctx.params.sc
    pool, tasks, idx = truth_tasks(genome, reads, sp.bw, limit)
    res, cig = ctx.align(pool, tasks, traceback=True)
    tc = [reads.truth_cigar[int(reads.truth_cig_off[i]):int(reads.truth_cig_off[i + 1])] for i in idx]
    toff = np.zeros(len(idx) + 1, np.uint64)
    toff[1:] = np.cumsum([len(x) for x in tc])
    tcat = np.concatenate(tc).astype(np.uint32) if len(tc) and toff[-1] else np.zeros(1, np.uint32)
    st = abi.TruthStats()
    _check(_lib.host().dnaln_truth_check(C.byref(sp), _p(pool), _p(tasks), len(tasks), _p(res), _p(cig), _p(tcat),
                                         _p(toff), C.byref(st)), "dnaln_truth_check")
    return {f: int(getattr(st, f)) for f, _ in abi.TruthStats._fields_}
