"""Python access to the C oracle and to the reference build (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this package, and
only as the checker or the timed CPU baseline; the product path never does.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_lib = None
_ref = None


class ScoreT(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("match", "mismatch", "gap_open", "gap_ext", "ambig", "bw")]


class ChainParamsT(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("max_gap", "band", "lookback", "min_score", "min_anchors")]


class MapParamsT(C.Structure):
    _fields_ = [("k", C.c_int), ("w", C.c_int), ("max_occ", C.c_uint32), ("sc", ScoreT), ("ch", ChainParamsT),
                ("max_cand", C.c_int), ("pri_ratio", C.c_double), ("min_aln_score", C.c_int)]


class MiniT(C.Structure):
    _fields_ = [("hash", C.c_uint64), ("pos", C.c_uint32), ("strand", C.c_uint32)]


class BucketT(C.Structure):
    _fields_ = [("key", C.c_uint64), ("offset", C.c_uint32), ("count", C.c_uint32)]


class IndexT(C.Structure):
    _fields_ = [("table", C.POINTER(BucketT)), ("mask", C.c_uint64), ("pos", C.POINTER(C.c_uint32)),
                ("n_pos", C.c_uint64), ("n_keys", C.c_uint64), ("cap", C.c_uint64), ("max_count", C.c_uint32),
                ("k", C.c_int), ("w", C.c_int)]


class AnchorT(C.Structure):
    _fields_ = [("rpos", C.c_uint32), ("qpos", C.c_int32), ("strand", C.c_uint32)]


class ChainT(C.Structure):
    _fields_ = [("score", C.c_int32), ("strand", C.c_uint32), ("n_anchors", C.c_int32), ("dmin", C.c_int64),
                ("dmax", C.c_int64), ("rpos", C.c_uint32)]


class ResultT(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("score", "qb", "qe", "tb", "te", "nm", "n_cigar")]


class RecordT(C.Structure):
    _fields_ = [("contig", C.c_int32), ("pad_", C.c_int32), ("pos", C.c_int64), ("rev", C.c_uint8),
                ("mapq", C.c_uint8), ("pad2_", C.c_uint8 * 2), ("score", C.c_int32), ("nm", C.c_int32),
                ("n_cand", C.c_int32), ("n_cigar", C.c_int32), ("cigar_off", C.c_uint64)]


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = HERE / "liboracle.so"
        if not path.exists():
            raise ImportError(f"{path} is not built (make -C oracle)")
        L = C.CDLL(str(path))
        vp = C.c_void_p
        L.or_hash64.argtypes = [C.c_uint64, C.c_uint64]
        L.or_hash64.restype = C.c_uint64
        L.or_fnv1a.argtypes = [C.c_char_p, C.c_size_t]
        L.or_fnv1a.restype = C.c_uint32
        L.or_sketch.argtypes = [vp, C.c_uint32, C.c_int, C.c_int, C.c_uint32, vp]
        L.or_sketch.restype = C.c_size_t
        L.or_index_build.argtypes = [C.POINTER(IndexT), vp, vp, vp, C.c_int, C.c_int, C.c_int]
        L.or_index_lookup.argtypes = [C.POINTER(IndexT), C.c_uint64, C.POINTER(C.c_uint32)]
        L.or_index_lookup.restype = C.POINTER(C.c_uint32)
        L.or_index_free.argtypes = [C.POINTER(IndexT)]
        L.or_chain.argtypes = [vp, C.c_size_t, C.c_int, C.POINTER(ChainParamsT), vp]
        L.or_chain.restype = C.c_size_t
        L.or_align.argtypes = [C.POINTER(ScoreT), vp, C.c_int, vp, C.c_int, C.c_int, C.POINTER(ResultT), vp, C.c_int]
        L.or_score_cigar.argtypes = [C.POINTER(ScoreT), vp, vp, C.c_int, C.c_int, vp, C.c_int]
        L.or_map.argtypes = [C.POINTER(MapParamsT), vp, vp, vp, C.c_int, C.POINTER(IndexT), vp, vp, vp, vp, vp, vp,
                             C.c_size_t, vp, vp, C.c_size_t, C.POINTER(C.c_uint64)]
        L.or_candidates.argtypes = [C.POINTER(MapParamsT), vp, C.c_int, C.POINTER(IndexT), vp, vp, vp, C.c_size_t,
                                    vp, vp]
        L.or_format_sam.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_char_p, C.c_int, C.POINTER(RecordT), vp,
                                    C.c_char_p, C.c_char_p, C.c_size_t]
        L.or_format_sam.restype = C.c_long
        L.or_default_params.argtypes = [C.POINTER(MapParamsT)]
        _lib = L
    return _lib


def default_params(**kw) -> MapParamsT:
    p = MapParamsT()
    lib().or_default_params(C.byref(p))
    for k, v in kw.items():
        if k in ("match", "mismatch", "gap_open", "gap_ext", "ambig", "bw"):
            setattr(p.sc, k, v)
        elif k in ("max_gap", "band", "lookback", "min_chain", "min_anchors"):
            setattr(p.ch, "min_score" if k == "min_chain" else k, v)
        else:
            setattr(p, k, v)
    return p


def score_t(match=2, mismatch=4, gap_open=4, gap_ext=2, ambig=1, bw=16) -> ScoreT:
    return ScoreT(match, mismatch, gap_open, gap_ext, ambig, bw)


def hash64(key: int, mask: int) -> int:
    return lib().or_hash64(key, mask)


def fnv1a(s: bytes) -> int:
    return lib().or_fnv1a(s, len(s))


def sketch(codes: np.ndarray, k: int, w: int, offset: int = 0) -> np.ndarray:
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    out = (MiniT * max(len(codes), 1))()
    n = lib().or_sketch(codes.ctypes.data, len(codes), k, w, offset, out)
    return np.array([(out[i].hash, out[i].pos, out[i].strand) for i in range(n)],
                    dtype=[("hash", "u8"), ("pos", "u4"), ("strand", "u4")])


class Index:
    def __init__(self, codes: np.ndarray, contig_off: np.ndarray, contig_len: np.ndarray, k: int = 15, w: int = 10):
        self._codes = np.ascontiguousarray(codes, dtype=np.uint8)
        self._off = np.ascontiguousarray(contig_off, dtype=np.uint64)
        self._len = np.ascontiguousarray(contig_len, dtype=np.uint32)
        self.idx = IndexT()
        if lib().or_index_build(C.byref(self.idx), self._codes.ctypes.data, self._off.ctypes.data,
                                self._len.ctypes.data, len(self._len), k, w) != 0:
            raise ValueError("bad k/w")

    def lookup(self, h: int) -> np.ndarray:
        n = C.c_uint32(0)
        p = lib().or_index_lookup(C.byref(self.idx), h, C.byref(n))
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0, np.uint32)

    def table(self) -> np.ndarray:
        buf = np.ctypeslib.as_array(C.cast(self.idx.table, C.POINTER(C.c_uint8)), shape=(self.idx.cap * 16,))
        return buf.view(np.dtype([("key", "u8"), ("offset", "u4"), ("count", "u4")])).copy()

    def positions(self) -> np.ndarray:
        if self.idx.n_pos == 0:
            return np.zeros(0, np.uint32)
        return np.ctypeslib.as_array(self.idx.pos, shape=(self.idx.n_pos,)).copy()

    def __del__(self):
        if _lib is not None and getattr(self, "idx", None) is not None and self.idx.table:
            _lib.or_index_free(C.byref(self.idx))


def chain(anchors: np.ndarray, k: int = 15, max_gap=5000, band=16, lookback=50, min_score=20, min_anchors=2):
    """anchors: structured (rpos u4, qpos i4, strand u4). Returns (sorted anchors, chains)."""
    a = np.ascontiguousarray(anchors, dtype=[("rpos", "u4"), ("qpos", "i4"), ("strand", "u4")]).copy()
    out = (ChainT * max(len(a), 1))()
    p = ChainParamsT(max_gap, band, lookback, min_score, min_anchors)
    n = lib().or_chain(a.ctypes.data, len(a), k, C.byref(p), out)
    chains = [(out[i].score, out[i].strand, out[i].n_anchors, out[i].dmin, out[i].dmax, out[i].rpos) for i in range(n)]
    return a, chains


def align(sp: ScoreT, q: np.ndarray, t: np.ndarray, traceback: bool):
    """Returns (score, qb, qe, tb, te, nm, cigar list)."""
    q = np.ascontiguousarray(q, dtype=np.uint8)
    t = np.ascontiguousarray(t, dtype=np.uint8)
    cap = 2 * len(q) + 4
    cig = np.zeros(cap, dtype=np.uint32)
    r = ResultT()
    lib().or_align(C.byref(sp), q.ctypes.data, len(q), t.ctypes.data, len(t), int(traceback), C.byref(r),
                   cig.ctypes.data, cap)
    return (r.score, r.qb, r.qe, r.tb, r.te, r.nm, [int(x) for x in cig[:max(r.n_cigar, 0)]])


def score_cigar(sp: ScoreT, q, t, qb: int, tb: int, cigar: list[int]) -> int:
    q = np.ascontiguousarray(q, dtype=np.uint8)
    t = np.ascontiguousarray(t, dtype=np.uint8)
    c = np.asarray(cigar, dtype=np.uint32)
    return lib().or_score_cigar(C.byref(sp), q.ctypes.data, t.ctypes.data, qb, tb, c.ctypes.data, len(c))


CIGAR_OPS = "MIDNSHP=X"


def cigar_str(ops) -> str:
    return "".join(f"{int(c) >> 4}{CIGAR_OPS[int(c) & 15]}" for c in ops) or "*"


def map_reads(params: MapParamsT, genome, index: Index, reads):
    """Runs the oracle mapper. Returns (records structured array, cigar pool, cells)."""
    names, noff, nlen = reads.name_arrays()
    n = len(reads)
    recs = (RecordT * max(n, 1))()
    cap = int(reads.lens.astype(np.int64).sum()) * 2 + 4 * n + 16
    pool = np.zeros(cap, dtype=np.uint32)
    cells = C.c_uint64(0)
    rc = lib().or_map(C.byref(params), genome.codes.ctypes.data, genome.contig_off.ctypes.data,
                      genome.contig_len.ctypes.data, genome.n_contigs, C.byref(index.idx), reads.seq.ctypes.data,
                      reads.offs.ctypes.data, reads.lens.ctypes.data, names, noff.ctypes.data, nlen.ctypes.data, n,
                      recs, pool.ctypes.data, cap, C.byref(cells))
    if rc != 0:
        raise RuntimeError("oracle map failed (CIGAR pool too small)")
    arr = np.ctypeslib.as_array(recs, shape=(max(n, 1),))[:n].copy()
    return arr, pool, cells.value


def candidates(params: MapParamsT, genome, index: Index, reads):
    """Candidates after seeding and chaining (src/mapper.cpp:14-63): (n_cand[n], cand[n, max_cand, 4])
    with columns rev, contig, diag and the score of the candidate's chain."""
    n = len(reads)
    mc = max(int(params.max_cand), 1)
    ncand = np.zeros(max(n, 1), np.int32)
    cand = np.zeros((max(n, 1), mc, 4), np.int64)
    lib().or_candidates(C.byref(params), genome.contig_off.ctypes.data, genome.n_contigs, C.byref(index.idx),
                        reads.seq.ctypes.data, reads.offs.ctypes.data, reads.lens.ctypes.data, n,
                        ncand.ctypes.data, cand.ctypes.data)
    return ncand[:n], cand[:n]


def record_fields(rec, pool, contig_names) -> tuple:
    """(FLAG, RNAME, POS, MAPQ, CIGAR, NM, AS) as SAM would print them."""
    if rec["contig"] < 0:
        return (4, "*", 0, 0, "*", None, None)
    ops = pool[int(rec["cigar_off"]):int(rec["cigar_off"]) + int(rec["n_cigar"])]
    return (16 if rec["rev"] else 0, contig_names[rec["contig"]], int(rec["pos"]) + 1, int(rec["mapq"]),
            cigar_str(ops), int(rec["nm"]), int(rec["score"]))


# ----------------------------------------------------------------------------- the reference build
def ref_lib():
    """oracle/_ref/libdnaln_ref.so, or None where the reference was not built."""
    global _ref
    if _ref is None:
        path = HERE / "_ref" / "libdnaln_ref.so"
        if not path.exists():
            return None
        L = C.CDLL(str(path))
        vp = C.c_void_p
        L.ref_align.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, C.c_size_t, C.c_int, vp, vp, vp, vp, C.c_size_t]
        L.ref_sketch.argtypes = [vp, C.c_uint32, C.c_int, C.c_int, C.c_uint32, vp, vp, vp]
        L.ref_sketch.restype = C.c_size_t
        L.ref_chain.argtypes = [vp, vp, vp, C.c_size_t, C.c_int, vp, vp]
        L.ref_chain.restype = C.c_size_t
        L.ref_ctx_create.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_int]
        L.ref_ctx_create.restype = vp
        L.ref_ctx_destroy.argtypes = [vp]
        L.ref_index_lookup.argtypes = [vp, C.c_uint64, vp, C.c_uint32]
        L.ref_index_lookup.restype = C.c_uint32
        L.ref_index_stats.argtypes = [vp, vp]
        L.ref_index_stats.restype = C.c_uint64
        L.ref_map.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, C.c_size_t, vp, vp, vp, vp, C.c_size_t, vp]
        _ref = L
    return _ref


def ref_binary() -> Path | None:
    p = HERE / "_ref" / "dnaln"
    return p if p.exists() else None
