"""ctypes mirror of include/dnaln_hip.h and include/dnaln_host.h (layouts must match the C structs)."""
from __future__ import annotations

import ctypes as C

import numpy as np

i32, u32, u64, i64, dbl, vp = C.c_int32, C.c_uint32, C.c_uint64, C.c_int64, C.c_double, C.c_void_p

OK, E_INVALID, E_CAPACITY, E_DEVICE, E_NOMEM, E_INTERNAL, E_PENDING, E_IO = 0, -1, -2, -3, -4, -5, -6, -7
MODE_PLAIN, MODE_BATCHED = 0, 1


class ScoreParams(C.Structure):
    _fields_ = [(n, i32) for n in ("match", "mismatch", "gap_open", "gap_ext", "ambig", "bw")]


class ChainParams(C.Structure):
    _fields_ = [(n, i32) for n in ("max_gap", "band", "lookback", "min_score", "min_anchors")]


class MapParams(C.Structure):
    _fields_ = [("k", i32), ("w", i32), ("max_occ", u32), ("sc", ScoreParams), ("ch", ChainParams),
                ("max_cand", i32), ("pri_ratio", dbl), ("min_aln_score", i32)]


class Contig(C.Structure):
    _fields_ = [("offset", u64), ("len", u32), ("reserved", u32)]


class IndexView(C.Structure):
    _fields_ = [("table", vp), ("n_buckets", u64), ("pos", vp), ("n_pos", u64), ("k", i32), ("w", i32)]


class Index(C.Structure):
    _fields_ = [("view", IndexView), ("n_keys", u64), ("max_count", u32)]


class PipelineStats(C.Structure):
    _fields_ = [(n, u64) for n in ("reads", "mapped", "bases", "sam_bytes", "batches")] + \
               [(n, C.c_double) for n in ("s_wall", "s_input", "s_device", "s_format", "s_write")] + \
               [("parallel_fastq", i32), ("reserved", i32)]


class EvalStats(C.Structure):
    _fields_ = [(n, u64) for n in ("n_truth", "records", "mapped", "correct", "bases", "bases_ok")] + \
               [("mq_mapped", u64 * 61), ("mq_wrong", u64 * 61)]


class TruthStats(C.Structure):
    _fields_ = [(n, u64) for n in ("n", "rescored_ok", "in_band", "optimal_ok", "majority_ok", "bases", "bases_ok")]


class SeedTraffic(C.Structure):
    _fields_ = [(n, u64) for n in ("read_bases", "minimizers", "buckets", "occurrences", "bytes")]


class Task(C.Structure):
    _fields_ = [("q_off", u64), ("t_off", u64), ("qlen", i32), ("tlen", i32)]


class MapStats(C.Structure):
    _fields_ = [(n, u64) for n in ("reads", "mapped", "anchors", "candidates", "cells", "rank_tasks", "tb_tasks")] + \
               [(n, dbl) for n in ("ms_seed", "ms_chain", "ms_rank", "ms_traceback", "ms_total")] + \
               [("global_path_reads", u64), ("fused_retry_reads", u64)]


class BatchConfig(C.Structure):
    _fields_ = [("mode", i32), ("chunk_reads", u32), ("max_reads", u64), ("max_bases", u64),
                ("anchors_per_read", u32), ("tune", i32), ("len_split", i32 * 2), ("dir_budget", u64),
                ("walk_kind", i32), ("graph", i32), ("sw_lanes", i32), ("reserved", i32)]


class TuneReport(C.Structure):
    _fields_ = [("len_split", i32 * 2), ("dir_budget", u64), ("walk_kind", i32), ("n_tried", i32), ("chosen", i32),
                ("sw_lanes", i32), ("ms", dbl * 16), ("cand_split", (i32 * 2) * 16), ("cand_budget", u64 * 16),
                ("cand_walk", i32 * 16), ("cand_lanes", i32 * 16)]


CAND_DTYPE = np.dtype([("rev", "<u4"), ("contig", "<i4"), ("diag", "<i8"), ("chain_score", "<i4"), ("reserved", "<i4")])
TASK_DTYPE = np.dtype([("q_off", "<u8"), ("t_off", "<u8"), ("qlen", "<i4"), ("tlen", "<i4")])
RESULT_DTYPE = np.dtype([("score", "<i4"), ("qb", "<i4"), ("qe", "<i4"), ("tb", "<i4"), ("te", "<i4"), ("nm", "<i4"),
                         ("cigar_off", "<u8"), ("cigar_len", "<u4"), ("reserved", "<u4")])
RECORD_DTYPE = np.dtype([("contig", "<i4"), ("rev", "u1"), ("mapq", "u1"), ("reserved", "<u2"), ("pos", "<i8"),
                         ("score", "<i4"), ("nm", "<i4"), ("n_cand", "<i4"), ("cigar_len", "<u4"),
                         ("cigar_off", "<u8")])
assert CAND_DTYPE.itemsize == 24 and TASK_DTYPE.itemsize == 24 and RESULT_DTYPE.itemsize == 40 and RECORD_DTYPE.itemsize == 40
assert C.sizeof(MapParams) == 80 and C.sizeof(BatchConfig) == 64 and C.sizeof(MapStats) == 112

# Every symbol include/dnaln_hip.h declares, with (restype, argtypes).
HIP_SYMBOLS = {
    "dnaln_hip_default_params": (None, [vp]),
    "dnaln_hip_create": (i32, [vp, vp, u64, vp, i32, vp, i32, vp]),
    "dnaln_hip_destroy": (None, [vp]),
    "dnaln_hip_last_error": (C.c_char_p, []),
    "dnaln_hip_align": (i32, [vp, vp, vp, u64, vp, u64, i32, vp, vp, u64]),
    "dnaln_hip_align_static": (i32, [vp, vp, vp, u64, vp, u64, i32, vp, vp, u64]),
    "dnaln_hip_map_batch": (i32, [vp, vp, vp, vp, vp, u64, vp, vp, u64, vp]),
    "dnaln_hip_default_batch_config": (None, [vp]),
    "dnaln_hip_batch_create": (i32, [vp, vp, vp]),
    "dnaln_hip_batch_destroy": (None, [vp]),
    "dnaln_hip_batch_upload": (i32, [vp, vp, vp, vp, vp, u64]),
    "dnaln_hip_batch_run": (i32, [vp]),
    "dnaln_hip_batch_poll": (i32, [vp, i32]),
    "dnaln_hip_batch_download": (i32, [vp, vp, vp, u64, vp]),
    "dnaln_hip_batch_device_bytes": (u64, [vp, vp]),
    "dnaln_hip_batch_timings": (i32, [vp, vp, vp]),
    "dnaln_hip_batch_tune": (i32, [vp, vp]),
    "dnaln_hip_device_memory": (i32, [vp, vp, vp]),
    "dnaln_hip_batch_reserve": (i32, [vp, u64, u64, i32, u32]),
    "dnaln_hip_context_params": (i32, [vp, vp]),
    "dnaln_hip_alloc_counters": (None, [vp, vp]),
    "dnaln_hip_peak_int16": (i32, [i32, i32, vp, vp]),
    "dnaln_hip_index_build": (i32, [vp, u64, vp, i32, i32, i32, i32, vp]),
    "dnaln_hip_index_info": (i32, [vp, vp, vp, vp, vp, vp]),
    "dnaln_hip_index_download": (i32, [vp, vp, vp]),
    "dnaln_hip_index_free": (None, [vp]),
    "dnaln_hip_context_index": (i32, [vp, vp, vp, vp]),
    "dnaln_hip_host_alloc": (i32, [u64, vp]),
    "dnaln_hip_host_free": (None, [vp]),
    "dnaln_hip_batch_candidates": (i32, [vp, vp, vp, u64]),
}

class ServeConfig(C.Structure):
    _fields_ = [("slots", i32), ("mode", i32), ("hbm_budget", u64), ("max_batch_reads", u64),
                ("max_batch_bases", u64), ("dir_budget", u64), ("max_read_len", i32), ("reserved", i32)]


class ServeStats(C.Structure):
    _fields_ = [(n, u64) for n in ("submitted", "completed", "batches", "reads", "queued_requests",
                                   "inflight_batches", "max_inflight_batches", "inflight_bytes",
                                   "max_inflight_bytes", "budget_bytes", "deferred_by_budget")]


SERVE_SYMBOLS = {
    "dnaln_serve_default_config": (None, [vp]),
    "dnaln_serve_create": (i32, [vp, vp, vp]),
    "dnaln_serve_destroy": (None, [vp]),
    "dnaln_serve_batch_bytes": (u64, [vp, u64, u64, i32]),
    "dnaln_serve_submit": (i32, [vp, u64, vp, vp, vp, vp, u64]),
    "dnaln_serve_progress": (i32, [vp]),
    "dnaln_serve_take": (i32, [vp, vp, vp, u64, vp, u64, vp, vp]),
    "dnaln_serve_stats_get": (i32, [vp, vp]),
    "dnaln_serve_last_error": (C.c_char_p, [vp]),
    "dnaln_serve_run_trace": (i32, [vp, vp, vp, vp, vp, u64, vp, vp, vp, u64, vp, vp, vp, u64, vp]),
}

HOST_SYMBOLS = {
    "dnaln_index_build": (i32, [vp, vp, i32, i32, i32, vp]),
    "dnaln_index_free": (None, [vp]),
    "dnaln_index_build_gpu": (i32, [vp, vp, i32, i32, i32, i32, vp, vp]),
    "dnaln_index_save": (i32, [vp, vp, vp, i32, C.c_char_p]),
    "dnaln_index_load": (i32, [C.c_char_p, vp, vp, i32, vp]),
    "dnaln_reference_digest": (u64, [vp, vp, i32]),
    "dnaln_seed_traffic": (i32, [vp, vp, vp, vp, u64, u32, i32, vp]),
    "dnaln_map_file": (i32, [vp, vp, C.c_char_p, i32, vp, i32, u64, i32, vp]),
    "dnaln_map_file_on": (i32, [vp, vp, C.c_char_p, i32, vp, i32, u64, i32, vp]),
    "dnaln_map_encoded": (i32, [vp, i32, vp, vp, vp, vp, u64, u64, vp, vp, u64, vp, vp]),
    "dnaln_host_last_error": (C.c_char_p, []),
    "dnaln_evaluate": (i32, [vp, vp, u64, vp, vp, vp, vp, vp, i32, vp]),
    "dnaln_truth_check": (i32, [vp, vp, vp, u64, vp, vp, vp, vp, vp]),
    "dnaln_fnv1a": (u32, [C.c_char_p, C.c_size_t]),
    "dnaln_fnv1a_many": (None, [vp, vp, vp, u64, vp]),
    "dnaln_encode": (None, [C.c_char_p, C.c_size_t, vp]),
    "dnaln_sam_format": (C.c_long, [C.c_char_p, C.c_size_t, C.c_char_p, C.c_char_p, C.c_size_t, vp, vp, C.c_char_p,
                                    vp, C.c_size_t]),
}


def declare(lib: C.CDLL, table: dict | None = None) -> None:
    for name, (res, args) in (table or HIP_SYMBOLS).items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
