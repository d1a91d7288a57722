"""ctypes loading of the in-tree native libraries (no fallback: a missing library is an error)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_DIR = Path(__file__).resolve().parent / "lib"

_host = None
_hip = None


def _load(name: str) -> C.CDLL:
    path = LIB_DIR / name
    if not path.exists():
        raise ImportError(f"{path} is not built; run `python -m dnaln_amd.build` (or __graft_entry__.build())")
    return C.CDLL(str(path), mode=C.RTLD_GLOBAL)


def hip() -> C.CDLL:
    """libdnaln_hip.so: the HIP kernels behind the C-ABI of include/dnaln_hip.h."""
    global _hip
    if _hip is None:
        lib = _load("libdnaln_hip.so")
        from . import abi
        abi.declare(lib)
        _hip = lib
    return _hip


def host() -> C.CDLL:
    """libdnaln_host.so: generator, host index build, SAM formatting."""
    global _host
    if _host is None:
        if (LIB_DIR / "libdnaln_hip.so").exists():
            hip()
        lib = _load("libdnaln_host.so")
        u64, i32, dbl, vp = C.c_uint64, C.c_int, C.c_double, C.c_void_p
        lib.synth_genome.argtypes = [i32, u64, i32, u64, dbl, dbl, dbl, vp, vp, vp]
        lib.synth_genome.restype = i32
        lib.synth_reads_capacity.argtypes = [vp]
        lib.synth_reads_capacity.restype = u64
        lib.synth_reads.argtypes = [vp, vp, vp, i32, vp, vp, vp, vp, vp, vp]
        lib.synth_reads.restype = C.c_int64
        lib.synth_reads_truth.argtypes = [vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, u64]
        lib.synth_reads_truth.restype = C.c_int64
        lib.synth_write_fasta.argtypes = [C.c_char_p, vp, vp, vp, i32]
        lib.synth_write_fasta.restype = i32
        lib.synth_write_fastq.argtypes = [C.c_char_p, vp, vp, vp, u64, u64]
        lib.synth_write_fastq.restype = i32
        from . import abi
        abi.declare(lib, abi.HOST_SYMBOLS)
        abi.declare(lib, abi.SERVE_SYMBOLS)
        _host = lib
    return _host
