"""The GPU index build's decomposition and the index file, on the CPU.

k_ix_count / k_ix_emit cut every contig into segments of k-mer starts and sketch each segment on
its own (sketch_segment). Run through the CPU build of that code at every segment length, the
occurrence list must equal the oracle's MinimizerIndex (src/index.cpp:10-45), including contigs
shorter than k or than one window, N runs, palindromic k-mers (even k) and w = 1. The index file
must round-trip exactly and refuse a truncated, corrupt or foreign file."""
from __future__ import annotations

import ctypes as C
import shutil

import numpy as np
import pytest

import oracle
from dnaln_amd import abi, api, synth
from tests.golden import sets
from tests.test_emu import emu  # noqa: F401  (fixture)

EMPTY = np.uint64(0xFFFFFFFFFFFFFFFF)


def occ_keys(table: np.ndarray, pos: np.ndarray) -> np.ndarray:
    """hash << 32 | pos word for every occurrence, in the table's (hash, pos) order."""
    t = table[table["key"] != EMPTY]
    t = t[np.argsort(t["offset"], kind="stable")]
    h = np.repeat(t["key"].astype(np.uint64), t["count"].astype(np.int64))
    assert len(h) == len(pos)
    return h << np.uint64(32) | pos.astype(np.uint64)


def ragged_genome(seed: int) -> synth.Genome:
    """Contigs of 3 .. 70,000 bp around the k and window boundaries, with N runs and single Ns."""
    rng = np.random.default_rng(seed)
    lens = [3, 14, 15, 16, 24, 25, 26, 33, 40, 1023, 1024, 1025, 1500, 4096, 70_000]
    codes = [rng.integers(0, 4, n, dtype=np.uint8) for n in lens]
    for c in codes:
        if len(c) > 100:
            for _ in range(len(c) // 700 + 1):
                a = int(rng.integers(0, len(c)))
                c[a:a + int(rng.integers(1, 40))] = 4
    codes[-1][0] = 4  # a contig that starts with N
    codes[-2][-1] = 4  # and one that ends with N
    # low-complexity stretches: runs of one base and of short periods (tied hashes in a window)
    codes[-1][20_000:20_400] = 2
    codes[-1][30_000:30_600] = np.tile(np.array([0, 1], np.uint8), 300)
    codes[-1][40_000:40_640] = np.tile(np.array([0, 0, 3, 3], np.uint8), 160)
    off = np.cumsum([0] + lens[:-1]).astype(np.uint64)
    return synth.Genome(np.concatenate(codes), off, np.array(lens, np.uint32), [f"c{i}" for i in range(len(lens))])


def emu_occ(emu, g: synth.Genome, k: int, w: int, seg: int) -> np.ndarray:
    v = C.c_void_p
    off = g.contig_off.astype(np.int64)
    ln = g.contig_len.astype(np.int64)
    cap = len(g.codes) + 16
    out = np.zeros(cap, np.uint64)
    emu.emu_index_occ.restype = C.c_long
    n = emu.emu_index_occ(v(g.codes.ctypes.data), v(off.ctypes.data), v(ln.ctypes.data), C.c_int(len(ln)),
                          C.c_int(k), C.c_int(w), C.c_uint32(seg), v(out.ctypes.data), C.c_size_t(cap))
    assert n >= 0
    return np.sort(out[:n])


@pytest.mark.parametrize("k,w", [(15, 10), (16, 10), (11, 10), (15, 5), (12, 1), (5, 17), (15, 25), (4, 2)])
@pytest.mark.parametrize("seg", [1, 7, 64, 1024])
def test_segmented_sketch_equals_oracle_index(emu, k, w, seg):
    g = ragged_genome(5)
    ref = oracle.Index(g.codes, g.contig_off, g.contig_len, k, w)
    want = occ_keys(ref.table(), ref.positions())
    got = emu_occ(emu, g, k, w, seg)
    assert len(got) == len(want), (len(got), len(want))
    assert np.array_equal(got, want)


@pytest.mark.parametrize("spec", [sets.UNIFORM_4M, sets.REPEATS_8M])
def test_segmented_sketch_equals_oracle_on_benchmark_genomes(emu, spec):
    g = sets.genome({"genome": spec})
    ref = oracle.Index(g.codes, g.contig_off, g.contig_len)
    assert np.array_equal(emu_occ(emu, g, 15, 10, 1024), occ_keys(ref.table(), ref.positions()))


def test_occurrence_keys_pin_the_oracle_layout():
    """occ_keys gives every (hash, pos) once, sorted: the order the device sort must reproduce."""
    g = ragged_genome(6)
    ref = oracle.Index(g.codes, g.contig_off, g.contig_len, 15, 10)
    keys = occ_keys(ref.table(), ref.positions())
    assert len(np.unique(keys)) == len(keys) and np.all(np.diff(keys.astype(np.float64)) >= 0)
    # every occurrence is an in-bounds, N-free k-mer whose hash is the table key
    p = (keys & np.uint64(0xFFFFFFFF)) >> np.uint64(1)
    assert np.all(p + 15 <= len(g.codes))


# ---------------------------------------------------------------------------------- index file
def test_index_file_round_trip(tmp_path):
    g = ragged_genome(7)
    a = api.Index(g, 13, 7)
    path = tmp_path / "a.dnx"
    a.save(path)
    b = api.Index.load(path, g)
    assert (b.k, b.w, b.n_keys, b.n_pos, b.max_count, b.n_buckets) == (13, 7, a.n_keys, a.n_pos, a.max_count,
                                                                       a.n_buckets)
    assert np.array_equal(a.table(), b.table()) and np.array_equal(a.positions(), b.positions())
    assert path.stat().st_size == 64 + 16 * a.n_buckets + 4 * a.n_pos


def _expect_io_error(path, g):
    with pytest.raises(api.DnalnError) as e:
        api.Index.load(path, g)
    assert e.value.code == abi.E_IO


def test_index_file_refuses_other_codes_truncation_and_corruption(tmp_path):
    g = ragged_genome(8)
    a = api.Index(g)
    path = tmp_path / "a.dnx"
    a.save(path)
    _expect_io_error(tmp_path / "missing.dnx", g)
    # another reference: one base changed, or the same codes cut into other contigs
    other = synth.Genome(g.codes.copy(), g.contig_off, g.contig_len, g.names)
    other.codes[12345] = (other.codes[12345] + 1) % 4
    _expect_io_error(path, other)
    recut = synth.Genome(g.codes, g.contig_off.copy(), g.contig_len.copy(), g.names)
    recut.contig_len[-1] -= 1
    _expect_io_error(path, recut)
    raw = path.read_bytes()
    cases = {
        "truncated": raw[:-4],
        "extended": raw + b"\0" * 4,
        "magic": b"X" + raw[1:],
        "version": raw[:8] + (2).to_bytes(4, "little") + raw[12:],
    }
    t = a.table()
    used = np.flatnonzero(t["key"] != EMPTY)
    body = bytearray(raw)
    slot = int(used[0])
    bad_off = a.n_pos  # slice past the end of pos[]
    body[64 + 16 * slot + 8:64 + 16 * slot + 12] = int(bad_off).to_bytes(4, "little")
    cases["slice"] = bytes(body)
    body = bytearray(raw)
    pos0 = 64 + 16 * a.n_buckets
    body[pos0:pos0 + 4] = (len(g.codes) << 1).to_bytes(4, "little")  # a position past the reference
    cases["position"] = bytes(body)
    body = bytearray(raw)
    empty = int(np.flatnonzero(t["key"] == EMPTY)[0])
    body[64 + 16 * empty:64 + 16 * empty + 16] = raw[64 + 16 * slot:64 + 16 * slot + 16]  # key count off by one
    cases["keys"] = bytes(body)
    for name, data in cases.items():
        p = tmp_path / f"{name}.dnx"
        p.write_bytes(data)
        _expect_io_error(p, g)
    # the untouched file still loads
    shutil.copy(path, tmp_path / "ok.dnx")
    assert api.Index.load(tmp_path / "ok.dnx", g).n_keys == a.n_keys


def test_reference_digest_depends_on_codes_and_contigs():
    g = ragged_genome(9)
    from dnaln_amd import _lib

    def digest(gg):
        arr = api.contig_array(gg)
        return _lib.host().dnaln_reference_digest(gg.codes.ctypes.data, C.cast(arr, C.c_void_p), len(arr))

    d0 = digest(g)
    assert d0 == digest(synth.Genome(g.codes.copy(), g.contig_off, g.contig_len, g.names))
    c = g.codes.copy()
    c[-1] = (c[-1] + 1) % 5
    assert digest(synth.Genome(c, g.contig_off, g.contig_len, g.names)) != d0
