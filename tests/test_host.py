"""CPU checks of the boundary and the host side: every declared C symbol is exported, the ctypes
layouts match, the host index equals the oracle's, SAM text matches the reference digest, and the
generator is deterministic."""
from __future__ import annotations

import ctypes as C
import hashlib
import re
from pathlib import Path

import numpy as np
import pytest

import oracle
from dnaln_amd import _lib, abi, api, synth
from tests import golden_io
from tests.golden import sets

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: str) -> set[str]:
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(dnaln_[a-z0-9_]+)\s*\(", text))


def test_hip_library_exports_every_declared_symbol():
    lib = _lib.hip()
    declared = _declared("dnaln_hip.h")
    assert declared == set(abi.HIP_SYMBOLS), declared ^ set(abi.HIP_SYMBOLS)
    for name in declared:
        assert hasattr(lib, name), name


def test_host_library_exports_every_declared_symbol():
    lib = _lib.host()
    declared = _declared("dnaln_host.h")
    assert declared == set(abi.HOST_SYMBOLS), declared ^ set(abi.HOST_SYMBOLS)
    for name in declared:
        assert hasattr(lib, name), name


def test_default_params_match_reference_defaults():
    p = api.map_params()
    assert (p.k, p.w, p.max_occ, p.max_cand, p.min_aln_score, p.pri_ratio) == (15, 10, 500, 10, 40, 0.5)
    assert (p.sc.match, p.sc.mismatch, p.sc.gap_open, p.sc.gap_ext, p.sc.ambig, p.sc.bw) == (2, 4, 4, 2, 1, 16)
    assert (p.ch.max_gap, p.ch.band, p.ch.lookback, p.ch.min_score, p.ch.min_anchors) == (5000, 16, 50, 20, 2)


def test_create_rejects_out_of_range_parameters_without_touching_the_device():
    g = synth.make_genome(0, 50_000, 1, 3)
    idx = api.Index(g)
    for bad in ({"k": 3}, {"bw": 0}, {"bw": 1001}, {"match": 0}, {"gap_ext": 0}, {"mismatch": -1}):
        with pytest.raises(api.DnalnError) as e:
            api.Context(g, idx, **bad)
        assert e.value.code == abi.E_INVALID


def test_generator_is_deterministic_and_follows_the_error_model():
    g1 = synth.make_genome(0, 100_000, 2, 5)
    g2 = synth.make_genome(0, 100_000, 2, 5)
    assert np.array_equal(g1.codes, g2.codes)
    assert abs(np.bincount(g1.codes, minlength=4) / 1e5 - 0.25).max() < 0.01
    r1 = synth.make_reads(g1, 2000, synth.LEN_FIXED, 150, seed=9)
    r2 = synth.make_reads(g1, 2000, synth.LEN_FIXED, 150, seed=9)
    assert np.array_equal(r1.seq, r2.seq) and np.array_equal(r1.truth_pos, r2.truth_pos)
    assert set(r1.lens.tolist()) == {150}
    assert 0.45 < r1.truth_rev.mean() < 0.55
    # reads stay inside their contig
    for i in range(len(r1)):
        c = int(np.searchsorted(g1.contig_off, r1.truth_pos[i], side="right")) - 1
        assert r1.truth_pos[i] + 140 <= g1.contig_off[c] + g1.contig_len[c]
    lg = synth.make_reads(g1, 3000, synth.LEN_LOGUNIFORM, 50, 1000, seed=3)
    assert lg.lens.min() >= 50 and lg.lens.max() <= 1000 and 250 < lg.lens.mean() < 400


def test_repeat_genome_has_segmental_copies():
    g = synth.make_genome(1, 2_000_000, 4, 7, 0.3, 0.0, 0.0)
    idx = oracle.Index(g.codes, g.contig_off, g.contig_len)
    # exact copies make many minimizers occur more than once
    assert idx.idx.n_pos > 1.15 * idx.idx.n_keys


@pytest.mark.parametrize("spec", [sets.UNIFORM_4M, sets.REPEATS_8M])
def test_host_index_equals_oracle_index(spec):
    g = sets.genome({"genome": spec})
    for k, w in ((15, 10), (13, 7)):
        mine = api.Index(g, k, w)
        ref = oracle.Index(g.codes, g.contig_off, g.contig_len, k, w)
        assert mine.n_keys == ref.idx.n_keys and mine.n_pos == ref.idx.n_pos
        t, pos = mine.table(), mine.positions()
        assert len(t) == ref.idx.cap
        assert np.array_equal(t, ref.table())  # identical layout: same keys, same probe order
        assert np.array_equal(pos, ref.positions())


def test_sam_format_matches_reference_digest():
    """The host SAM writer reproduces the reference's SAM body byte for byte from the oracle's
    records (the records themselves are checked on the GPU)."""
    name = "edge_repeats_bw32"
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    idx = oracle.Index(g.codes, g.contig_off, g.contig_len)
    recs, pool, _ = oracle.map_reads(oracle.default_params(**sets.map_params(spec)), g, idx, r)
    h = hashlib.sha256()
    buf = C.create_string_buffer(1 << 16)
    host = _lib.host()
    for i in range(len(r)):
        m = np.zeros(1, abi.RECORD_DTYPE)
        m["contig"], m["rev"], m["mapq"], m["pos"] = recs[i]["contig"], recs[i]["rev"], recs[i]["mapq"], recs[i]["pos"]
        m["score"], m["nm"], m["n_cand"] = recs[i]["score"], recs[i]["nm"], recs[i]["n_cand"]
        m["cigar_len"], m["cigar_off"] = recs[i]["n_cigar"], recs[i]["cigar_off"]
        s = r.read(i)
        nm = f"r{i}".encode()
        cname = g.names[int(recs[i]["contig"])].encode() if recs[i]["contig"] >= 0 else None
        n = host.dnaln_sam_format(nm, len(nm), s, b"I" * len(s), len(s), m.ctypes.data, pool.ctypes.data, cname,
                                  C.cast(buf, C.c_void_p), len(buf))
        h.update(buf.raw[:n])
    assert h.hexdigest() == golden_io.manifest()["sets"][name]["sam_body_sha256"]


def test_fnv1a_many_matches_oracle():
    names = [f"r{i}".encode() for i in range(1000)] + [b"", b"read/1 with space"]
    blob = b"".join(names)
    lens = np.array([len(x) for x in names], np.int32)
    offs = np.concatenate([[0], np.cumsum(lens[:-1])]).astype(np.uint64)
    out = np.zeros(len(names), np.uint32)
    nb = np.frombuffer(blob, np.uint8)
    _lib.host().dnaln_fnv1a_many(nb.ctypes.data, offs.ctypes.data, lens.ctypes.data, len(names), out.ctypes.data)
    assert out.tolist() == [oracle.fnv1a(x) for x in names]


def test_serving_library_exports_every_declared_symbol():
    lib = _lib.host()
    declared = _declared("dnaln_serve.h")
    assert declared == set(abi.SERVE_SYMBOLS), declared ^ set(abi.SERVE_SYMBOLS)
    for name in declared:
        assert hasattr(lib, name), name
    cfg = abi.ServeConfig()
    lib.dnaln_serve_default_config(C.byref(cfg))
    assert cfg.slots >= 1 and cfg.mode == abi.MODE_BATCHED and cfg.max_batch_reads >= 4096


def test_request_trace_follows_the_serving_model():
    first, arrive = synth.request_trace(200_000, 1000.0, 3)
    sizes = np.diff(first.astype(np.int64))
    assert sizes.min() >= 1 and sizes.max() <= 4096 and np.all(np.diff(arrive) >= 0)
    # log-uniform sizes: mean (4097 - 1) / ln 4097 = 492
    assert 380 < sizes.mean() < 620
    # Poisson arrivals at 1000/s: mean gap 1 ms
    assert 0.8e-3 < np.diff(arrive).mean() < 1.25e-3


def test_seed_traffic_model_counts_match_the_oracle():
    """dnaln_seed_traffic against a count made from the oracle: its sketch of every read, and for
    each minimizer the buckets linear probing visits in the oracle's table (the home slot through
    the matching or the first empty one) and the occurrences of keys with 1 <= count <= max_occ.
    A repeat-rich genome makes probe chains and multi-occurrence keys common."""
    g = synth.make_genome(0, 300_000, 3, 77, repeat_frac=0.3)
    r = synth.make_reads(g, 400, synth.LEN_UNIFORM, 20, 400, 0.02, 0.003, 6, 0.5, 78, n_rate=0.003)
    rb = api.prepare(r)
    k, w, max_occ = 15, 10, 3
    ix = api.Index(g, k, w)
    got = ix.seed_traffic(rb, max_occ=max_occ, threads=3)
    oi = oracle.Index(g.codes, g.contig_off, g.contig_len, k, w)
    tab = oi.table()
    mask = len(tab) - 1
    empty = np.uint64(0xFFFFFFFFFFFFFFFF)
    mins = buckets = occ = 0
    for i in range(len(rb)):
        o, L = int(rb.offs[i]), int(rb.lens[i])
        for m in oracle.sketch(rb.codes[o:o + L], k, w):
            mins += 1
            s = int(m["hash"]) & mask
            while True:
                buckets += 1
                if tab["key"][s] == m["hash"]:
                    c = int(tab["count"][s])
                    occ += c if 1 <= c <= max_occ else 0
                    break
                if tab["key"][s] == empty:
                    break
                s = (s + 1) & mask
    bases = int(rb.lens.astype(np.int64).sum())
    assert (got["minimizers"], got["buckets"], got["occurrences"], got["read_bases"]) == (mins, buckets, occ, bases)
    assert got["bytes"] == bases + 16 * buckets + 4 * occ
    assert occ > 0 and buckets > mins  # the case exercises multi-occurrence keys and probe chains


def test_map_encoded_refuses_bad_arguments_without_a_device():
    """dnaln_map_encoded checks its batches before touching a device: none, null ones or the same
    batch twice give DNALN_E_INVALID with a message."""
    import ctypes as C
    from dnaln_amd import _lib, abi
    lib = _lib.host()
    used = C.c_uint64()
    for arr, n in (((C.c_void_p * 2)(None, None), 2), ((C.c_void_p * 1)(None), 0)):
        rc = lib.dnaln_map_encoded(arr, n, None, None, None, None, 0, 0, None, None, 0, C.byref(used), None)
        assert rc == abi.E_INVALID
        assert lib.dnaln_host_last_error()
    fake = (C.c_void_p * 2)(C.c_void_p(8), C.c_void_p(8))  # never dereferenced: refused as a duplicate
    assert lib.dnaln_map_encoded(fake, 2, None, None, None, None, 0, 0, None, None, 0, C.byref(used), None) \
        == abi.E_INVALID
