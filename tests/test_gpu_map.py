"""GPU parity of Mapper::map through the C ABI: every golden record set written by the reference
binary, in the plain and batched forms, and batch invariance (same record alone and in any chunk)."""
from __future__ import annotations

import numpy as np
import pytest

from dnaln_amd import abi, api
from tests import golden_io
from tests.golden import sets

pytestmark = pytest.mark.gpu

_CTX: dict = {}


def _context(spec):
    mp = sets.map_params(spec)
    key = (id(sets.genome(spec)), tuple(sorted(mp.items())))
    if key not in _CTX:
        _CTX[key] = api.Context(sets.genome(spec), **mp)
    return _CTX[key]


def _compare(name, recs, pool, g):
    gold = golden_io.records(name)
    bad = [i for i in range(len(gold)) if api.record_fields(recs[i], pool, g.names) != gold[i][1:]]
    assert not bad, (name, len(bad), [(i, api.record_fields(recs[i], pool, g.names), gold[i][1:]) for i in bad[:3]])


@pytest.mark.parametrize("name", list(sets.SETS))
def test_map_matches_reference_records(name):
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    print(f"set {name} n {len(r)} flags {spec['flags']}", flush=True)
    ctx = _context(spec)
    recs, pool, stats = ctx.map(r)
    _compare(name, recs, pool, g)
    assert stats.reads == len(r)


@pytest.mark.parametrize("name", list(sets.SETS))
def test_batched_form_matches_reference_records(name):
    """The batched form (fused LDS seed+chain, overflow reads on the global path) gives the
    reference's records too, and reports how many reads took the global path."""
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    print(f"set {name} batched", flush=True)
    b = _context(spec).batch(mode=abi.MODE_BATCHED)
    b.upload(api.prepare(r))
    b.run()
    recs, pool, stats = b.download()
    b.close()
    _compare(name, recs, pool, g)
    assert stats.global_path_reads <= len(r) and stats.fused_retry_reads <= len(r)
    print(f"global path {stats.global_path_reads}, second fused pass {stats.fused_retry_reads}", flush=True)


@pytest.mark.parametrize("name", ["uniform4m_150bp_10k", "edge_repeats_bw32", "edge_twins_bw16"])
@pytest.mark.parametrize("mode,chunk", [(abi.MODE_PLAIN, 1000), (abi.MODE_PLAIN, 777), (abi.MODE_BATCHED, 0)])
def test_map_is_batch_invariant(name, mode, chunk):
    """Each read's record depends only on the read: any chunking gives the reference's records."""
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    print(f"set {name} mode {mode} chunk {chunk}", flush=True)
    ctx = _context(spec)
    b = ctx.batch(mode=mode, chunk_reads=chunk or 65536)
    b.upload(api.prepare(r))
    b.run()
    recs, pool, _ = b.download()
    b.close()
    _compare(name, recs, pool, g)


def test_map_subsets_equal_full_batch():
    spec = sets.SETS["edge_twins_bw16"]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    ctx = _context(spec)
    full, fpool, _ = ctx.map(r)
    for lo, hi in ((0, 1), (17, 18), (100, 613), (5000, 6000)):
        part, ppool, _ = ctx.map(r.subset(lo, hi))
        for i in range(hi - lo):
            assert api.record_fields(part[i], ppool, g.names) == api.record_fields(full[lo + i], fpool, g.names)


def test_map_empty_batch():
    spec = sets.SETS["uniform4m_150bp_10k"]
    g = sets.genome(spec)
    r = sets.reads(spec, g).subset(0, 0)
    recs, pool, stats = _context(spec).map(r)
    assert len(recs) == 0 and stats.reads == 0


@pytest.mark.parametrize("name", ["edge_uniform_bw16", "edge_repeats_params", "edge_twins_sr_bw5"])
@pytest.mark.parametrize("split", [(0, 0), (0, -1), (-1, 0), (-1, -1), (60, 120)])
def test_every_length_bucketing_gives_reference_records(name, split):
    """Whichever reads chain one per lane in the fused LDS kernel (and in which shape) and which
    one per wave through global memory, the records are the reference's."""
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    print(f"set {name} split {split}", flush=True)
    b = _context(spec).batch(mode=abi.MODE_BATCHED, len_split=split, dir_budget=1 << 24)
    b.upload(api.prepare(r))
    b.run()
    recs, pool, _ = b.download()
    b.close()
    _compare(name, recs, pool, g)


def test_autotuned_batch_gives_reference_records():
    spec = sets.SETS["edge_uniform_bw64"]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    b = _context(spec).batch(mode=abi.MODE_BATCHED)
    b.upload(api.prepare(r))
    rep = b.tune()
    print("tune", rep, flush=True)
    assert 1 <= len(rep["tried"]) <= 16 and 0 <= rep["chosen"] < len(rep["tried"])
    assert min(t["ms"] for t in rep["tried"]) == rep["tried"][rep["chosen"]]["ms"]
    b.run()
    recs, pool, _ = b.download()
    b.close()
    _compare("edge_uniform_bw64", recs, pool, g)


@pytest.mark.parametrize("name", ["uniform4m_150bp_10k", "edge_uniform_bw64", "edge_twins_sr_bw5", "edge_repeats_bw32"])
@pytest.mark.parametrize("walk_kind", [0, 1])
@pytest.mark.parametrize("mode", [abi.MODE_PLAIN, abi.MODE_BATCHED])
def test_both_traceback_walks_give_reference_records(name, walk_kind, mode):
    """One thread per task reading flag words from HBM, or one block per fill wave through LDS
    windows: the same CIGARs, coordinates and NM."""
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    print(f"set {name} walk {walk_kind} mode {mode}", flush=True)
    b = _context(spec).batch(mode=mode, walk_kind=walk_kind, dir_budget=1 << 22)
    b.upload(api.prepare(r))
    b.run()
    recs, pool, _ = b.download()
    b.close()
    _compare(name, recs, pool, g)


@pytest.mark.parametrize("name,lanes", [("edge_uniform_bw64", 4), ("edge_repeats_bw32", 2)])
@pytest.mark.parametrize("sw_lanes", [1, 0])
def test_split_fill_gives_reference_records(name, lanes, sw_lanes):
    """The packed fill with one pair of alignments per lane, and split over 2 (bw 32) or 4
    (bw 64) lanes per pair, the neighbours exchanging their edge cells by DPP; a small direction
    budget cuts the traceback fill into several launches. The batched form's default is the split."""
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    print(f"set {name} sw_lanes {sw_lanes}", flush=True)
    b = _context(spec).batch(mode=abi.MODE_BATCHED, sw_lanes=sw_lanes or lanes, dir_budget=1 << 22)
    b.upload(api.prepare(r))
    b.run()
    recs, pool, _ = b.download()
    b.close()
    _compare(name, recs, pool, g)


def test_split_lanes_the_band_lacks_are_refused():
    spec = sets.SETS["edge_uniform_bw16"]
    with pytest.raises(api.DnalnError):
        _context(spec).batch(mode=abi.MODE_BATCHED, sw_lanes=2)


def _oracle_candidates(spec, g, r):
    import oracle
    mp = sets.map_params(spec)
    idx = oracle.Index(g.codes, g.contig_off, g.contig_len, mp["k"], mp["w"])
    return oracle.candidates(oracle.default_params(**mp), g, idx, r)


@pytest.mark.parametrize("name", ["edge_uniform_bw64", "edge_uniform_bw16", "edge_repeats_bw32", "edge_repeats_bwa_bw8",
                                  "edge_repeats_params"])
@pytest.mark.parametrize("mode,split", [(abi.MODE_PLAIN, (0, 0)), (abi.MODE_BATCHED, (0, 0)),
                                        (abi.MODE_BATCHED, (-1, -1))])
def test_device_candidates_and_chain_scores_equal_the_oracle(name, mode, split):
    """The candidates the device's seed and chain stages leave for the rank pass — strand, contig,
    diagonal and the score of the chain each came from (src/mapper.cpp:14-63, src/chain.cpp:12-71)
    — equal the oracle's for every read: one thread per read over global memory (plain form), the
    fused LDS kernel in its shapes, and one wave per read (batched form, every bucket off). Chain
    scores depend on every gap cost the reads' indels (up to 12 bp) cross, which the records
    alone rarely reveal."""
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    print(f"set {name} mode {mode} split {split}", flush=True)
    b = _context(spec).batch(mode=mode, chunk_reads=len(r), len_split=split)
    b.upload(api.prepare(r))
    b.run()
    dn, dc = b.candidates()
    b.close()
    on, oc = _oracle_candidates(spec, g, r)
    assert np.array_equal(dn, on), np.flatnonzero(dn != on)[:5]
    cols = np.stack([dc["rev"].astype(np.int64), dc["contig"].astype(np.int64), dc["diag"],
                     dc["chain_score"].astype(np.int64)], axis=-1)
    bad = [i for i in range(len(r)) if not np.array_equal(cols[i, :on[i]], oc[i, :on[i]])]
    assert not bad, (len(bad), [(i, cols[i, :on[i]].tolist(), oc[i, :on[i]].tolist()) for i in bad[:3]])
