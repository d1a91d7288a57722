"""The batched form (graph = 0, the default): the whole batch as one chunk, length buckets for the
fused seed + chain kernel, length-sorted SW tasks. Repeated runs of one batch with other device work
between them, batches of many sizes through one batch object, a first anchor capacity far too small,
the int32 fill for a band without a packed kernel, and the refusal of the device-count forms
(graph = 1, 2), which stay off on the device. Every case must give the reference binary's records
(or the oracle's, where no golden set covers the band)."""
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


def _bad(name, recs, pool, g, lo=0):
    gold = golden_io.records(name)
    return [i for i in range(len(recs))
            if api.record_fields(recs[i], pool, g.names) != gold[lo + i][1:]]


@pytest.mark.parametrize("name", ["uniform4m_150bp_10k", "edge_uniform_bw64", "edge_repeats_bw32"])
def test_repeated_runs_give_reference_records(name):
    """Three runs of one resident batch, with other device work (a plain-form map) between them."""
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    ctx = _context(spec)
    b = ctx.batch(mode=abi.MODE_BATCHED)
    b.upload(api.prepare(r))
    for rep in range(3):
        print(f"set {name} run {rep}", flush=True)
        b.run()
        recs, pool, stats = b.download()
        assert _bad(name, recs, pool, g) == []
        assert stats.reads == len(r)
        ctx.map(r.subset(0, 500))
    b.close()


@pytest.mark.parametrize("name", ["edge_uniform_bw16", "edge_twins_bw16"])
def test_stage_timings_and_reference_records(name):
    """Per-stage timing events around the launches; the records are the reference's."""
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    print(f"set {name}", flush=True)
    b = _context(spec).batch(mode=abi.MODE_BATCHED)
    b.upload(api.prepare(r))
    b.run()
    recs, pool, _ = b.download()
    ms, launches = b.timings()
    b.close()
    assert _bad(name, recs, pool, g) == []
    assert ms[5] > 0 and ms[0] > 0 and ms[3] > 0
    print("timings", [round(x, 3) for x in ms], launches, flush=True)


@pytest.mark.parametrize("graph", [1, 2])
def test_device_count_forms_are_refused_while_switched_off(graph):
    """graph = 2 (the device-count sequence) and graph = 1 (it, captured) are switched off in the
    device build: their first run faulted the GPU and the cause is not found. The batch refuses
    them at creation, before anything reaches the device."""
    spec = sets.SETS["edge_uniform_bw16"]
    with pytest.raises(api.DnalnError) as e:
        _context(spec).batch(mode=abi.MODE_BATCHED, graph=graph)
    assert e.value.code == abi.E_INVALID
    assert "switched off" in str(e.value)


def test_batches_of_many_sizes_through_one_batch_object():
    """Batch sizes that need plans of other grid sizes (and back), each the reference's records."""
    name = "edge_uniform_bw16"
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    b = _context(spec).batch(mode=abi.MODE_BATCHED)
    for lo, hi in ((0, 100), (100, 4000), (7, 307), (0, 4000), (3999, 4000), (50, 60)):
        print(f"reads [{lo}, {hi})", flush=True)
        b.upload(api.prepare(r.subset(lo, hi)))
        b.run()
        recs, pool, _ = b.download()
        assert _bad(name, recs, pool, g, lo) == []
    b.close()


@pytest.mark.parametrize("name", ["edge_repeats_bw32", "edge_uniform_bw64"])
def test_small_first_anchor_capacity(name):
    """A first anchor capacity of one per read is far too small for the global path (every read
    on it); the run grows its buffers and gives the reference's records."""
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    b = _context(spec).batch(mode=abi.MODE_BATCHED, anchors_per_read=1, len_split=(-1, -1))
    b.upload(api.prepare(r))
    b.run()
    recs, pool, stats = b.download()
    b.close()
    assert stats.anchors > len(r)
    assert _bad(name, recs, pool, g) == []


def test_band_without_packed_kernel_in_the_batched_form():
    """bw 7 has no packed specialisation: the batched form runs the int32 fill (one task per
    thread of a fixed grid); its records equal the oracle's."""
    import oracle
    spec = dict(sets.SETS["edge_uniform_bw16"])
    spec["flags"] = {"--bw": 7}
    g = sets.genome(spec)
    r = sets.reads(spec, g).subset(0, 1500)
    mp = sets.map_params(spec)
    ctx = api.Context(g, **mp)
    b = ctx.batch(mode=abi.MODE_BATCHED)
    b.upload(api.prepare(r))
    b.run()
    recs, pool, _ = b.download()
    b.close()
    idx = oracle.Index(g.codes, g.contig_off, g.contig_len, mp["k"], mp["w"])
    orecs, opool, _ = oracle.map_reads(oracle.default_params(**mp), g, idx, r)
    bad = [i for i in range(len(r))
           if api.record_fields(recs[i], pool, g.names) != oracle.record_fields(orecs[i], opool, g.names)]
    ctx.close()
    assert bad == []
