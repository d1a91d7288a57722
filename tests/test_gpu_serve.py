"""Serving loop on the GPU (dnaln_serve_*): continuous batching gives every request the records it
gets alone, which are the reference's, whatever it was batched with; the HBM budget bounds what
is in flight; batches overlap on several slots."""
from __future__ import annotations

import numpy as np
import pytest

from dnaln_amd import abi, api, synth
from tests import golden_io
from tests.golden import sets

pytestmark = pytest.mark.gpu


def _requests(n_reads: int, seed: int, max_size: int):
    first, _ = synth.request_trace(n_reads, 1.0, seed, max_size=max_size)
    if first[-1] < n_reads:  # the tail as one last request
        first = np.append(first, np.uint64(n_reads))
    return first


def _sub(rb: api.ReadBatch, a: int, b: int) -> api.ReadBatch:
    base = int(rb.offs[a]) if b > a else 0
    return api.ReadBatch(rb.codes[base:], rb.offs[a:b] - np.uint64(base), rb.lens[a:b], rb.fnv[a:b])


def _fields(recs, pool, g, i):
    return api.record_fields(recs[i], pool, g.names)


@pytest.mark.parametrize("name", ["uniform4m_150bp_10k", "edge_repeats_bw32", "edge_uniform_bw64"])
def test_each_request_same_alone_and_in_the_mix(name):
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    rb = api.prepare(r)
    ctx = api.Context(g, **sets.map_params(spec))
    first = _requests(len(rb), 5, 600)
    n_req = len(first) - 1
    print(f"set {name} requests {n_req}", flush=True)
    # alone: one slot, one request at a time
    alone = api.Server(ctx, slots=1)
    alone_recs = {}
    for q in range(n_req):
        alone.submit(q, _sub(rb, int(first[q]), int(first[q + 1])))
        got = None
        while got is None:
            alone.progress()
            got = alone.take()
        assert got[0] == q
        alone_recs[q] = got
    alone.close()
    # in the mix: every request arrives at once; four slots and a budget that holds two full
    # batches make the server pack requests together and run two batches at a time
    mix = api.Server(ctx, slots=4, max_batch_reads=1024)
    budget_probe = mix.batch_bytes(1024, 1024 * int(rb.lens.max()), int(rb.lens.max()))
    mix.close()
    mix = api.Server(ctx, slots=4, max_batch_reads=1024, hbm_budget=2 * budget_probe + 1)
    arrivals = np.zeros(n_req)
    allocs = api.device_allocs()
    lat, wall, recs, pool = mix.run_trace(rb, first, arrivals, want_records=True)
    allocs_after = api.device_allocs()
    st = mix.stats()
    mix.close()
    print("stats", st, "wall", wall, "device allocations while serving", allocs_after[0] - allocs[0], flush=True)
    # slots are sized when the server starts: serving allocates no device memory
    assert allocs_after == allocs
    assert st["completed"] == n_req and st["max_inflight_bytes"] <= st["budget_bytes"]
    assert st["max_inflight_batches"] >= 2  # batches overlapped
    assert np.all(lat >= 0)
    gold = golden_io.records(name)
    for q in range(n_req):
        _, arec, apool = alone_recs[q]
        for j, i in enumerate(range(int(first[q]), int(first[q + 1]))):
            a = api.record_fields(arec[j], apool, g.names)
            assert a == _fields(recs, pool, g, i), (q, i)
            assert a == gold[i][1:], (q, i)
    ctx.close()


def test_budget_defers_admission_and_rejects_oversized_requests():
    spec = sets.SETS["uniform4m_150bp_10k"]
    g = sets.genome(spec)
    rb = api.prepare(sets.reads(spec, g))
    ctx = api.Context(g, **sets.map_params(spec))
    probe = api.Server(ctx, slots=1)
    one = probe.batch_bytes(512, 512 * 150, 150)
    probe.close()
    s = api.Server(ctx, slots=4, max_batch_reads=512, hbm_budget=one)  # room for one batch at a time
    first = np.arange(0, 4097, 512, dtype=np.uint64)
    lat, wall, _, _ = s.run_trace(rb, first, np.zeros(len(first) - 1), want_records=False)
    st = s.stats()
    assert st["completed"] == len(first) - 1
    assert st["max_inflight_batches"] == 1 and st["deferred_by_budget"] > 0
    with pytest.raises(RuntimeError):
        s.submit(99, _sub(rb, 0, 2048))  # larger than the whole budget
    s.close()
    ctx.close()
