"""GPU index build (dnaln_hip_index_build, and dnaln_hip_create without an index) against the
host build of MinimizerIndex (src/index.cpp:10-45), itself pinned slot for slot to the oracle.

The device places keys with concurrent linear probing, so a key may sit in another slot of its
probe chain. What must hold: the same bucket count, key count, largest count and pos[] as the
host build; the same set of (key, offset, count) buckets; and every key reachable by linear
probing from its home slot before an empty slot (src/index.hpp:24-38), so every lookup returns
the same slice. Contexts on a GPU-built or file-loaded index give the reference's records."""
from __future__ import annotations

import numpy as np
import pytest

from dnaln_amd import abi, api
from tests import golden_io
from tests.golden import sets
from tests.test_index_build import EMPTY, ragged_genome

pytestmark = pytest.mark.gpu


def probe_reachable(table: np.ndarray) -> bool:
    """Every occupied slot is reached from its key's home slot (key & mask) without crossing an
    empty slot, probing upward with wrap-around."""
    cap = len(table)
    occ = table["key"] != EMPTY
    e0 = int(np.flatnonzero(~occ)[0])
    rot = np.roll(np.arange(cap), -e0)  # rotated order starts at an empty slot: no chain wraps
    occ_r = occ[rot]
    idx = np.arange(cap)
    last_empty = np.maximum.accumulate(np.where(~occ_r, idx, 0))
    slots = np.flatnonzero(occ_r)
    home = (table["key"][rot[slots]] & np.uint64(cap - 1)).astype(np.int64)
    home_r = (home - e0) % cap
    return bool(np.all((home_r <= slots) & (home_r > last_empty[slots])))


def buckets(table: np.ndarray) -> np.ndarray:
    t = table[table["key"] != EMPTY]
    return np.sort(t, order="key")


def assert_same_index(gpu: api.Index, host: api.Index):
    assert (gpu.n_buckets, gpu.n_keys, gpu.n_pos, gpu.max_count) == (host.n_buckets, host.n_keys, host.n_pos,
                                                                     host.max_count)
    assert np.array_equal(gpu.positions(), host.positions())
    tg, th = gpu.table(), host.table()
    assert np.array_equal(buckets(tg), buckets(th))
    assert probe_reachable(tg) and probe_reachable(th)


@pytest.mark.parametrize("spec", [sets.UNIFORM_4M, sets.REPEATS_8M, sets.TWINS_4M])
@pytest.mark.parametrize("k,w", [(15, 10), (13, 7)])
def test_gpu_index_equals_host_index(spec, k, w):
    g = sets.genome({"genome": spec})
    print(f"genome {spec['total']} bp, {spec['contigs']} contigs, k {k} w {w}", flush=True)
    gpu = api.Index(g, k, w, build="gpu")
    assert_same_index(gpu, api.Index(g, k, w))
    assert gpu.build_ms > 0


@pytest.mark.parametrize("k,w", [(15, 10), (16, 10), (11, 10), (15, 5), (12, 1), (5, 17), (15, 25), (4, 2)])
def test_gpu_index_on_ragged_contigs_with_n(k, w):
    """Contigs shorter than k or than one window, N runs, low-complexity runs, even k."""
    g = ragged_genome(5)
    print(f"ragged genome, k {k} w {w}", flush=True)
    assert_same_index(api.Index(g, k, w, build="gpu"), api.Index(g, k, w))


def test_gpu_index_rejects_k_above_16():
    g = ragged_genome(5)
    with pytest.raises(api.DnalnError) as e:
        api.Index(g, 17, 10, build="gpu")
    assert e.value.code == abi.E_INVALID
    with pytest.raises(api.DnalnError) as e:
        api.Context(g, "gpu", k=17)
    assert e.value.code == abi.E_INVALID


def _records_equal(name, ctx):
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    recs, pool, _ = ctx.map(r)
    gold = golden_io.records(name)
    bad = [i for i in range(len(gold)) if api.record_fields(recs[i], pool, g.names) != gold[i][1:]]
    assert not bad, (name, len(bad))
    b = ctx.batch(mode=abi.MODE_BATCHED)
    b.upload(api.prepare(r))
    b.run()
    recs, pool, _ = b.download()
    b.close()
    bad = [i for i in range(len(gold)) if api.record_fields(recs[i], pool, g.names) != gold[i][1:]]
    assert not bad, (name, "batched", len(bad))


@pytest.mark.parametrize("name", ["uniform4m_150bp_10k", "edge_repeats_bw32", "edge_twins_sr_bw5",
                                  "edge_repeats_params"])
def test_context_with_device_built_index_gives_reference_records(name):
    spec = sets.SETS[name]
    g = sets.genome(spec)
    print(f"set {name} on a device-built index", flush=True)
    ctx = api.Context(g, "gpu", **sets.map_params(spec))
    info = ctx.index_info()
    host = api.Index(g, ctx.params.k, ctx.params.w)
    assert (info["n_buckets"], info["n_pos"]) == (host.n_buckets, host.n_pos) and info["build_ms"] > 0
    _records_equal(name, ctx)
    ctx.close()


def test_context_on_loaded_index_file_gives_reference_records(tmp_path):
    name = "edge_repeats_bw32"
    spec = sets.SETS[name]
    g = sets.genome(spec)
    gpu = api.Index(g, 15, 10, build="gpu")
    gpu.save(tmp_path / "r.dnx")
    loaded = api.Index.load(tmp_path / "r.dnx", g)
    assert_same_index(loaded, api.Index(g))
    ctx = api.Context(g, loaded, **sets.map_params(spec))
    assert ctx.index_info()["build_ms"] == 0
    _records_equal(name, ctx)
    ctx.close()
