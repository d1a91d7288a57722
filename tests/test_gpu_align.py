"""GPU parity of Aligner::align through the C ABI (dnaln_hip_align): reference kernel vectors, the
oracle on random, long and N-rich tasks, and edge cases."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from dnaln_amd import abi, api, synth
from tests import golden_io
from tests.golden import sets

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    g = synth.make_genome(0, 20_000, 1, 5)
    c = api.Context(g)
    yield c
    c.close()


def _tasks_array(qo, ql, to, tl):
    t = np.zeros(len(ql), abi.TASK_DTYPE)
    t["q_off"], t["qlen"], t["t_off"], t["tlen"] = qo, ql, to, tl
    return t


def _score(sp):
    return dict(match=int(sp[0]), mismatch=int(sp[1]), gap_open=int(sp[2]), gap_ext=int(sp[3]), ambig=int(sp[4]),
                bw=int(sp[5]))


@pytest.mark.parametrize("si", range(len(sets.SCHEMES)))
def test_align_matches_reference_vectors(ctx, si):
    kv = golden_io.kernel_vectors()
    sp = kv[f"s{si}_sp"]
    tasks = _tasks_array(kv[f"s{si}_qoff"], kv[f"s{si}_qlen"], kv[f"s{si}_toff"], kv[f"s{si}_tlen"])
    print(f"scheme {si} bw {sp[5]} n {len(tasks)}", flush=True)
    for tb in (0, 1):
        out, cig = ctx.align(kv[f"s{si}_pool"], tasks, bool(tb), _score(sp))
        exp = kv[f"s{si}_tb{tb}_res"]
        got = np.stack([out[k] for k in ("score", "qb", "qe", "tb", "te", "nm")], axis=1)
        bad = np.where((got != exp).any(1))[0]
        assert len(bad) == 0, (si, tb, bad[:5], got[bad[:1]], exp[bad[:1]])
        assert np.array_equal(out["cigar_len"], kv[f"s{si}_tb{tb}_cn"].astype(np.uint32))
        flat = np.concatenate([cig[o:o + n] for o, n in zip(out["cigar_off"], out["cigar_len"])] or [np.zeros(0, np.uint32)])
        assert np.array_equal(flat, kv[f"s{si}_tb{tb}_cig"])


@pytest.mark.parametrize("bw", [16, 32, 64, 7])
def test_align_long_and_n_rich_tasks_match_oracle(ctx, bw):
    rng = np.random.default_rng(1000 + bw)
    sc = dict(match=2, mismatch=4, gap_open=4, gap_ext=2, ambig=1, bw=bw)
    tasks = sets.random_tasks(rng, 300, bw, max_len=1200)
    for q, t in tasks[:30]:
        q[rng.random(len(q)) < 0.3] = 4
    tasks.append((np.zeros(0, np.uint8), np.zeros(2 * bw, np.uint8)))   # empty query
    tasks.append((rng.integers(0, 4, 50, dtype=np.uint8), np.zeros(0, np.uint8)))  # empty target
    pool, qo, ql, to, tl = sets.pack_tasks(tasks)
    print(f"bw {bw} n {len(tasks)}", flush=True)
    out, cig = ctx.align(pool, _tasks_array(qo, ql, to, tl), True, sc)
    ospt = oracle.score_t(**sc)
    for i, (q, t) in enumerate(tasks):
        exp = oracle.align(ospt, q, t, True)
        got = tuple(int(out[k][i]) for k in ("score", "qb", "qe", "tb", "te", "nm"))
        assert got == tuple(exp[:6]), (i, got, exp[:6])
        o, n = int(out["cigar_off"][i]), int(out["cigar_len"][i])
        assert cig[o:o + n].tolist() == exp[6], i


def test_align_no_tasks_is_ok(ctx):
    out, cig = ctx.align(np.zeros(1, np.uint8), np.zeros(0, abi.TASK_DTYPE), True)
    assert len(out) == 0


def test_align_rejects_task_outside_pool(ctx):
    t = np.zeros(1, abi.TASK_DTYPE)
    t["q_off"], t["qlen"], t["t_off"], t["tlen"] = 0, 100, 0, 10
    with pytest.raises(api.DnalnError) as e:
        ctx.align(np.zeros(20, np.uint8), t, True)
    assert e.value.code == abi.E_INVALID
