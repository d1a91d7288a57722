"""dnaln_hip_align_static, the entry point that allocates no device memory (every buffer of a
call in one static array per GPU): through the C ABI, its results and CIGARs equal those of
dnaln_hip_align bit for bit, on two calls in a row where the second is smaller (so it runs
over what the first left in the array), for score-only and traceback passes, packed and int32
band widths; it makes no device allocation, and a call too large for the array is refused."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from dnaln_amd import _lib, abi, api, synth
from tests.golden import sets

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    g = synth.make_genome(0, 20_000, 1, 5)
    c = api.Context(g)
    yield c
    c.close()


def _pack(tasks):
    pool, qo, ql, to, tl = sets.pack_tasks(tasks)
    t = np.zeros(len(ql), abi.TASK_DTYPE)
    t["q_off"], t["qlen"], t["t_off"], t["tlen"] = qo, ql, to, tl
    return pool, t


def _align_static(ctx, pool, tasks, tb, score):
    n = len(tasks)
    out = np.zeros(n, abi.RESULT_DTYPE)
    cap = int(tasks["qlen"].astype(np.int64).sum()) * 2 + 4 * n + 4
    cig = np.zeros(cap, np.uint32)
    sp = abi.ScoreParams(score["match"], score["mismatch"], score["gap_open"], score["gap_ext"], 1, score["bw"])
    rc = _lib.hip().dnaln_hip_align_static(ctx.ctx, C.byref(sp), api._p(tasks), n, api._p(pool), len(pool), int(tb),
                                           api._p(out), api._p(cig), cap)
    return rc, out, cig


def _flat(out, cig):
    parts = [cig[o:o + n] for o, n in zip(out["cigar_off"], out["cigar_len"])]
    return np.concatenate(parts) if parts else np.zeros(0, np.uint32)


@pytest.mark.parametrize("bw", [16, 64, 7])
def test_static_align_equals_align_on_two_calls_in_a_row(ctx, bw):
    rng = np.random.default_rng(4400 + bw)
    score = dict(match=2, mismatch=4, gap_open=4, gap_ext=2, bw=bw)
    big = sets.random_tasks(rng, 700, bw, max_len=600)
    small = sets.random_tasks(rng, 90, bw, max_len=250)
    for q, _ in big[:25]:
        q[rng.random(len(q)) < 0.2] = 4
    for tb in (1, 0):
        for name, tasks in (("large", big), ("smaller", small)):
            pool, t = _pack(tasks)
            print(f"bw {bw} traceback {tb} {name} call: {len(t)} tasks", flush=True)
            ref_out, ref_cig = ctx.align(pool, t, bool(tb), score)
            e0, _ = api.device_allocs()
            rc, out, cig = _align_static(ctx, pool, t, tb, score)
            e1, _ = api.device_allocs()
            assert rc == abi.OK, _lib.hip().dnaln_hip_last_error()
            assert e1 == e0  # no device allocation
            for k in ("score", "qb", "qe", "tb", "te", "nm", "cigar_len", "cigar_off"):
                assert np.array_equal(out[k], ref_out[k]), (k, np.flatnonzero(out[k] != ref_out[k])[:5])
            assert np.array_equal(_flat(out, cig), _flat(ref_out, ref_cig))


def test_static_align_refuses_a_call_beyond_its_array(ctx):
    # 4,000 tasks of 3,000 bp at bw 64: about 1.6 GB of traceback directions, far over 64 MiB.
    rng = np.random.default_rng(3)
    q = rng.integers(0, 4, 3000, dtype=np.uint8)
    pool = np.concatenate([q, q])
    t = np.zeros(4000, abi.TASK_DTYPE)
    t["q_off"], t["qlen"], t["t_off"], t["tlen"] = 0, 3000, 3000, 3000
    rc, _, _ = _align_static(ctx, pool, t, 1, dict(match=2, mismatch=4, gap_open=4, gap_ext=2, bw=64))
    assert rc == abi.E_CAPACITY
    assert b"scratch" in _lib.hip().dnaln_hip_last_error()
