"""Accuracy of the GPU path against simulated truth (src/evaluate.cpp, src/bench.cpp:272-314):
the reference's own bars (src/tests/test_main.cpp:408-410) on the device's records, the same
statistics as the oracle's records of the same reads, and the true-locus kernel checks
(re-scoring, optimality against the true path) on the device aligner's traceback results."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from dnaln_amd import abi, api, synth
from tests.test_evaluate import _oracle_records

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bw,lens", [(16, (150, 150)), (64, (50, 1000))])
def test_gpu_mapping_accuracy_equals_the_oracle_and_meets_the_bars(bw, lens):
    g = synth.make_genome(0, 4_641_652, 1, 1)
    mode = synth.LEN_FIXED if lens[0] == lens[1] else synth.LEN_LOGUNIFORM
    r = synth.make_reads(g, 4000, mode, lens[0], lens[1], 0.01, 0.001, 3, seed=11, truth=True)
    print(f"bw {bw} lens {lens}", flush=True)
    ctx = api.Context(g, bw=bw)
    b = ctx.batch(mode=abi.MODE_BATCHED)
    b.upload(api.prepare(r))
    b.run()
    recs, pool, _ = b.download()
    b.close()
    e = api.evaluate(recs, pool, r, g)
    idx = oracle.Index(g.codes, g.contig_off, g.contig_len)
    ro, po, _ = oracle.map_reads(oracle.default_params(bw=bw), g, idx, r)
    assert e == api.evaluate(_oracle_records(ro), po, r, g)
    assert e["accuracy"] > 0.98 and e["base_concordance"] > 0.98
    assert e["wrong_mapq60"] <= e["mapped_mapq60"] / 200
    ctx.close()


@pytest.mark.parametrize("bw", [5, 16, 64])
def test_gpu_alignments_at_the_true_locus(bw):
    g = synth.make_genome(1, 2_000_000, 4, 7, 0.3)
    r = synth.make_reads(g, 3000, synth.LEN_UNIFORM, 60, 600, 0.02, 0.005, 8, seed=3, truth=True)
    ctx = api.Context(g, bw=bw)
    s = api.truth_check(ctx, r, g)
    print(bw, s, flush=True)
    assert s["n"] == len(r) and s["rescored_ok"] == s["n"] and s["optimal_ok"] == s["in_band"]
    assert s["in_band"] > 0.3 * s["n"] and s["bases_ok"] > 0.8 * s["bases"]
    ctx.close()
