"""Accuracy against simulated truth (src/evaluate.cpp, src/groundtruth.cpp, src/bench.cpp:272-314)
on the CPU: the generator's true CIGARs, the evaluator's closed forms, and both checks on the
oracle's mapping and alignments. The GPU results go through the same code in
test_gpu_evaluate.py."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from dnaln_amd import abi, api, synth

COMP = np.array([3, 2, 1, 0, 4], np.uint8)


@pytest.fixture(scope="module")
def genome():
    return synth.make_genome(0, 400_000, 3, 21)


def _ops(cig):
    return [(int(c) & 15, int(c) >> 4) for c in cig]


@pytest.mark.parametrize("sub", [0.0, 0.02])
def test_true_cigar_rebuilds_each_read(genome, sub):
    r = synth.make_reads(genome, 3000, synth.LEN_UNIFORM, 60, 400, sub, 0.01, 8, seed=5, random_prob=0.03,
                         truth=True)
    plain = synth.make_reads(genome, 3000, synth.LEN_UNIFORM, 60, 400, sub, 0.01, 8, seed=5, random_prob=0.03)
    assert np.array_equal(r.seq, plain.seq) and np.array_equal(r.truth_pos, plain.truth_pos)
    codes = api.encode(r.seq)
    mism = m_bases = 0
    for i in range(len(r)):
        ops = _ops(r.truth_cigar[int(r.truth_cig_off[i]):int(r.truth_cig_off[i + 1])])
        if r.truth_pos[i] < 0:
            assert not ops
            continue
        L = int(r.lens[i])
        q = codes[int(r.offs[i]):int(r.offs[i]) + L]
        if r.truth_rev[i]:
            q = COMP[q[::-1]]
        assert ops[0][0] == 0 and ops[-1][0] == 0  # no indel at the read ends
        assert sum(n for o, n in ops if o in (0, 1)) == L
        gp, qp = int(r.truth_pos[i]), 0
        for o, n in ops:
            if o == 0:
                g = genome.codes[gp:gp + n]
                mism += int((g != q[qp:qp + n]).sum())
                m_bases += n
                gp, qp = gp + n, qp + n
            elif o == 1:
                qp += n
            else:
                gp += n
    rate = mism / m_bases
    assert (rate == 0) if sub == 0 else abs(rate - sub) < 0.003


def _truth_records(genome, r):
    """Records that state the truth itself: every read correct, every base concordant."""
    c, start, rev = api._truth_arrays(genome, r)
    recs = np.zeros(len(r), abi.RECORD_DTYPE)
    recs["contig"] = np.where(start >= 0, c, -1)
    recs["pos"], recs["rev"], recs["mapq"] = start, rev, 60
    recs["cigar_off"] = r.truth_cig_off[:-1]
    recs["cigar_len"] = np.diff(r.truth_cig_off)
    return recs, r.truth_cigar


def test_evaluator_closed_forms(genome):
    r = synth.make_reads(genome, 2000, synth.LEN_FIXED, 150, 150, 0.01, 0.002, 4, seed=6, random_prob=0.05,
                         truth=True)
    recs, cig = _truth_records(genome, r)
    n_truth = int((r.truth_pos >= 0).sum())
    # aligned (M) bases of each read's truth: the bases that have a true reference position
    m_bases = np.array([sum(n for o, n in _ops(r.truth_cigar[int(r.truth_cig_off[i]):int(r.truth_cig_off[i + 1])])
                            if o == 0) for i in range(len(r))])
    e = api.evaluate(recs, cig, r, genome)
    assert e["n_truth"] == e["records"] == e["mapped"] == e["correct"] == n_truth < len(r)
    assert e["accuracy"] == 1.0 and e["base_concordance"] == 1.0 and e["bases"] == m_bases.sum()
    assert 0 < 150 * n_truth - m_bases.sum() < 0.01 * 150 * n_truth  # a few inserted bases
    assert e["mapped_mapq60"] == n_truth and e["wrong_mapq1"] == 0
    # the tolerance: 10 bp off is still correct, 11 is wrong; a flipped strand or contig is wrong
    real = np.flatnonzero(r.truth_pos >= 0)
    for shift, correct in ((10, True), (-10, True), (11, False), (-11, False)):
        bad = recs.copy()
        bad["pos"][real[:100]] += shift
        e = api.evaluate(bad, cig, r, genome)
        assert e["correct"] == n_truth - (0 if correct else 100)
        if correct:  # shifted records put every base 10 columns off
            assert e["bases_ok"] == e["bases"] - m_bases[real[:100]].sum()
    bad = recs.copy()
    bad["rev"][real[:7]] ^= 1
    bad["mapq"][real[:7]] = 30
    e = api.evaluate(bad, cig, r, genome)
    assert e["correct"] == n_truth - 7 and e["wrong_mapq30"] == 7 and e["wrong_mapq60"] == 0
    assert e["mapped_mapq60"] == n_truth - 7 and e["mapped_mapq30"] == n_truth
    bad = recs.copy()
    bad["contig"][real[:5]] = -1
    assert api.evaluate(bad, cig, r, genome)["mapped"] == n_truth - 5


def test_evaluator_unclipped_start_and_clipped_bases(genome):
    """One read: truth 100M at 1000. A record 5S95M at 1005 is correct (its unclipped start is
    1000) and places 95 of 100 bases."""
    r = synth.Reads(np.frombuffer(b"A" * 100, np.uint8).copy(), np.zeros(1, np.uint64), np.array([100], np.int32),
                    np.array([1000], np.int64), np.zeros(1, np.uint8), 0, np.array([100 << 4], np.uint32),
                    np.array([0, 1], np.uint64))
    rec = np.zeros(1, abi.RECORD_DTYPE)
    rec["contig"], rec["pos"], rec["mapq"], rec["cigar_len"] = 0, 1005, 60, 2
    cig = np.array([5 << 4 | 4, 95 << 4], np.uint32)
    e = api.evaluate(rec, cig, r, genome)
    assert e["correct"] == 1 and (e["bases"], e["bases_ok"]) == (100, 95)
    rec["pos"] = 1016  # unclipped start 1011: 11 bp off
    assert api.evaluate(rec, cig, r, genome)["correct"] == 0


def _oracle_records(recs_o):
    out = np.zeros(len(recs_o), abi.RECORD_DTYPE)
    for f in ("contig", "pos", "rev", "mapq", "score", "nm", "n_cand", "cigar_off"):
        out[f] = recs_o[f]
    out["cigar_len"] = recs_o["n_cigar"]
    return out


def test_oracle_mapping_meets_the_reference_accuracy_bars():
    """src/tests/test_main.cpp:408-410 on the oracle's records: accuracy and base concordance
    above 98%, at most 0.5% of MAPQ-60 records wrong."""
    g = synth.make_genome(0, 4_641_652, 1, 1)
    r = synth.make_reads(g, 4000, synth.LEN_FIXED, 150, 150, 0.01, 0.001, 3, seed=11, truth=True)
    idx = oracle.Index(g.codes, g.contig_off, g.contig_len)
    recs, pool, _ = oracle.map_reads(oracle.default_params(), g, idx, r)
    e = api.evaluate(_oracle_records(recs), pool, r, g)
    assert e["accuracy"] > 0.98 and e["base_concordance"] > 0.98
    assert e["wrong_mapq60"] <= e["mapped_mapq60"] / 200


def test_truth_check_on_oracle_alignments(genome):
    """The true-locus checks on the oracle's traceback results: every CIGAR re-scores, every
    in-band true path scores at most the optimum; a planted score error is caught."""
    r = synth.make_reads(genome, 1500, synth.LEN_UNIFORM, 80, 300, 0.02, 0.005, 6, seed=8, truth=True)
    sp = api.map_params().sc
    pool, tasks, idx = api.truth_tasks(genome, r, sp.bw)
    osp = oracle.score_t(bw=sp.bw)
    res = np.zeros(len(tasks), abi.RESULT_DTYPE)
    cigs = []
    off = 0
    for k, t in enumerate(tasks):
        qo, to = int(t["q_off"]), int(t["t_off"])
        q = pool[qo:qo + int(t["qlen"])]
        tt = pool[to:to + int(t["tlen"])]
        score, qb, qe, tb, te, nm, cig = oracle.align(osp, q, tt, True)
        res[k] = (score, qb, qe, tb, te, nm, off, len(cig), 0)
        cigs += cig
        off += len(cig)
    cig = np.array(cigs or [0], np.uint32)
    tc = [r.truth_cigar[int(r.truth_cig_off[i]):int(r.truth_cig_off[i + 1])] for i in idx]
    toff = np.zeros(len(idx) + 1, np.uint64)
    toff[1:] = np.cumsum([len(x) for x in tc])
    tcat = np.concatenate(tc).astype(np.uint32)

    def check(rr):
        import ctypes as C
        from dnaln_amd import _lib
        st = abi.TruthStats()
        assert _lib.host().dnaln_truth_check(C.byref(sp), pool.ctypes.data, tasks.ctypes.data, len(tasks),
                                             rr.ctypes.data, cig.ctypes.data, tcat.ctypes.data, toff.ctypes.data,
                                             C.byref(st)) == 0
        return {f: int(getattr(st, f)) for f, _ in abi.TruthStats._fields_}

    s = check(res)
    assert s["n"] == len(tasks) and s["rescored_ok"] == s["n"] and s["optimal_ok"] == s["in_band"]
    assert s["in_band"] >= 0.95 * s["n"] and s["bases_ok"] >= 0.97 * s["bases"]
    bad = res.copy()
    bad["score"][:3] += 1
    assert check(bad)["rescored_ok"] == s["n"] - 3
