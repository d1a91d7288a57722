"""CPU build of the device lane programs (libdnaln_emu.so) against the reference kernel vectors and
the oracle: the packed int16 tricks (v_perm scoring, sign-bit flags, saved best row, lock-step
walk) are exact before they reach a GPU."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import oracle
from tests import golden_io
from tests.golden import sets

LIB = Path(__file__).resolve().parent.parent / "dnaln_amd" / "lib" / "libdnaln_emu.so"


@pytest.fixture(scope="module")
def emu():
    if not LIB.exists():
        pytest.fail("libdnaln_emu.so is not built (python -m dnaln_amd.build emu)")
    L = C.CDLL(str(LIB))
    L.emu_align.restype = C.c_int
    return L


def _run(emu, sp, pool, qo, ql, to, tl, tb, force):
    n = len(ql)
    res = np.zeros((n, 6), np.int32)
    cap = int(ql.astype(np.int64).sum()) * 2 + 8 * n + 8
    cig = np.zeros(cap, np.uint32)
    co = np.zeros(n, np.int64)
    cn = np.zeros(n, np.int32)
    v = C.c_void_p
    rc = emu.emu_align(v(sp.ctypes.data), v(pool.ctypes.data), v(qo.ctypes.data), v(ql.ctypes.data),
                       v(to.ctypes.data), v(tl.ctypes.data), C.c_size_t(n), tb, v(res.ctypes.data),
                       v(cig.ctypes.data), v(co.ctypes.data), v(cn.ctypes.data), C.c_size_t(cap), force)
    assert rc >= 0, rc
    return rc, res, cig[:int(cn.sum())], cn


@pytest.mark.parametrize("si", range(len(sets.SCHEMES)))
@pytest.mark.parametrize("force", [0, 1])
def test_emulated_kernels_match_reference_vectors(emu, si, force):
    kv = golden_io.kernel_vectors()
    sp = kv[f"s{si}_sp"].astype(np.int32)
    args = [kv[f"s{si}_{k}"] for k in ("pool", "qoff", "qlen", "toff", "tlen")]
    for tb in (0, 1):
        rc, res, cig, cn = _run(emu, sp, *args, tb, force)
        if force == 0 and int(sp[5]) in (5, 8, 16, 21, 24, 32, 64):
            assert rc == 1  # the packed kernel ran
        assert np.array_equal(res, kv[f"s{si}_tb{tb}_res"])
        assert np.array_equal(cn, kv[f"s{si}_tb{tb}_cn"])
        assert np.array_equal(cig, kv[f"s{si}_tb{tb}_cig"])


def _run_lanes(emu, sp, pool, qo, ql, to, tl, rev, tlo, tb, lanes):
    n = len(ql)
    res = np.zeros((n, 6), np.int32)
    cap = int(ql.astype(np.int64).sum()) * 2 + 8 * n + 8
    cig = np.zeros(cap, np.uint32)
    co = np.zeros(n, np.int64)
    cn = np.zeros(n, np.int32)
    v = C.c_void_p
    rc = emu.emu_align_lanes(v(sp.ctypes.data), v(pool.ctypes.data), v(qo.ctypes.data), v(ql.ctypes.data),
                             v(to.ctypes.data), v(tl.ctypes.data), v(rev.ctypes.data), v(tlo.ctypes.data),
                             C.c_size_t(n), tb, v(res.ctypes.data), v(cig.ctypes.data), v(co.ctypes.data),
                             v(cn.ctypes.data), C.c_size_t(cap), 2, lanes)
    assert rc == 1, rc  # the split fill ran
    return res, cig[:int(cn.sum())], cn


@pytest.mark.parametrize("si,lanes", [(4, 2), (5, 4)])
def test_emulated_split_fill_matches_reference_vectors(emu, si, lanes):
    """The fill split over 2 (bw 32) or 4 (bw 64) lanes per pair of alignments, lanes in lock
    step with their neighbour exchanges, against the reference binary's kernel vectors."""
    kv = golden_io.kernel_vectors()
    sp = kv[f"s{si}_sp"].astype(np.int32)
    args = [kv[f"s{si}_{k}"] for k in ("pool", "qoff", "qlen", "toff", "tlen")]
    n = len(args[2])
    for tb in (0, 1):
        res, cig, cn = _run_lanes(emu, sp, *args, np.zeros(n, np.uint8), np.zeros(n, np.int32), tb, lanes)
        assert np.array_equal(res, kv[f"s{si}_tb{tb}_res"])
        assert np.array_equal(cn, kv[f"s{si}_tb{tb}_cn"])
        assert np.array_equal(cig, kv[f"s{si}_tb{tb}_cig"])


@pytest.mark.parametrize("bw,lanes", [(32, 2), (64, 4)])
def test_emulated_split_fill_on_reverse_clipped_and_n_rich_tasks(emu, bw, lanes):
    """Reverse-complemented queries, windows starting inside the target and N-rich queries
    through the split fill: the oracle on the equivalent plain task decides."""
    rng = np.random.default_rng(300 + bw)
    sc = dict(match=2, mismatch=4, gap_open=4, gap_ext=2, ambig=1, bw=bw)
    tasks = sets.random_tasks(rng, 150, bw, max_len=900)
    for q, t in tasks[:15]:
        q[rng.random(len(q)) < 0.2] = 4
    pool, qo, ql, to, tl = sets.pack_tasks(tasks)
    n = len(tasks)
    rev = (rng.random(n) < 0.5).astype(np.uint8)
    tlo = np.where(rng.random(n) < 0.3, rng.integers(0, 40, n), 0).astype(np.int32)
    sp = np.array(list(sc.values()), np.int32)
    res, cig, cn = _run_lanes(emu, sp, pool, qo, ql, to, tl, rev, tlo, 1, lanes)
    sres, _, _ = _run_lanes(emu, sp, pool, qo, ql, to, tl, rev, tlo, 0, lanes)
    ospt = oracle.score_t(**sc)
    offs = np.concatenate([[0], np.cumsum(cn)])
    for i, (q, t) in enumerate(tasks):
        qq = np.where(q[::-1] < 4, 3 - q[::-1], 4).astype(np.uint8) if rev[i] else q
        tt = t.copy()
        tt[:min(int(tlo[i]), len(tt))] = 4
        exp = oracle.align(ospt, qq, tt, True)
        assert tuple(res[i]) == tuple(exp[:6]), (bw, i)
        assert cig[offs[i]:offs[i + 1]].tolist() == exp[6], (bw, i)
        assert sres[i][0] == exp[0], (bw, i)


def test_emulated_packed_kernel_on_long_and_n_rich_tasks(emu):
    rng = np.random.default_rng(77)
    for bw in (16, 32):
        sc = dict(match=2, mismatch=4, gap_open=4, gap_ext=2, ambig=1, bw=bw)
        tasks = sets.random_tasks(rng, 130, bw, max_len=1500)
        for q, t in tasks[:20]:
            q[rng.random(len(q)) < 0.2] = 4
        pool, qo, ql, to, tl = sets.pack_tasks(tasks)
        sp = np.array(list(sc.values()), np.int32)
        rc, res, cig, cn = _run(emu, sp, pool, qo, ql, to, tl, 1, 2)
        assert rc == 1
        ospt = oracle.score_t(**sc)
        co = np.concatenate([[0], np.cumsum(cn)])
        for i, (q, t) in enumerate(tasks):
            exp = oracle.align(ospt, q, t, True)
            assert tuple(res[i]) == tuple(exp[:6]), i
            assert cig[co[i]:co[i + 1]].tolist() == exp[6], i


def test_emulated_packed_kernel_reverse_queries_and_clipped_windows(emu):
    """Reverse-complemented queries stream backwards and windows may start inside the target
    (positions below tlo read as N): both against the oracle on the equivalent plain task."""
    rng = np.random.default_rng(91)
    for bw in (8, 16, 32):
        sc = dict(match=2, mismatch=4, gap_open=4, gap_ext=2, ambig=1, bw=bw)
        tasks = sets.random_tasks(rng, 260, bw, max_len=700)
        pool, qo, ql, to, tl = sets.pack_tasks(tasks)
        n = len(tasks)
        rev = (rng.random(n) < 0.5).astype(np.uint8)
        tlo = np.where(rng.random(n) < 0.3, rng.integers(0, 40, n), 0).astype(np.int32)
        res = np.zeros((n, 6), np.int32)
        cap = int(ql.astype(np.int64).sum()) * 2 + 8 * n + 8
        cig = np.zeros(cap, np.uint32)
        co = np.zeros(n, np.int64)
        cn = np.zeros(n, np.int32)
        v = C.c_void_p
        sp = np.array(list(sc.values()), np.int32)
        rc = emu.emu_align_ex(v(sp.ctypes.data), v(pool.ctypes.data), v(qo.ctypes.data), v(ql.ctypes.data),
                              v(to.ctypes.data), v(tl.ctypes.data), v(rev.ctypes.data), v(tlo.ctypes.data),
                              C.c_size_t(n), 1, v(res.ctypes.data), v(cig.ctypes.data), v(co.ctypes.data),
                              v(cn.ctypes.data), C.c_size_t(cap), 2)
        assert rc == 1
        ospt = oracle.score_t(**sc)
        offs = np.concatenate([[0], np.cumsum(cn)])
        for i, (q, t) in enumerate(tasks):
            qq = np.where(q[::-1] < 4, 3 - q[::-1], 4).astype(np.uint8) if rev[i] else q
            tt = t.copy()
            tt[:min(int(tlo[i]), len(tt))] = 4
            exp = oracle.align(ospt, qq, tt, True)
            assert tuple(res[i]) == tuple(exp[:6]), (bw, i)
            assert cig[offs[i]:offs[i + 1]].tolist() == exp[6], (bw, i)


def _gap_costs(k, band):
    """cost[dd] = int(0.01 k dd + 0.5 ilog2 dd) for dd = 1..band, 0 at dd = 0 (src/chain.cpp:36)."""
    return np.array([0] + [int(0.01 * k * dd + 0.5 * (dd.bit_length() - 1)) for dd in range(1, band + 1)], np.int32)


def _emu_seed_chain(emu, spec, g, r, cap, maxmin, band=None, cost=None, want_cands=False, table=1):
    from dnaln_amd import api
    mp = sets.map_params(spec)
    k, w, bw = mp["k"], mp["w"], mp["bw"]
    idx = api.Index(g, k, w)
    tab, pos = idx.table(), idx.positions()
    rb = api.prepare(r)
    if band is None:
        band = min(16, bw)  # ChainParams::band, clamped to the SW band (src/mapper.cpp:9)
    if cost is None:
        cost = _gap_costs(k, band)
    m = np.array([k, w, mp["max_occ"], bw, mp["max_cand"], mp["min_chain"], 50, 5000, band, 2, table], np.int32)
    st = np.zeros(3, np.uint64)
    mc = max(mp["max_cand"], 1)
    out_n = np.zeros(max(len(r), 1), np.int32)
    out_c = np.zeros((max(len(r), 1), mc, 4), np.int64)
    v = C.c_void_p
    emu.emu_seed_chain.restype = C.c_long
    d = emu.emu_seed_chain(v(rb.codes.ctypes.data), v(rb.offs.ctypes.data), v(rb.lens.ctypes.data),
                           C.c_uint64(len(r)), v(tab.ctypes.data), C.c_uint64(len(tab)), v(pos.ctypes.data),
                           v(g.contig_off.ctypes.data), g.n_contigs, v(m.ctypes.data), C.c_double(0.5),
                           v(cost.ctypes.data), cap, maxmin, v(st.ctypes.data),
                           v(out_n.ctypes.data) if want_cands else None, v(out_c.ctypes.data) if want_cands else None)
    if want_cands:
        return d, int(st[0]), out_n[:len(r)], out_c[:len(r)]
    return d, int(st[0])


def _candidate_mismatches(n_a, c_a, n_b, c_b):
    """Reads whose candidate lists (count, then every column of every candidate) differ."""
    bad = []
    for i in range(len(n_a)):
        if n_a[i] != n_b[i] or not np.array_equal(c_a[i, :n_a[i]], c_b[i, :n_b[i]]):
            bad.append(i)
    return bad


@pytest.mark.parametrize("name", ["edge_uniform_bw64", "edge_uniform_bw16", "edge_repeats_bwa_bw8",
                                  "edge_repeats_params"])
@pytest.mark.parametrize("table", [0, 1])
def test_emulated_candidates_and_chain_scores_equal_the_oracle(emu, name, table):
    """Every read's candidates — strand, contig, diagonal and the score of the chain each came
    from — as the device's seed and chain programs compute them with the parameter set-up the
    device path uses (init_chain_args), equal the oracle's (src/mapper.cpp:14-63). The reads
    carry indels of up to 12 bp, so chains cross drifts with a non-zero gap cost: the same
    check with the gap-cost table zeroed finds differing reads, so a lost gap cost cannot pass."""
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    mp = sets.map_params(spec)
    oidx = oracle.Index(g.codes, g.contig_off, g.contig_len, mp["k"], mp["w"])
    on, oc = oracle.candidates(oracle.default_params(**mp), g, oidx, r)
    assert (on > 0).sum() > len(r) // 2
    _, _, en, ec = _emu_seed_chain(emu, spec, g, r, 64, 56, want_cands=True, table=table)
    assert _candidate_mismatches(en, ec, on, oc) == []
    band = min(16, mp["bw"])
    _, _, zn, zc = _emu_seed_chain(emu, spec, g, r, 64, 56, cost=np.zeros(band + 1, np.int32), want_cands=True,
                                   table=table)
    assert len(_candidate_mismatches(zn, zc, on, oc)) > 0


@pytest.mark.parametrize("name", ["uniform4m_150bp_10k", "edge_twins_bw16", "edge_repeats_params",
                                  "edge_twins_sr_bw5", "edge_uniform_linear_bw21"])
@pytest.mark.parametrize("cap,maxmin,table", [(32, 32, 1), (64, 56, 1), (32, 24, 1), (64, 56, 0)])
def test_emulated_fused_seed_chain_matches_global_path(emu, name, cap, maxmin, table):
    """The fused one-read-per-thread program (register van Herk sketch, LDS minimizer list,
    batched probes, LDS chain DP and extraction) gives every read the candidates of the
    global-memory program, or declines the read (it then takes the global path)."""
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    differ, declined = _emu_seed_chain(emu, spec, g, r, cap, maxmin, table=table)
    assert differ == 0
    assert declined < len(r)


def test_emulated_fused_seed_chain_with_a_chain_band_above_16(emu):
    """A chain band above 16 (a band option of the reference's ChainParams) takes the fused DP's
    byte-table gap costs instead of the four-bit table; candidates still match the global path."""
    spec = sets.SETS["edge_uniform_linear_bw21"]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    differ, declined = _emu_seed_chain(emu, spec, g, r, 64, 56, band=21)
    assert differ == 0
    assert declined < len(r)


# ------------------------------------------------------------------ the 8-byte probe table
_BUCKET = np.dtype([("key", "<u8"), ("offset", "<u4"), ("count", "<u4")])


def _linear_probe_table(n, keys, counts, rng):
    """A table of n buckets (power of two) holding keys in insertion order with linear probing
    from key & (n - 1) (src/index.cpp:32-44, src/index.hpp:24-38), and its occurrence array."""
    tab = np.zeros(n, _BUCKET)
    tab["key"] = np.iinfo(np.uint64).max
    pos = []
    for key, c in zip(keys, counts):
        i = int(key) & (n - 1)
        while tab["key"][i] != np.iinfo(np.uint64).max:
            i = (i + 1) & (n - 1)
        tab[i] = (key, len(pos), c)
        pos.extend(sorted(int(x) for x in rng.integers(0, 2 ** 32, c, dtype=np.uint64)))
    return tab, np.array(pos + [0], np.uint32)


def _probe_check(emu, tab, pos, max_occ, queries):
    st = np.zeros(2, np.uint64)
    q = np.ascontiguousarray(queries, np.uint64)
    emu.emu_probe_check.restype = C.c_long
    bad = emu.emu_probe_check(C.c_void_p(tab.ctypes.data), C.c_uint64(len(tab)), C.c_void_p(pos.ctypes.data),
                              C.c_uint32(max_occ), C.c_void_p(q.ctypes.data), C.c_uint64(len(q)),
                              C.c_void_p(st.ctypes.data))
    return bad, int(st[0]), int(st[1])


@pytest.mark.parametrize("n,fill,max_occ", [(1, 0, 500), (2, 1, 500), (4, 2, 500), (4, 3, 1), (64, 40, 500),
                                            (64, 63, 2), (1024, 512, 500), (1024, 900, 3)])
def test_probe_table_lookups_equal_the_bucket_table(emu, n, fill, max_occ):
    """The 8-byte probe table (probe_table.hpp) returns, for every key and for absent hashes, the
    occurrences the 16-byte table's linear probe returns after the max_occ filter. Keys cluster
    at the top of the table so probe chains wrap past the last slot, chains run longer than the
    four slots a probe loads at once, and counts sit at 1, 2, max_occ and max_occ + 1."""
    rng = np.random.default_rng(n * 1000 + fill)
    # 30-bit hashes (k = 15), most of them homed in the top eighth of the table
    hashes = rng.choice(2 ** 30, size=4 * fill + 8, replace=False).astype(np.uint64)
    top = (hashes & ~np.uint64(n - 1)) | (np.uint64(n - 1) - (hashes % np.uint64(max(n // 8, 1))))
    keys = list(dict.fromkeys(int(x) for x in top))[:fill]
    counts = [int(rng.choice([1, 2, 3, max_occ, max_occ + 1, 7])) for _ in keys]
    tab, pos = _linear_probe_table(n, keys, counts, rng)
    absent = [int(x) for x in rng.integers(0, 2 ** 30, 200, dtype=np.uint64) if int(x) not in set(keys)]
    near = [(k ^ (1 << 29)) for k in keys]  # same home slot, different key
    queries = keys + absent + near
    bad, found, long_chains = _probe_check(emu, tab, pos, max_occ, queries)
    assert bad == 0
    assert found == len(keys)
    if fill >= 40:
        assert long_chains > 0


def test_probe_table_lookups_on_a_genome_index(emu):
    """Every minimizer of a genome index and as many absent hashes: the same occurrences through
    both tables (the index's own sorted-key insertion order and load)."""
    from dnaln_amd import api
    spec = sets.SETS["edge_repeats_params"]
    g = sets.genome(spec)
    mp = sets.map_params(spec)
    idx = api.Index(g, mp["k"], mp["w"])
    tab, pos = idx.table(), idx.positions()
    tab = np.frombuffer(tab.tobytes(), _BUCKET)
    keys = tab["key"][tab["key"] != np.iinfo(np.uint64).max]
    rng = np.random.default_rng(3)
    absent = rng.integers(0, 2 ** (2 * mp["k"]), len(keys), dtype=np.uint64)
    for max_occ in (mp["max_occ"], 2, 1):
        bad, found, _ = _probe_check(emu, tab, np.ascontiguousarray(pos, np.uint32), max_occ,
                                     np.concatenate([keys, absent]))
        assert bad == 0
        assert found >= len(keys)
