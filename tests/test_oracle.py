"""Pins the C oracle: closed forms, brute force, textbook DP, invariants, the reference's own KATs
(src/../tests/test_main.cpp) and the golden records written by the reference binary."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from dnaln_amd import synth
from tests import golden_io
from tests.golden import sets


# ---------------------------------------------------------------------------------------- hash64
def _hash64_py(key: int, mask: int) -> int:
    """The Thomas Wang 64-bit mix restricted to `mask` (minimizer.hpp:21-30), in Python ints."""
    key = (~key + (key << 21)) & mask
    key ^= key >> 24
    key = (key + (key << 3) + (key << 8)) & mask
    key ^= key >> 14
    key = (key + (key << 2) + (key << 4)) & mask
    key ^= key >> 28
    key = (key + (key << 31)) & mask
    return key


def test_hash64_closed_form_and_bijection():
    rng = np.random.default_rng(0)
    for k in (4, 5, 8, 15, 28):
        mask = (1 << (2 * k)) - 1
        for key in rng.integers(0, mask + 1, 200, dtype=np.uint64):
            assert oracle.hash64(int(key), mask) == _hash64_py(int(key), mask)
    for k in (4, 5, 6):  # a bijection on [0, mask]
        mask = (1 << (2 * k)) - 1
        vals = {oracle.hash64(x, mask) for x in range(mask + 1)}
        assert vals == set(range(mask + 1))


def test_encode_revcomp_kat():
    # test_main.cpp:155-166: "ACGTNacgtx" -> 0 1 2 3 4 0 1 2 3 4
    import ctypes as C
    s = b"ACGTNacgtx"
    out = np.zeros(len(s), np.uint8)
    oracle.lib().or_encode(s, len(s), out.ctypes.data_as(C.c_void_p))
    assert out.tolist() == [0, 1, 2, 3, 4, 0, 1, 2, 3, 4]
    rc = np.zeros(len(s), np.uint8)
    oracle.lib().or_revcomp(out.ctypes.data_as(C.c_void_p), len(s), rc.ctypes.data_as(C.c_void_p))
    assert rc.tolist() == [4, 0, 1, 2, 3, 4, 0, 1, 2, 3]


def test_fnv1a_kat():
    # FNV-1a 32-bit published vectors
    assert oracle.fnv1a(b"") == 0x811C9DC5
    assert oracle.fnv1a(b"a") == 0xE40C292C
    assert oracle.fnv1a(b"foobar") == 0xBF9CF968


# ---------------------------------------------------------------------------------------- sketch
def _sketch_brute(s: np.ndarray, k: int, w: int, offset: int):
    """test_main.cpp:184-219: leftmost valid window minimum inside N-free runs, deduplicated."""
    mask = (1 << (2 * k)) - 1
    runs = [[]]
    for i in range(len(s)):
        if s[i] > 3:
            runs.append([])
            continue
        if i + 1 < k or any(s[i - j] > 3 for j in range(k)):
            continue
        f = r = 0
        for j in range(k):
            f = (f << 2) | int(s[i + 1 - k + j])
            r = (r << 2) | (3 - int(s[i - j]))
        runs[-1].append((i + 1 - k, _hash64_py(min(f, r), mask), 1 if r < f else 0, f != r))
    exp = []
    for run in runs:
        for st in range(0, len(run) - w + 1):
            best = None
            for j in range(st, st + w):
                if run[j][3] and (best is None or run[j][1] < best[1]):
                    best = run[j]
            if best and (not exp or exp[-1][0] != best[0] + offset):
                exp.append((best[0] + offset, best[1], best[2]))
    return exp


def test_sketch_matches_brute_force():
    rng = np.random.default_rng(7)
    for trial in range(120):
        k, w = int(rng.integers(5, 25)), int(rng.integers(1, 21))
        n = int(rng.integers(50, 300))
        s = rng.integers(0, 4, n, dtype=np.uint8)
        if trial % 3 == 0:
            s[rng.random(n) < 0.01] = 4
        got = oracle.sketch(s, k, w, 1000)
        exp = _sketch_brute(s, k, w, 1000)
        assert [(int(m["pos"]), int(m["hash"]), int(m["strand"])) for m in got] == exp, (trial, k, w)


def test_sketch_even_k_palindromes_never_selected():
    # ACGT is its own reverse complement: with k=4 every k-mer of (ACGT)* alternates with one.
    s = np.array([0, 1, 2, 3] * 20, np.uint8)
    got = oracle.sketch(s, 4, 3, 0)
    assert [(int(m["pos"]), int(m["hash"]), int(m["strand"])) for m in got] == _sketch_brute(s, 4, 3, 0)
    assert all(int(m["pos"]) % 4 != 0 for m in got)


def test_sketch_strand_symmetry():
    rng = np.random.default_rng(11)
    s = rng.integers(0, 4, 5000, dtype=np.uint8)
    rc = (3 - s[::-1]).astype(np.uint8)
    a = np.sort(oracle.sketch(s, 15, 10)["hash"])
    b = np.sort(oracle.sketch(rc, 15, 10)["hash"])
    assert len(np.intersect1d(a, b)) >= 0.95 * len(a)


def test_sketch_matches_reference_build():
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("reference build not present")
    rng = np.random.default_rng(5)
    for trial in range(200):
        k, w = int(rng.integers(4, 29)), int(rng.integers(1, 60))
        n = int(rng.integers(0, 2000))
        s = rng.integers(0, 4, n, dtype=np.uint8)
        s[rng.random(n) < (0.02 if trial % 2 else 0.0)] = 4
        h = np.zeros(max(n, 1), np.uint64)
        p = np.zeros(max(n, 1), np.uint32)
        st = np.zeros(max(n, 1), np.uint32)
        m = ref.ref_sketch(s.ctypes.data, n, k, w, 7, h.ctypes.data, p.ctypes.data, st.ctypes.data)
        got = oracle.sketch(s, k, w, 7)
        assert len(got) == m
        assert np.array_equal(got["hash"], h[:m]) and np.array_equal(got["pos"], p[:m])
        assert np.array_equal(got["strand"], st[:m])


# ---------------------------------------------------------------------------------------- index
def test_index_finds_every_minimizer_and_matches_dictionary():
    rng = np.random.default_rng(3)
    parts, off, ln = [], [], []
    o = 0
    for c in range(3):
        n = 20000 + int(rng.integers(5000))
        s = rng.integers(0, 4, n, dtype=np.uint8)
        s[rng.random(n) < 0.001] = 4
        parts.append(s)
        off.append(o)
        ln.append(n)
        o += n
    codes = np.concatenate(parts)
    idx = oracle.Index(codes, np.array(off, np.uint64), np.array(ln, np.uint32), 15, 10)
    d: dict[int, list[int]] = {}
    total = 0
    for c in range(3):
        for m in oracle.sketch(codes[off[c]:off[c] + ln[c]], 15, 10, off[c]):
            d.setdefault(int(m["hash"]), []).append(int(m["pos"]) << 1 | int(m["strand"]))
            total += 1
    assert idx.idx.n_pos == total and idx.idx.n_keys == len(d)
    for h, v in d.items():
        assert idx.lookup(h).tolist() == sorted(v)
    assert len(idx.lookup(1 << 40)) == 0
    # load factor <= 0.5 and power-of-two capacity
    assert idx.idx.cap >= 2 * idx.idx.n_keys and idx.idx.cap & (idx.idx.cap - 1) == 0


# ---------------------------------------------------------------------------------------- SW
def _naive_banded_sw(q, t_in, sp) -> int:
    """test_main.cpp:84-104: full-matrix affine-gap local DP, cells outside the band are H=0, E=F=-inf."""
    ql, W = len(q), 2 * sp.bw + 1
    tl = ql + W - 1
    t = np.full(max(tl, 0), 4, np.uint8)
    t[:min(len(t_in), tl)] = t_in[:min(len(t_in), tl)]
    NEG = -10 ** 9
    oe, ge = sp.gap_open + sp.gap_ext, sp.gap_ext
    H = [[0] * (tl + 1) for _ in range(ql + 1)]
    E = [[NEG] * (tl + 1) for _ in range(ql + 1)]
    F = [[NEG] * (tl + 1) for _ in range(ql + 1)]
    best = 0
    for i in range(1, ql + 1):
        for j in range(1, tl + 1):
            d = j - i
            if d < 0 or d >= W:
                continue
            E[i][j] = max(H[i - 1][j] - oe, E[i - 1][j] - ge)
            F[i][j] = max(H[i][j - 1] - oe, F[i][j - 1] - ge)
            a, b = int(q[i - 1]), int(t[j - 1])
            s = -sp.ambig if (a > 3 or b > 3) else (sp.match if a == b else -sp.mismatch)
            H[i][j] = max(0, H[i - 1][j - 1] + s, E[i][j], F[i][j])
            best = max(best, H[i][j])
    return best


@pytest.mark.parametrize("si", range(4))
def test_scalar_kernel_matches_naive_dp(si):
    sc = sets.SCHEMES[si]
    sp = oracle.score_t(**sc)
    rng = np.random.default_rng(50 + si)
    for q, t in sets.random_tasks(rng, 60, sc["bw"], max_len=50):
        assert oracle.align(sp, q, t, False)[0] == _naive_banded_sw(q, t, sp)


@pytest.mark.parametrize("si", range(len(sets.SCHEMES)))
def test_traceback_is_consistent(si):
    sc = sets.SCHEMES[si]
    sp = oracle.score_t(**sc)
    rng = np.random.default_rng(90 + si)
    for q, t in sets.random_tasks(rng, 200, sc["bw"]):
        score, qb, qe, tb, te, nm, cig = oracle.align(sp, q, t, True)
        assert score == oracle.align(sp, q, t, False)[0]
        if score == 0:
            assert (qb, qe, tb, te, nm, cig) == (0, 0, 0, 0, 0, [])
            continue
        assert oracle.score_cigar(sp, q, t, qb, tb, cig) == score
        qspan = sum(c >> 4 for c in cig if c & 15 != 2)
        tspan = sum(c >> 4 for c in cig if c & 15 != 1)
        assert qspan == qe - qb and tspan == te - tb
        assert cig[0] & 15 == 0 and cig[-1] & 15 == 0
        # NM = mismatches (N counts) + gap lengths
        i, j, exp_nm = qb, tb, 0
        for c in cig:
            L, op = c >> 4, c & 15
            if op == 0:
                exp_nm += sum(1 for x in range(L) if q[i + x] != t[j + x] or q[i + x] > 3)
                i += L
                j += L
            elif op == 1:
                exp_nm += L
                i += L
            else:
                exp_nm += L
                j += L
        assert nm == exp_nm


def test_perfect_match_kat():
    # test_main.cpp:340-360
    sp = oracle.score_t()
    rng = np.random.default_rng(17)
    q = rng.integers(0, 4, 150, dtype=np.uint8)
    t = np.concatenate([rng.integers(0, 4, 16, dtype=np.uint8), q, rng.integers(0, 4, 16, dtype=np.uint8)])
    score, qb, qe, tb, te, nm, cig = oracle.align(sp, q, t, True)
    assert (score, qb, qe, tb, nm, oracle.cigar_str(cig)) == (300, 0, 150, 16, 0, "150M")


def test_kernel_vectors_match_reference_fixture():
    kv = golden_io.kernel_vectors()
    for si in range(len(sets.SCHEMES)):
        sp = oracle.score_t(*[int(x) for x in kv[f"s{si}_sp"]])
        pool = kv[f"s{si}_pool"]
        qo, ql, to, tl = kv[f"s{si}_qoff"], kv[f"s{si}_qlen"], kv[f"s{si}_toff"], kv[f"s{si}_tlen"]
        for tb in (0, 1):
            res, cig, cn = kv[f"s{si}_tb{tb}_res"], kv[f"s{si}_tb{tb}_cig"], kv[f"s{si}_tb{tb}_cn"]
            co = np.concatenate([[0], np.cumsum(cn)])
            for i in range(len(ql)):
                q = pool[qo[i]:qo[i] + ql[i]]
                t = pool[to[i]:to[i] + tl[i]]
                got = oracle.align(sp, q, t, bool(tb))
                assert tuple(got[:6]) == tuple(int(x) for x in res[i]), (si, tb, i)
                assert got[6] == [int(x) for x in cig[co[i]:co[i + 1]]], (si, tb, i)


# ---------------------------------------------------------------------------------------- chaining
def test_chain_kat():
    # test_main.cpp:362-375: 10 colinear anchors with a 3 bp deletion plus 2 noise anchors
    a = [(1000 + i * 12 + (3 if i >= 5 else 0), i * 12, 0) for i in range(10)] + [(50000, 30, 0), (90000, 60, 1)]
    _, chains = oracle.chain(np.array(a, dtype=[("rpos", "u4"), ("qpos", "i4"), ("strand", "u4")]))
    assert chains and chains[0][2] == 10 and chains[0][3] == 1000 and chains[0][4] == 1003


def _chain_brute(a, k, max_gap=5000, band=16, lookback=50, min_score=20, min_anchors=2):
    """The DP and extraction of chain.cpp, in Python, on anchors already sorted."""
    import math
    n = len(a)
    f, prev = [0] * n, [-1] * n
    for i in range(n):
        best, bj = k, -1
        for j in range(i - 1, max(0, i - lookback) - 1, -1):
            if a[j][2] != a[i][2]:
                break
            dr = a[i][0] - a[j][0]
            if dr > max_gap:
                break
            dq = a[i][1] - a[j][1]
            if dr <= 0 or dq <= 0:
                continue
            dd = abs(dr - dq)
            if dd > band:
                continue
            cost = int(0.01 * k * dd + 0.5 * int(math.log2(dd))) if dd else 0
            sc = f[j] + min(dq, dr, k) - cost
            if sc > best:
                best, bj = sc, j
        f[i], prev[i] = best, bj
    used = [False] * n
    out = []
    for end in sorted(range(n), key=lambda x: (-f[x], x)):
        if used[end]:
            continue
        i, cnt, ds = end, 0, []
        while i >= 0 and not used[i]:
            used[i] = True
            cnt += 1
            ds.append(a[i][0] - a[i][1])
            i = prev[i]
        score = f[end] - (f[i] if i >= 0 else 0)
        if score >= min_score and cnt >= min_anchors:
            out.append((score, a[end][2], cnt, min(ds), max(ds), a[end][0]))
    out.sort(key=lambda c: -c[0])  # stable
    return out


def test_chain_matches_brute_force_and_reference():
    rng = np.random.default_rng(21)
    ref = oracle.ref_lib()
    for trial in range(150):
        n = int(rng.integers(0, 120))
        base = int(rng.integers(0, 100000))
        rows = []
        for _ in range(n):
            if rng.random() < 0.7:
                q = int(rng.integers(0, 300))
                rows.append((base + q + int(rng.integers(-20, 21)) + int(rng.integers(0, 3)) * 1000, q,
                             int(rng.integers(0, 2))))
            else:
                rows.append((int(rng.integers(0, 200000)), int(rng.integers(0, 300)), int(rng.integers(0, 2))))
        rows = list({(r, q, s): None for r, q, s in rows})  # unique keys, as anchors are
        arr = np.array(rows, dtype=[("rpos", "u4"), ("qpos", "i4"), ("strand", "u4")]) if rows else \
            np.zeros(0, dtype=[("rpos", "u4"), ("qpos", "i4"), ("strand", "u4")])
        k = int(rng.integers(10, 20))
        srt, chains = oracle.chain(arr, k)
        srt_list = [(int(x["rpos"]), int(x["qpos"]), int(x["strand"])) for x in srt]
        assert srt_list == sorted(srt_list, key=lambda t: (t[2], t[0], t[1]))
        assert chains == _chain_brute(srt_list, k), trial
        if ref is not None and n:
            rp, qp, st = arr["rpos"].copy(), arr["qpos"].copy(), arr["strand"].copy()
            out = np.zeros((n, 6), np.int64)
            cp = np.array([5000, 16, 50, 20, 2], np.int32)
            m = ref.ref_chain(rp.ctypes.data, qp.ctypes.data, st.ctypes.data, n, k, cp.ctypes.data, out.ctypes.data)
            assert [tuple(int(x) for x in r) for r in out[:m]] == [tuple(c) for c in chains]


# ---------------------------------------------------------------------------------------- mapper
@pytest.mark.parametrize("name", list(sets.SETS))
def test_oracle_mapper_matches_reference_records(name):
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    mp = sets.map_params(spec)
    idx = oracle.Index(g.codes, g.contig_off, g.contig_len, mp["k"], mp["w"])
    recs, pool, _ = oracle.map_reads(oracle.default_params(**mp), g, idx, r)
    gold = golden_io.records(name)
    assert len(gold) == len(r)
    bad = [i for i in range(len(r)) if oracle.record_fields(recs[i], pool, g.names) != gold[i][1:]]
    assert not bad, (name, len(bad), bad[:5])


def test_sam_formatting_matches_reference_digest():
    """or_format_sam reproduces the reference's SAM body byte for byte (manifest digest)."""
    import ctypes as C
    import hashlib
    name = "edge_uniform_bw16"
    spec = sets.SETS[name]
    g = sets.genome(spec)
    r = sets.reads(spec, g)
    idx = oracle.Index(g.codes, g.contig_off, g.contig_len)
    recs, pool, _ = oracle.map_reads(oracle.default_params(**sets.map_params(spec)), g, idx, r)
    h = hashlib.sha256()
    buf = C.create_string_buffer(1 << 16)
    for i in range(len(r)):
        s = r.read(i)
        rec = oracle.RecordT.from_buffer_copy(recs[i].tobytes())
        nm = f"r{i}".encode()
        cname = g.names[rec.contig].encode() if rec.contig >= 0 else b""
        n = oracle.lib().or_format_sam(nm, len(nm), s, b"I" * len(s), len(s), C.byref(rec), pool.ctypes.data,
                                       cname, buf, len(buf))
        h.update(buf.raw[:n])
    assert h.hexdigest() == golden_io.manifest()["sets"][name]["sam_body_sha256"]
