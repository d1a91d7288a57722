"""The register sort network of the wave chain (csrc/device/chain_wave.hpp, reg_sort) restated
lane by lane: key e in slot e / 64 of lane e % 64, partner e ^ m with m = k - 1 on a stage's
first step and j after it, the lower index keeping the smaller key. It must sort every size up
to 64 * R for R = 1, 2, 4 with ~0 padding, duplicates included; the kernel's own results are
checked against the oracle by the GPU chaining tests."""
from __future__ import annotations

import numpy as np
import pytest

PAD = np.uint64(0xFFFFFFFFFFFFFFFF)


def reg_sort(v):
    """v[r][lane] holds key 64 r + lane; returns the registers after the network."""
    R = v.shape[0]
    N = 64 * R
    k = 2
    while k <= N:
        j = k >> 1
        while j > 0:
            m = k - 1 if j == (k >> 1) else j
            ms, ml = m >> 6, m & 63
            pv = np.empty_like(v)
            for r in range(R):
                src = v[r ^ ms]
                pv[r] = src[np.arange(64) ^ ml] if ml else src
            nv = v.copy()
            for r in range(R):
                for lane in range(64):
                    lower = (r & (j >> 6)) == 0 if j >= 64 else (lane & j) == 0
                    lt = v[r][lane] < pv[r][lane]
                    nv[r][lane] = v[r][lane] if lower == lt else pv[r][lane]
            v = nv
            j >>= 1
        k <<= 1
    return v


@pytest.mark.parametrize("R", [1, 2, 4])
def test_reg_sort_network_sorts(R):
    rng = np.random.default_rng(R)
    sizes = [0, 1, 64 * R - 1, 64 * R] + [int(x) for x in rng.integers(0, 64 * R + 1, 12)]
    for t, n in enumerate(sizes):
        a = rng.integers(0, 1 << 40 if t % 2 else 50, n, dtype=np.uint64)
        v = np.full((R, 64), PAD, dtype=np.uint64)
        for e in range(n):
            v[e // 64][e % 64] = a[e]
        out = reg_sort(v)
        got = np.array([out[e // 64][e % 64] for e in range(n)], dtype=np.uint64)
        assert np.array_equal(got, np.sort(a)), (R, n)
