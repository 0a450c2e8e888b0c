"""Pins the oracle's sampler and sin/cos against closed forms and independent restatements."""

from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O

PRIMES = [3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59]


def radical_inverse(base: int, n: int) -> Fraction:
    """Van der Corput radical inverse, exact."""
    r, f = Fraction(0), Fraction(1, base)
    while n > 0:
        r += f * (n % base)
        n //= base
        f /= base
    return r


def test_halton_known_answers():
    # block 0, base 3 -> 1/3, 2/3, 1/9, 4/9, 7/9 (prrtc_planner.cu:144-157, state starts at 1)
    want = [1 / 3, 2 / 3, 1 / 9, 4 / 9, 7 / 9, 2 / 9, 5 / 9, 8 / 9]
    got = [O.halton(3, n) for n in range(1, 9)]
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-7)
    assert O.halton(2, 1) == 0.5 and O.halton(2, 2) == 0.25 and O.halton(2, 3) == 0.75
    assert O.halton(5, 0) == 0.0


@pytest.mark.parametrize("base", PRIMES)
def test_halton_matches_radical_inverse(base):
    for n in list(range(1, 300)) + [1000, 4096, 10_000, 99_999, 1_000_000]:
        exact = float(radical_inverse(base, n))
        got = O.halton(base, n)
        # FP32 accumulation of at most log_base(n) terms: a few ulp of 1.0
        assert abs(got - exact) <= 4 * np.finfo(np.float32).eps, (base, n, got, exact)


def _lcg_shuffle(explorer: int, seed: int) -> list[int]:
    """Independent restatement of the per-block prime shuffle (prrtc_planner.cu:745-765)."""
    p = list(PRIMES)
    if explorer == 0 and seed == 0:
        return p
    rng = ((explorer + seed * 65536) * 1664525 + 1013904223) % 2**32
    for i in range(15, 0, -1):
        rng = (rng * 1664525 + 1013904223) % 2**32
        j = rng % (i + 1)
        p[i], p[j] = p[j], p[i]
    return p


@pytest.mark.parametrize("explorer,seed", [(0, 0), (1, 0), (2, 0), (127, 0), (511, 0), (0, 1), (5, 77), (3, 2**31 + 5)])
def test_prime_order(explorer, seed):
    got = list(O.prime_order(explorer, seed))
    assert got == _lcg_shuffle(explorer, seed)
    assert sorted(got) == PRIMES


def test_prime_order_block_zero_is_canonical():
    assert list(O.prime_order(0, 0)) == PRIMES
    assert list(O.prime_order(1, 0)) != PRIMES


def test_sincos_accuracy():
    xs = np.concatenate([np.linspace(-20, 20, 20001), np.array([0.0, -0.0, 1e-30, np.pi / 4, np.pi / 2, np.pi, 3 * np.pi / 2,
                                                                 1e3, -1e3 + 0.25])]).astype(np.float32)
    for x in xs:
        s, c = O.sincos(float(x))
        xd = float(x)
        assert abs(s - math.sin(xd)) <= 3e-7 * max(1.0, abs(xd) / 8), (x, s, math.sin(xd))
        assert abs(c - math.cos(xd)) <= 3e-7 * max(1.0, abs(xd) / 8), (x, c, math.cos(xd))


def test_sincos_symmetries():
    for x in np.linspace(-7, 7, 1001).astype(np.float32):
        s, c = O.sincos(float(x))
        assert abs(s * s + c * c - 1.0) < 5e-7
        s2, c2 = O.sincos(float(-x))
        assert s2 == -s and c2 == c


@pytest.mark.parametrize("n,dim", [(1, 7), (37, 7), (1000, 7), (513, 14)])
def test_nearest_matches_brute_force_with_ties(n, dim):
    """oc_nearest is the lowest-index argmin of the squared distance (prrtc_robogpu.py:82-87
    semantics): brute force in float64, exact duplicates injected so ties must go low."""
    rng = np.random.default_rng(n * 31 + dim)
    max_nodes = n + 5
    pts = rng.uniform(-3, 3, size=(n, dim)).astype(np.float32)
    if n > 10:
        pts[n // 2] = pts[3]          # duplicate node: index 3 must win
        pts[n - 1] = pts[n // 3]      # and another pair
    tree = np.zeros((dim, max_nodes), np.float32)
    tree[:, :n] = pts.T
    tree[:, n:] = -7.0                # beyond size: must be ignored
    qs = rng.uniform(-3, 3, size=(64, dim)).astype(np.float32)
    if n > 10:
        qs[0] = pts[n // 2]           # exact hit on a duplicated node
        qs[1] = pts[n - 1]
    idx, dsq = O.nearest(tree, n, qs)
    d64 = ((qs[:, None, :].astype(np.float64) - pts[None].astype(np.float64)) ** 2).sum(-1)
    want = d64.argmin(axis=1)         # numpy: first (lowest) index among equal minima
    for k in range(qs.shape[0]):
        srt = np.sort(d64[k])
        gap = srt[1] - srt[0] if srt.size > 1 else np.inf
        if gap > 1e-5 * max(srt[0], 1e-12):   # clear winner: FP32 must find it
            assert idx[k] == want[k], k
        else:                                  # FP32 may pick either of two near-equal nodes
            assert d64[k, idx[k]] - srt[0] <= 1e-5 * max(srt[0], 1e-12), k
        assert abs(dsq[k] - d64[k, idx[k]]) <= 1e-5 * max(d64[k, idx[k]], 1.0)
    if n > 10:
        assert idx[0] == 3 and dsq[0] == 0.0
        assert idx[1] == n // 3 and dsq[1] == 0.0
