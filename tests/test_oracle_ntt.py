"""Pins the oracle's NTT against the exact DFT (Python big-int at small n,
naive O(n^2) C at medium n, sampled O(n) evaluation and closed forms at large
n), mirroring the reference's test_ntt contract (ntt/src/ntt/mod.rs:128-139)."""
import random

import numpy as np
import pytest

P = 0xFFFFFFFF00000001


def bigint_dft(x, inverse=False):
    n = len(x)
    w = pow(2717, (P - 1) // n, P)
    if inverse:
        w = pow(w, P - 2, P)
    out = []
    for k in range(n):
        wk = pow(w, k, P)
        acc, wj = 0, 1
        for v in x:
            acc = (acc + int(v) * wj) % P
            wj = wj * wk % P
        out.append(acc * (pow(n, P - 2, P) if inverse else 1) % P)
    return out


def small_sizes(oracle, limit):
    return [n for n in oracle.divisors() if 2 <= n <= limit]


def test_naive_vs_bigint(oracle):
    for n in small_sizes(oracle, 64):
        x = oracle.fill_random(n, n)
        assert list(oracle.naive_ntt(x)) == bigint_dft(x), n
        assert list(oracle.naive_intt(x)) == bigint_dft(x, inverse=True), n


def test_fast_vs_naive_all_small(oracle):
    # every legal size up to 4096 (every codelet size, CT, GT, Rader 5/17/257)
    for n in small_sizes(oracle, 4096):
        x = oracle.fill_random(n, 1000 + n)
        assert np.array_equal(oracle.ntt(x), oracle.naive_ntt(x)), n


def test_fast_vs_naive_composites(oracle):
    for n in (257 * 16, 257 * 3, 17 * 256, 15 * 256, 5 * 2**10, 3 * 2**12, 2**13):
        x = oracle.fill_random(n, n)
        assert np.array_equal(oracle.ntt(x), oracle.naive_ntt(x)), n


def test_roundtrip_and_intt(oracle):
    for n in (1, 2, 3, 5, 17, 64, 257, 1024, 2**15, 2**16 + 0, 3 * 2**15):
        x = oracle.fill_random(n, 5)
        y = oracle.ntt(x)
        assert np.array_equal(oracle.intt(y), x), n
        if n <= 1024:
            assert np.array_equal(oracle.intt(x), oracle.naive_intt(x)), n


def test_closed_forms(oracle):
    for n in (16, 240, 257 * 4, 2**16, 65537):
        e0 = np.zeros(n, dtype=np.uint64)
        e0[0] = 1
        assert np.all(oracle.ntt(e0) == 1)
        ones = np.ones(n, dtype=np.uint64)
        expect = np.zeros(n, dtype=np.uint64)
        expect[0] = n % P
        assert np.array_equal(oracle.ntt(ones), expect)
        # delta at 1 gives the powers of the root
        e1 = np.zeros(n, dtype=np.uint64)
        e1[1] = 1
        w = oracle.root(n)
        got = oracle.ntt(e1)
        for k in (0, 1, 2, n // 2, n - 1):
            assert int(got[k]) == pow(w, k, P)


def test_linearity(oracle):
    n = 2**14 * 3
    x = oracle.fill_random(n, 1)
    y = oracle.fill_random(n, 2)
    s = np.array([(int(a) + int(b)) % P for a, b in zip(x, y)], dtype=np.uint64)
    fx, fy, fs = oracle.ntt(x), oracle.ntt(y), oracle.ntt(s)
    assert all((int(a) + int(b)) % P == int(c) for a, b, c in zip(fx[:4096], fy[:4096], fs[:4096]))


@pytest.mark.parametrize("n", [2**16, 2**20, 17 * 2**14, 257 * 2**10, 15 * 2**16])
def test_large_sampled(oracle, n):
    """n > 2^15 takes the on-the-fly twiddle path (cooley_tukey.rs:113-138)."""
    x = oracle.fill_random(n, 11)
    y = oracle.ntt(x)
    rng = random.Random(n)
    for k in [0, 1, n - 1, n // 2] + [rng.randrange(n) for _ in range(4)]:
        assert int(y[k]) == oracle.eval_dft(x, k), (n, k)


def test_batch_and_errors(oracle):
    n, count = 240, 7
    x = oracle.fill_random(n * count, 3)
    y = oracle.ntt_batch(x, n)
    for t in range(count):
        assert np.array_equal(y[t * n:(t + 1) * n], oracle.ntt(x[t * n:(t + 1) * n]))
    with pytest.raises(ValueError):
        oracle.ntt(np.zeros(7, dtype=np.uint64))
    with pytest.raises(ValueError):
        oracle.ntt(np.zeros(0, dtype=np.uint64))
    # intt accepts n <= 1 as a no-op (mod.rs:105-107)
    assert oracle.intt(np.zeros(0, dtype=np.uint64)).size == 0
    assert list(oracle.ntt(np.array([5], dtype=np.uint64))) == [5]


def test_convolution_kat(oracle):
    """convolve.rs:44-60: circular([1,2,3],[4,5,6]) == [31,31,28] (values)."""
    a = np.array([1, 2, 3], dtype=np.uint64)
    b = np.array([4, 5, 6], dtype=np.uint64)
    fa, fb = oracle.ntt(a), oracle.ntt(b)
    prod = np.array([oracle.mul(int(u), int(v)) for u, v in zip(fa, fb)], dtype=np.uint64)
    assert list(oracle.intt(prod)) == [31, 31, 28]


def test_convolve(oracle):
    """or_convolve (convolve.rs:13-20) on Montgomery words: the KAT
    [1,2,3]*[4,5,6] = [31,31,28] (convolve.rs:44-60) and, for small n, the
    cyclic sum of Field (Montgomery) products computed with big integers."""
    mont = lambda v: np.array([(x << 64) % P for x in v], dtype=np.uint64)
    assert list(oracle.convolve(mont([1, 2, 3]), oracle.ntt(mont([4, 5, 6])))) == list(mont([31, 31, 28]))
    rinv = pow(2**64, -1, P)
    for n in (1, 2, 3, 5, 16, 17, 60, 64, 257):
        a = oracle.fill_random(n, 40 + n)
        b = oracle.fill_random(n, 80 + n)
        want = [sum(int(a[j]) * int(b[(i - j) % n]) for j in range(n)) * rinv % P for i in range(n)]
        assert list(oracle.convolve(a, oracle.ntt(b))) == want, n


@pytest.mark.slow
def test_rader_65537(oracle):
    """rader.rs:125-129 skips this (naive takes minutes); sampled DFT instead."""
    n = 65537
    x = oracle.fill_random(n, 9)
    y = oracle.ntt(x)
    for k in (0, 1, 3, 65536, 40000):
        assert int(y[k]) == oracle.eval_dft(x, k)
