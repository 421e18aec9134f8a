"""Proof-of-work search and circular convolution (reference KATs)."""
import os

import numpy as np
import pytest

from gpu_util import from_torch, note, to_torch

pytestmark = pytest.mark.gpu
P = 0xFFFFFFFF00000001


@pytest.fixture(scope="module")
def gl():
    import goldilocks_amd as g
    assert g.lib().gl_init(-1) == 0, g.lib().gl_last_error()
    return g


def test_pow_kats(gl):
    ch = bytes([0x55] * 32)
    note("pow 0/1/16")
    # pcs/src/pow.rs:67-99
    assert all(gl.pow_verify(ch, 0, i) for i in range(4))
    assert [gl.pow_verify(ch, 1 << 63, i) for i in range(4)] == [False, False, False, True]
    assert gl.pow_solve(ch, 1 << 63) == 3
    assert gl.pow_solve(ch, gl.threshold_from_bits(16.0)) == 18455


def test_pow_32_bits(gl):
    """pcs/src/pow.rs:101-106: 4.4e8 hashes, the reference's slow KAT."""
    note("pow 32 bits")
    ch = bytes([0xAA] * 32)
    assert gl.pow_solve(ch, gl.threshold_from_bits(32.0)) == 440258853


def test_convolve_kat(gl, oracle):
    """ntt/src/convolve.rs:44-60 on Montgomery words: [1,2,3] * [4,5,6] = [31,31,28]."""
    note("convolve KAT")
    mont = lambda v: np.array([(x << 64) % P for x in v], dtype=np.uint64)
    a = mont([1, 2, 3])
    b = mont([4, 5, 6])
    gl.ntt(b)
    gl.convolve(a, b)
    assert list(a) == list(mont([31, 31, 28]))


def test_convolve_vs_naive(gl, oracle):
    note("convolve random")
    rinv = pow(2**64, -1, P)
    for n in (64, 1 << 12, 3 << 10, 257):
        a = oracle.fill_random(n, n)
        b = oracle.fill_random(n, n + 1)
        expect = [0] * n
        if n <= 257:
            for i in range(n):
                acc = 0
                for j in range(n):
                    acc += int(a[j]) * int(b[(i - j) % n])
                expect[i] = acc * rinv % P  # Field product of Montgomery words
        bn = b.copy()
        gl.ntt(bn)
        got = a.copy()
        gl.convolve(got, bn)
        if n <= 257:
            assert list(got) == expect
        ref = oracle.ntt(a.copy())
        bo = oracle.ntt(b.copy())
        prod = np.array([oracle.mont_mul(int(x), int(y)) for x, y in zip(ref, bo)], dtype=np.uint64)
        assert np.array_equal(got, oracle.intt(prod))


def _last_kernel(gl, n):
    info = gl.plan_info(n)
    return info[-1]["kernel"] if info else "none"


def test_convolve_fused_and_separate_pointwise(gl, oracle):
    """The pointwise product rides on the forward transform's last pass when
    that is one two- or three-stage kernel (2^12, 257*4 and 2^24: two-stage;
    2^19, 2^20: three-stage), and is a separate kernel after a radix-257 or
    radix-65537 last pass (257, 65537) or with no pass (n = 1): every result
    equals the oracle's ntt, Montgomery product, intt; device and host entry
    points agree."""
    import torch
    oracle.set_threads(min(16, os.cpu_count() or 1))
    try:
        kinds = {}
        for n in (1, 1 << 12, 257 * 4, 257, 65537, 1 << 19, 1 << 20, 1 << 24):
            kinds[n] = _last_kernel(gl, n)
            note(f"convolve n={n} last pass {kinds[n]}")
            a = oracle.fill_random(n, 3 * n + 1)
            b = oracle.fill_random(n, 3 * n + 2)
            bo = oracle.ntt(b.copy())
            want = oracle.convolve(a, bo)
            da, db = to_torch(a), to_torch(bo)
            gl.convolve(da, db)
            torch.cuda.synchronize()
            assert np.array_equal(from_torch(da), want), n
            if n <= 1 << 16:
                got = a.copy()
                gl.convolve(got, bo)
                assert np.array_equal(got, want), n
        assert "pass3" in kinds[1 << 20] and "pass_kernel" in kinds[1 << 24] and "rader257" in kinds[257]
    finally:
        oracle.set_threads(1)

