"""Every legal length up to 2^24 words (the 301 divisors of p - 1 up to
16777216, radix 65537 included): forward NTT compared element by element
with the oracle, inverse back to the input, both on device-resident words.
The reference's own tests use the exact DFT as the contract
(ntt/src/ntt/mod.rs:128-139); the oracle is pinned to it for n <= 4096 and
by sampled X[k] above."""
import os

import numpy as np
import pytest

from gpu_util import from_torch, note, to_torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gl(oracle):
    import goldilocks_amd as g
    assert g.lib().gl_init(-1) == 0, g.lib().gl_last_error()
    oracle.set_threads(min(16, os.cpu_count() or 1))
    yield g
    oracle.set_threads(1)


def test_every_divisor_up_to_2_22(gl, oracle):
    import torch
    sizes = [n for n in oracle.divisors() if n <= 1 << 22]
    assert len(sizes) == 254
    for n in sizes:
        note(f"divisor n={n}")
        x = oracle.fill_random(n, 5000 + n % 9973)
        t = to_torch(x)
        gl.ntt(t)
        torch.cuda.synchronize()
        assert np.array_equal(from_torch(t), oracle.ntt(x)), f"forward n={n}"
        gl.intt(t)
        torch.cuda.synchronize()
        assert np.array_equal(from_torch(t), x), f"roundtrip n={n}"


def test_every_divisor_from_2_22_to_2_24(gl, oracle):
    import torch
    sizes = [n for n in oracle.divisors() if 1 << 22 < n <= 1 << 24]
    assert len(sizes) == 47
    for n in sizes:
        note(f"divisor n={n}")
        x = oracle.fill_random(n, 7000 + n % 9973)
        t = to_torch(x)
        gl.ntt(t)
        torch.cuda.synchronize()
        assert np.array_equal(from_torch(t), oracle.ntt(x)), f"forward n={n}"
        gl.intt(t)
        torch.cuda.synchronize()
        assert np.array_equal(from_torch(t), x), f"roundtrip n={n}"
