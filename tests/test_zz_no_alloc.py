"""gl_ntt_dev_noalloc (the entry point that allocates nothing) against
gl_ntt_dev through the C ABI: two calls in a row, the second with a smaller
input, forward and inverse; sizes past its scratch are rejected."""
import ctypes

import numpy as np
import pytest

from gpu_util import from_torch, note

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gl():
    import goldilocks_amd as g
    assert g.lib().gl_init(-1) == 0, g.lib().gl_last_error()
    return g


def _noalloc(gl, t, n, count, inverse):
    return gl.lib().gl_ntt_dev_noalloc(ctypes.c_void_p(t.data_ptr()), n, count, int(inverse), None)


def _alloc(gl, t, n, count, inverse):
    return gl.lib().gl_ntt_dev(ctypes.c_void_p(t.data_ptr()), n, count, int(inverse), None)


def test_noalloc_equals_dev_two_calls(gl, oracle):
    import torch
    # (n, count): the first call is the larger; both lengths are warmed by
    # the allocating entry point (the no-allocation call builds no tables)
    for (n1, c1), (n2, c2) in (((1 << 24, 1), (1 << 20, 3)), ((3 << 18, 4), (257 << 8, 2)), ((1 << 16, 7), (15 << 6, 5))):
        for inverse in (0, 1):
            note(f"noalloc n1={n1}x{c1} n2={n2}x{c2} inverse={inverse}")
            outs = []
            for n, c, seed in ((n1, c1, 11), (n2, c2, 12)):
                x = torch.empty(n * c, dtype=torch.int64, device="cuda")
                gl.fill_random(x, seed + n)
                ref = x.clone()
                assert _alloc(gl, ref, n, c, inverse) == 0, gl.lib().gl_last_error()
                outs.append((x, ref, n, c))
            torch.cuda.synchronize()
            for x, ref, n, c in outs:  # two no-allocation calls in a row
                assert _noalloc(gl, x, n, c, inverse) == 0, gl.lib().gl_last_error()
            torch.cuda.synchronize()
            for x, ref, n, c in outs:
                assert torch.equal(x, ref), (n, c, inverse)
    # one small case against the oracle as well
    x = torch.empty(4096, dtype=torch.int64, device="cuda")
    gl.fill_random(x, 5)
    host = from_torch(x)
    assert _alloc(gl, x.clone(), 4096, 1, 0) == 0  # builds the length's tables
    assert _noalloc(gl, x, 4096, 1, 0) == 0, gl.lib().gl_last_error()
    torch.cuda.synchronize()
    assert np.array_equal(from_torch(x), oracle.ntt(host))


def test_noalloc_rejects(gl):
    import torch
    x = torch.empty(1 << 12, dtype=torch.int64, device="cuda")
    # beyond the scratch: 2^25 words (pointer is never dereferenced)
    assert _noalloc(gl, x, 1 << 24, 2, 0) == 4  # GL_E_ARG
    assert b"2^24" in gl.lib().gl_last_error()
    # n * count words beyond the address space
    assert _noalloc(gl, x, 1 << 12, 1 << 62, 0) == 4
    assert b"overflow" in gl.lib().gl_last_error()
    # a length with a factor 65537
    assert _noalloc(gl, x, 65537, 1, 0) == 5  # GL_E_UNSUPPORTED
    # a length whose tables were never built
    gl.clear_cache()
    rc = _noalloc(gl, x, 1 << 12, 1, 0)
    assert rc == 5 and b"not resident" in gl.lib().gl_last_error(), gl.lib().gl_last_error()
    assert _alloc(gl, x, 1 << 12, 1, 0) == 0
    assert _noalloc(gl, x, 1 << 12, 1, 0) == 0
    torch.cuda.synchronize()
