"""GPU permutations (the reference's permute::transpose and bit_reverse,
ntt/src/permute/mod.rs:53-74) against the oracle and numpy: host and
device paths, in place and out of place, square and rectangular shapes,
shapes that are not multiples of the 32 x 32 tile, and the reference's own
known answers."""
import ctypes

import numpy as np
import pytest

from gpu_util import from_torch, note, to_torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gl():
    import goldilocks_amd as g
    assert g.lib().gl_init(-1) == 0, g.lib().gl_last_error()
    return g


SHAPES = [(1, 1), (1, 9), (9, 1), (2, 3), (3, 3), (31, 33), (33, 65), (100, 37), (64, 64), (1000, 1000),
          (512, 2048), (2048, 512), (3000, 5000), (4096, 4096)]


def _expect(m, rows, cols):
    return np.ascontiguousarray(m.reshape(rows, cols).T).reshape(-1)


def test_transpose_kats(gl):
    # permute/mod.rs test_copy_2_3 and test_copy_3_3
    a = np.arange(6, dtype=np.uint64)
    gl.transpose(a, 2, 3)
    assert list(a) == [0, 3, 1, 4, 2, 5]
    b = np.arange(9, dtype=np.uint64)
    gl.transpose(b, 3, 3)
    assert list(b) == [0, 3, 6, 1, 4, 7, 2, 5, 8]


@pytest.mark.parametrize("rows,cols", SHAPES)
def test_transpose_host_and_device(gl, oracle, rows, cols):
    import torch
    note(f"transpose {rows}x{cols}")
    m = oracle.fill_random(rows * cols, rows * 7919 + cols)
    want = _expect(m, rows, cols)
    if rows * cols <= 1 << 20:
        assert np.array_equal(want, oracle.transpose(m, rows, cols))
    h = m.copy()
    gl.transpose(h, rows, cols)
    assert np.array_equal(h, want), "host in place"
    t = to_torch(m)
    gl.transpose(t, rows, cols)
    torch.cuda.synchronize()
    assert np.array_equal(from_torch(t), want), "device in place"
    src = to_torch(m)
    dst = torch.empty_like(src)
    gl.transpose(src, rows, cols, out=dst)
    torch.cuda.synchronize()
    assert np.array_equal(from_torch(dst), want), "device out of place"
    assert np.array_equal(from_torch(src), m), "source untouched"
    # back again
    gl.transpose(dst, cols, rows)
    torch.cuda.synchronize()
    assert np.array_equal(from_torch(dst), m), "round trip"


def _rev_index(k):
    n = 1 << k
    i = np.arange(n, dtype=np.uint64)
    r = np.zeros(n, dtype=np.uint64)
    for b in range(k):
        r |= ((i >> np.uint64(b)) & np.uint64(1)) << np.uint64(k - 1 - b)
    return r


@pytest.mark.parametrize("k", [0, 1, 2, 5, 9, 10, 11, 12, 15, 20, 24])
def test_bit_reverse_host_and_device(gl, oracle, k):
    import torch
    n = 1 << k
    note(f"bit_reverse 2^{k}")
    v = oracle.fill_random(n, 300 + k)
    want = v[_rev_index(k)]
    if k <= 16:
        assert np.array_equal(want, oracle.bit_reverse(v))
    h = v.copy()
    gl.bit_reverse(h)
    assert np.array_equal(h, want), "host"
    t = to_torch(v)
    gl.bit_reverse(t)
    torch.cuda.synchronize()
    assert np.array_equal(from_torch(t), want), "device"
    gl.bit_reverse(t)
    torch.cuda.synchronize()
    assert np.array_equal(from_torch(t), v), "involution"


def test_permute_errors(gl):
    import torch
    for n in (0, 3, 6, 1000):
        with pytest.raises(ValueError):
            gl.bit_reverse(np.zeros(n, dtype=np.uint64))
    t = torch.zeros(64, dtype=torch.int64, device="cuda")
    rc = gl.lib().gl_transpose_oop_dev(ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(t.data_ptr()), 8, 8, None)
    assert rc == 4  # GL_E_ARG: out of place needs two buffers
    # empty shapes are no-ops; a shape whose byte size overflows is rejected
    assert gl.lib().gl_transpose_dev(None, 0, 5, None) == 0
    assert gl.lib().gl_transpose_dev(ctypes.c_void_p(t.data_ptr()), 1 << 40, 1 << 40, None) == 4
    assert gl.lib().gl_transpose_dev(ctypes.c_void_p(t.data_ptr()), 1 << 20, 1 << 20, None) == 1  # GL_E_SIZE: > 2^34 words
