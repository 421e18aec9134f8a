"""Host-buffer entry points around the pinned staging of large copies.

Copies above 4 MiB go through 8 MiB pinned chunks, three per set (see
gl_hostio.cpp); these cases put the copy sizes below, at and across those
boundaries (one partial chunk, exactly one chunk, exactly the ring, a ring
that wraps, a partial last chunk) and run host copies from two threads at
once.  Each result is compared with the device-resident path on the same
words (itself checked against the oracle in test_gpu_ntt.py) or with the
oracle directly."""
import threading

import numpy as np
import pytest

from gpu_util import from_torch, note, to_torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gl():
    import goldilocks_amd as g
    assert g.lib().gl_init(-1) == 0, g.lib().gl_last_error()
    return g


def _device_ntt(gl, x, inverse=False):
    import torch
    t = to_torch(x)
    (gl.intt if inverse else gl.ntt)(t)
    torch.cuda.synchronize()
    return from_torch(t)


# words: 4 MiB (pageable path), 6 MiB (one partial chunk), 8 MiB (one chunk),
# 24 MiB (three chunks), 30 MiB (3.75 chunks), 32 MiB (the ring wraps),
# 136 MiB (17 chunks)
SIZES = [2**19, 3 * 2**18, 2**20, 3 * 2**20, 15 * 2**18, 2**22, 17 * 2**20]


@pytest.mark.parametrize("n", SIZES)
def test_host_ntt_across_staging_sizes(gl, oracle, n):
    note(f"host staging ntt n={n}")
    x = oracle.fill_random(n, 77 + n)
    y = x.copy()
    gl.ntt(y)
    assert np.array_equal(y, _device_ntt(gl, x)), f"forward n={n}"
    gl.intt(y)
    assert np.array_equal(y, x), f"roundtrip n={n}"


def test_host_ntt_vs_oracle_staged(gl, oracle):
    n = 3 * 2**18  # staged, and small enough for the oracle
    note(f"host staging vs oracle n={n}")
    x = oracle.fill_random(n, 5)
    y = x.copy()
    gl.ntt(y)
    assert np.array_equal(y, oracle.ntt(x))


def test_host_copies_from_two_threads(gl, oracle):
    cases = [(2**22, 1), (15 * 2**18, 2), (2**20, 3), (3 * 2**20, 4)]
    refs = {}
    for n, seed in cases:
        x = oracle.fill_random(n, 900 + seed)
        refs[seed] = (x, _device_ntt(gl, x))
    errors = []

    def worker(items):
        try:
            for _ in range(3):
                for n, seed in items:
                    x, ref = refs[seed]
                    y = x.copy()
                    gl.ntt(y)
                    if not np.array_equal(y, ref):
                        errors.append(f"forward n={n}")
                    gl.intt(y)
                    if not np.array_equal(y, x):
                        errors.append(f"roundtrip n={n}")
        except Exception as e:  # surfaced below
            errors.append(repr(e))

    note("host staging: two threads")
    ts = [threading.Thread(target=worker, args=(cases[0::2],)), threading.Thread(target=worker, args=(cases[1::2],))]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors


def test_host_commit_and_tree_staged(gl, oracle):
    import torch
    rows, cols = 2**14, 64  # 8 MiB in, 16 MiB of LDE out, tree layers of 3.2 MiB
    note(f"host staging commit {rows}x{cols}")
    m = oracle.fill_random(rows * cols, 31).reshape(rows, cols)
    lde_h, layers_h = gl.commit(m)
    lde_d, layers_d = gl.commit(to_torch(m).view(rows, cols))
    torch.cuda.synchronize()
    assert np.array_equal(lde_h, from_torch(lde_d).reshape(lde_h.shape))
    assert b"".join(layers_h) == layers_d.cpu().numpy().tobytes()
    note("host staging merkle_tree 12 MiB")
    data = oracle.fill_random(3 * 2**19, 32)
    layers = gl.merkle_tree(data)
    assert layers == oracle.merkle_tree(data.tobytes())


# (n, count): batches of at least 64 MiB run as a pipeline of 32 MiB chunks of
# whole transforms (or one larger transform) over two device buffers; these
# take chunks of 4 + 4 + 1, 5 + 5 + 1 and 1 + 1 + 1 transforms, 20 chunks
# that cycle both buffers, and three 64 MiB transforms of one chunk each
PIPELINED = [(2**20, 9), (3 * 2**18, 11), (2**22, 3), (5 * 2**12, 4000), (2**23, 3)]


@pytest.mark.parametrize("n,count", PIPELINED)
def test_host_batch_pipelined(gl, oracle, n, count):
    """strategy(n).ntt on a host array of many transforms (the pipelined
    path) against the device path on the same words, forward and inverse;
    the first and last transforms also against the oracle."""
    import torch
    note(f"pipelined host batch n={n} count={count}")
    x = oracle.fill_random(n * count, 91 + n)
    for inverse in (False, True):
        y = x.copy()
        gl.strategy(n).ntt(y, inverse=inverse)
        t = to_torch(x)
        gl.strategy(n).ntt(t, inverse=inverse)
        torch.cuda.synchronize()
        assert np.array_equal(y, from_torch(t)), (n, count, inverse)
        if n <= 1 << 20:
            for j in (0, count - 1):
                want = (oracle.intt if inverse else oracle.ntt)(x[j * n:(j + 1) * n].copy())
                assert np.array_equal(y[j * n:(j + 1) * n], want), (n, count, inverse, j)
    y = x.copy()
    s = gl.strategy(n)
    s.ntt(y)
    s.ntt(y, inverse=True)
    assert np.array_equal(y, x), (n, count)


def test_pipelined_host_batches_from_two_threads(gl, oracle):
    """Two host threads run pipelined host batches at once (each with its
    own streams and copy thread, sharing the staging pool): every result
    equals the device path's."""
    import torch
    cases = [(2**20, 9, 1), (3 * 2**18, 11, 2)]
    refs = {}
    for n, count, seed in cases:
        x = oracle.fill_random(n * count, 950 + seed)
        t = to_torch(x)
        gl.strategy(n).ntt(t)
        torch.cuda.synchronize()
        refs[seed] = (x, from_torch(t))
    errors = []

    def worker(n, count, seed):
        try:
            x, ref = refs[seed]
            for _ in range(3):
                y = x.copy()
                gl.strategy(n).ntt(y)
                if not np.array_equal(y, ref):
                    errors.append(f"forward n={n} count={count}")
                gl.strategy(n).ntt(y, inverse=True)
                if not np.array_equal(y, x):
                    errors.append(f"roundtrip n={n} count={count}")
        except Exception as e:  # surfaced below
            errors.append(repr(e))

    note("pipelined host batches: two threads")
    ts = [threading.Thread(target=worker, args=c) for c in cases]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
