"""Seeded random sweep over the planner's combinations, every output word
against the oracle: uniform batches of random divisors of p - 1 in both
directions, column transforms with random column counts, commits with
random shapes and blowups (not only 2), and mixed device batches.  Each
case is drawn from a fixed seed, so a failure names a reproducible case."""
import random

import numpy as np
import pytest

from gpu_util import from_torch, note, to_torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gl():
    import goldilocks_amd as g
    assert g.lib().gl_init(-1) == 0, g.lib().gl_last_error()
    return g


@pytest.fixture(scope="module")
def divs(oracle):
    return oracle.divisors()


def _draw(rng, divs, cap):
    return rng.choice([d for d in divs if d <= cap])


def test_random_uniform_batches(gl, oracle, divs):
    """strategy(n).ntt over count consecutive transforms, forward or inverse,
    on device tensors, against the oracle's batch transform."""
    import torch
    rng = random.Random(20241204)
    for case in range(200):
        n = _draw(rng, divs, 1 << 18)
        count = rng.choice([1, 1, 2, 3, 5])
        while n * count > 1 << 20:
            count -= 1
        inverse = rng.random() < 0.5
        note(f"uniform case {case}: n={n} count={count} inverse={inverse}")
        x = oracle.fill_random(n * count, 1000 + case)
        t = to_torch(x)
        gl.strategy(n).ntt(t, inverse=inverse)
        torch.cuda.synchronize()
        assert np.array_equal(from_torch(t), oracle.ntt_batch(x, n, inverse)), (case, n, count, inverse)


def test_random_column_transforms(gl, oracle, divs):
    """Column transforms of row-major matrices with random row lengths and
    column counts (1-70), forward against the oracle column by column, then
    the inverse back to the input."""
    import torch
    rng = random.Random(4242)
    for case in range(80):
        rows = _draw(rng, divs, 1 << 15)
        cols = rng.randint(1, 70)
        while rows * cols > 1 << 20:
            cols //= 2
        cols = max(cols, 1)
        note(f"columns case {case}: {rows}x{cols}")
        m = oracle.fill_random(rows * cols, 2000 + case).reshape(rows, cols)
        t = to_torch(m)
        gl.ntt_cols(t)
        torch.cuda.synchronize()
        want = oracle.ntt_batch(np.ascontiguousarray(m.T).reshape(-1), rows).reshape(cols, rows).T
        assert np.array_equal(from_torch(t).reshape(rows, cols), want), (case, rows, cols)
        gl.ntt_cols(t, inverse=True)
        torch.cuda.synchronize()
        assert np.array_equal(from_torch(t).reshape(rows, cols), m), (case, rows, cols)


def test_random_commits(gl, oracle, divs):
    """Commits of random shapes with blowups 2, 3, 4, 5 and 8 (any blowup
    for which blowup * rows divides p - 1): the LDE and every tree layer
    against the oracle; every other case also as the row-tree commit from a
    device tensor."""
    rng = random.Random(777)
    done = 0
    case = 0
    while done < 48:
        case += 1
        rows = _draw(rng, divs, 1 << 12)
        blowup = rng.choice([2, 2, 3, 4, 5, 8])
        if (oracle.P - 1) % (rows * blowup):
            continue
        cols = rng.randint(1, 40)
        note(f"commit case {case}: {rows}x{cols} blowup {blowup}")
        m = oracle.fill_random(rows * cols, 3000 + case).reshape(rows, cols)
        lde, layers = gl.commit(m, blowup)
        lde_o, layers_o = oracle.commit(m, blowup)
        assert np.array_equal(lde, lde_o), (case, rows, cols, blowup)
        assert layers == layers_o, (case, rows, cols, blowup)
        if done % 2:
            note(f"row-tree commit case {case}: {rows}x{cols} blowup {blowup}")
            rl_o, rt_o = oracle.commit_rows(m, blowup)
            assert np.array_equal(rl_o, lde_o)
            dl, dt = gl.commit_rows(to_torch(m), blowup)
            assert np.array_equal(from_torch(dl).reshape(lde_o.shape), lde_o), (case, rows, cols, blowup)
            assert dt.cpu().numpy().tobytes() == b"".join(rt_o), (case, rows, cols, blowup)
        done += 1


def test_random_mixed_batches(gl, oracle, divs):
    """Grouped device batches of 3-25 transforms of random lengths, forward
    or inverse, each transform against the oracle."""
    import torch
    rng = random.Random(99)
    for case in range(24):
        k = rng.randint(3, 25)
        ns = [_draw(rng, divs, 1 << 16) for _ in range(k)]
        inverse = rng.random() < 0.5
        note(f"mixed case {case}: {k} transforms inverse={inverse}")
        hs = [oracle.fill_random(n, 4000 + 37 * case + i) for i, n in enumerate(ns)]
        ts = [to_torch(h) for h in hs]
        gl.ntt_batch_mixed(ts, inverse=inverse)
        torch.cuda.synchronize()
        for i, (n, h, t) in enumerate(zip(ns, hs, ts)):
            want = oracle.intt(h) if inverse else oracle.ntt(h)
            assert np.array_equal(from_torch(t), want), (case, i, n, inverse)


def test_random_transposes(gl, oracle):
    """Device transposes of random shapes (1-3000 on a side), in place and
    out of place, against numpy's transpose; every fourth shape square."""
    import torch
    rng = random.Random(31337)
    for case in range(40):
        rows = rng.randint(1, 3000)
        cols = rows if case % 4 == 0 else rng.randint(1, 3000)
        note(f"transpose case {case}: {rows}x{cols}")
        m = oracle.fill_random(rows * cols, 5000 + case)
        want = np.ascontiguousarray(m.reshape(rows, cols).T).reshape(-1)
        src = to_torch(m)
        dst = torch.empty_like(src)
        gl.transpose(src, rows, cols, out=dst)
        gl.transpose(src, rows, cols)
        torch.cuda.synchronize()
        assert np.array_equal(from_torch(dst), want), (case, rows, cols)
        assert np.array_equal(from_torch(src), want), (case, rows, cols)
