"""Mixed-size batches: the grouped (batched) path equals each transform run
alone, bit for bit, and equals the oracle."""
import random

import numpy as np
import pytest

from gpu_util import from_torch, note, supported_by_planner, to_torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gl():
    import goldilocks_amd as g
    assert g.lib().gl_init(-1) == 0, g.lib().gl_last_error()
    return g


def mixed_sizes(count, lo, hi, seed, factors=(1, 3, 5, 15, 17, 257)):
    rng = random.Random(seed)
    out = []
    while len(out) < count:
        n = (1 << rng.randint(lo, hi)) * rng.choice(factors)
        if supported_by_planner(n):
            out.append(n)
    return out


def test_mixed_host_vs_oracle(gl, oracle):
    sizes = mixed_sizes(40, 0, 12, 1)
    note(f"mixed host {len(sizes)} transforms")
    xs = [oracle.fill_random(n, i) for i, n in enumerate(sizes)]
    ys = [x.copy() for x in xs]
    gl.ntt_batch_mixed(ys)
    for x, y, n in zip(xs, ys, sizes):
        assert np.array_equal(y, oracle.ntt(x)), n
    gl.ntt_batch_mixed(ys, inverse=True)
    for x, y in zip(xs, ys):
        assert np.array_equal(x, y)


def test_mixed_inverse_two_level_twiddles(gl, oracle):
    """Grouped device batch, forward then inverse, with every output twiddle
    from the two-level table: the inverse then folds 1/n into the first
    pass's twiddle recurrence (two-sub-step and three-stage first passes)."""
    import torch
    sizes = mixed_sizes(24, 8, 14, 5)
    note(f"mixed device inverse, two-level twiddles, {len(sizes)} transforms")
    xs = [oracle.fill_random(n, 100 + i) for i, n in enumerate(sizes)]
    ts = [to_torch(x) for x in xs]
    prev = gl.set_twiddle_full_max(0)
    try:
        gl.ntt_batch_mixed(ts)
        torch.cuda.synchronize()
        for x, t, n in zip(xs, ts, sizes):
            assert np.array_equal(from_torch(t), oracle.ntt(x)), n
        gl.ntt_batch_mixed(ts, inverse=True)
        torch.cuda.synchronize()
        for x, t, n in zip(xs, ts, sizes):
            assert np.array_equal(from_torch(t), x), n
    finally:
        gl.set_twiddle_full_max(prev)


def test_mixed_device_alone_equals_in_mix(gl, oracle):
    import torch
    sizes = mixed_sizes(64, 10, 16, 2)
    note(f"mixed device {len(sizes)} transforms, max {max(sizes)}")
    xs = [torch.empty(n, dtype=torch.int64, device="cuda") for n in sizes]
    for i, x in enumerate(xs):
        gl.fill_random(x, 100 + i)
    batched = [x.clone() for x in xs]
    gl.ntt_batch_mixed(batched)
    alone = [x.clone() for x in xs]
    gl.ntt_batch_mixed(alone, batched=False)
    torch.cuda.synchronize()
    for a, b in zip(alone, batched):
        assert torch.equal(a, b)
    for i in (0, 1, len(sizes) - 1):
        assert np.array_equal(from_torch(batched[i]), oracle.ntt(from_torch(xs[i])))


@pytest.mark.parametrize("chunk_bytes", [1 << 16, 1 << 20])
def test_chunked_batches_and_captured_plan(gl, oracle, chunk_bytes):
    """Small chunks force many grouped chunks plus stand-alone transforms
    (larger than a chunk); both the eager path and a captured batch plan
    (replayed several times, with eager work and syncs between replays)
    match the oracle."""
    import torch
    sizes = mixed_sizes(48, 4, 15, 3 + chunk_bytes)
    prev = gl.set_batch_chunk_bytes(chunk_bytes)
    try:
        note(f"chunked batch chunk={chunk_bytes} {len(sizes)} transforms max {max(sizes)}")
        hs = [oracle.fill_random(n, 500 + i) for i, n in enumerate(sizes)]
        want = [oracle.ntt(h) for h in hs]
        ts = [to_torch(h) for h in hs]
        gl.ntt_batch_mixed(ts)
        torch.cuda.synchronize()
        for t, w, n in zip(ts, want, sizes):
            assert np.array_equal(from_torch(t), w), f"eager n={n}"
        note(f"captured batch plan chunk={chunk_bytes}")
        bufs = [to_torch(h) for h in hs]
        plan = gl.BatchPlan(bufs)
        try:
            for rep in range(3):
                for b, h in zip(bufs, hs):
                    b.copy_(to_torch(h))
                y = to_torch(hs[0])
                gl.ntt(y)  # eager work between replays
                torch.cuda.synchronize()
                plan.launch()
                torch.cuda.synchronize()
                for b, w, n in zip(bufs, want, sizes):
                    assert np.array_equal(from_torch(b), w), f"replay {rep} n={n}"
        finally:
            plan.close()
    finally:
        gl.set_batch_chunk_bytes(prev)


def test_mixed_batch_refuses_graph_capture(gl):
    """gl_ntt_mixed_dev on a capturing stream returns GL_E_UNSUPPORTED (its
    launch descriptors live in a reused pinned slab; gl_batch_plan_create is
    the capturable form) and the batch still runs afterwards."""
    import torch
    note("mixed batch under graph capture")
    sizes = [1 << 10, 3 << 8, 17 << 6]
    bufs = [torch.empty(n, dtype=torch.int64, device="cuda") for n in sizes]
    for i, b in enumerate(bufs):
        gl.fill_random(b, 40 + i)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        g.capture_begin()
        try:
            with pytest.raises(Exception) as e:
                gl.ntt_batch_mixed(bufs)
        finally:
            g.capture_end()
    assert getattr(e.value, "code", None) == 5, e.value  # GL_E_UNSUPPORTED
    want = [b.clone() for b in bufs]
    for w in want:
        gl.ntt(w)
    gl.ntt_batch_mixed(bufs)
    torch.cuda.synchronize()
    for b, w in zip(bufs, want):
        assert torch.equal(b, w)
