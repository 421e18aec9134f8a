"""Thread safety of the C ABI (every entry point may be called from several
host threads, as the reference's `Ntt: Sync + Send` plans and its global
plan cache allow, ntt/src/ntt/mod.rs:23-78).

- two threads transforming different buffers on the same (NULL) stream get
  the same words as one thread alone (each call owns its scratch);
- NTTs, Merkle trees, row-tree commits and clear_cache issued concurrently
  from several threads stay bit-exact (tables a call is using are never
  freed under it);
- a captured batch plan keeps its twiddle tables across clear_cache;
- grouped mixed batches from more threads than the pool of pinned
  descriptor slabs holds stay bit-exact;
- removing an explicit schedule, or clearing the caches, puts the tune
  file's schedule back in force.
"""
import ctypes
import threading

import numpy as np
import pytest

from gpu_util import from_torch, note

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gl():
    import goldilocks_amd as g
    assert g.lib().gl_init(-1) == 0, g.lib().gl_last_error()
    return g


def _run_threads(fns):
    errors = []

    def wrap(fn):
        try:
            fn()
        except Exception as e:  # reported after join
            errors.append(e)

    ts = [threading.Thread(target=wrap, args=(f,)) for f in fns]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=300)
    assert not any(t.is_alive() for t in ts), "a thread did not finish"
    if errors:
        raise errors[0]


def test_two_threads_null_stream(gl, oracle):
    """gl_ntt_dev from two threads on the NULL stream, 2^20 words each,
    forward and inverse in turn: every result equals the oracle's."""
    import torch
    n = 1 << 20
    note("two threads on the NULL stream, 2^20")
    xs = [torch.empty(n, dtype=torch.int64, device="cuda") for _ in range(2)]
    for i, x in enumerate(xs):
        gl.fill_random(x, 70 + i)
    torch.cuda.synchronize()
    hosts = [from_torch(x) for x in xs]
    fwd = [oracle.ntt(h) for h in hosts]
    lib = gl.lib()
    rounds = 41  # odd: ends after a forward transform

    def work(i):
        def run():
            p = ctypes.c_void_p(xs[i].data_ptr())
            for r in range(rounds):
                rc = lib.gl_ntt_dev(p, n, 1, r % 2, None)
                assert rc == 0, lib.gl_last_error()
        return run

    _run_threads([work(0), work(1)])
    torch.cuda.synchronize()
    for i in range(2):
        assert np.array_equal(from_torch(xs[i]), fwd[i]), f"thread {i}"


def test_concurrent_ntt_merkle_clear_cache(gl, oracle):
    """Threads running NTTs (own streams and the NULL stream), Merkle trees,
    row-tree commits and clear_cache at the same time: every result is
    bit-exact."""
    import torch
    note("concurrent ntt / merkle / clear_cache")
    sizes = [1 << 16, 3 << 14, 257 << 6, 1 << 18]
    xs = [torch.empty(n, dtype=torch.int64, device="cuda") for n in sizes]
    for i, x in enumerate(xs):
        gl.fill_random(x, 300 + i)
    data = oracle.fill_random(1 << 15, 5).tobytes()
    torch.cuda.synchronize()
    hosts = [from_torch(x) for x in xs]
    expect = [oracle.ntt(oracle.ntt(h)) for h in hosts]  # two forward transforms
    tree_expect = oracle.merkle_tree(data)
    lib = gl.lib()
    results = {}

    def ntt_worker(i, own_stream):
        def run():
            s = torch.cuda.Stream() if own_stream else None
            sp = ctypes.c_void_p(s.cuda_stream) if s is not None else None
            y = xs[i]
            p = ctypes.c_void_p(y.data_ptr())
            for _ in range(2):
                assert lib.gl_ntt_dev(p, sizes[i], 1, 0, sp) == 0, lib.gl_last_error()
            if s is not None:
                s.synchronize()
        return run

    def merkle_worker():
        for r in range(6):
            results[r] = gl.merkle_tree(data)

    def clear_worker():
        for _ in range(8):
            gl.clear_cache()

    def commit_rows_worker():
        for r in range(3):
            row_results[r] = gl.commit_rows(mat, 2)

    row_results = {}
    mat = oracle.fill_random(256 * 12, 9).reshape(256, 12)
    rows_expect = oracle.commit_rows(mat, 2)
    _run_threads([ntt_worker(0, True), ntt_worker(1, False), ntt_worker(2, True), ntt_worker(3, False),
                  merkle_worker, clear_worker, commit_rows_worker])
    torch.cuda.synchronize()
    for i in range(len(sizes)):
        assert np.array_equal(from_torch(xs[i]), expect[i]), f"n={sizes[i]}"
    for r, layers in results.items():
        assert layers == tree_expect, r
    for r, (lde, layers) in row_results.items():
        assert np.array_equal(lde, rows_expect[0]) and layers == rows_expect[1], r


def test_batch_plan_survives_clear_cache(gl):
    """A captured batch plan launched after clear_cache still reads valid
    twiddle tables: its results equal the eager batch."""
    import torch
    sizes = [1 << 12, 3 << 10, 1 << 16, 17 << 8, 257 << 4, 1 << 20]
    note(f"batch plan across clear_cache {sizes}")
    xs = [torch.empty(n, dtype=torch.int64, device="cuda") for n in sizes]
    for i, x in enumerate(xs):
        gl.fill_random(x, 600 + i)
    bufs = [x.clone() for x in xs]
    plan = gl.BatchPlan(bufs)
    gl.clear_cache()
    torch.cuda.synchronize()
    # tables the eager path builds now are new allocations; reuse of the
    # freed memory must not disturb the plan's own tables
    junk = torch.full((1 << 22,), -1, dtype=torch.int64, device="cuda")
    plan.launch()
    gl.ntt_batch_mixed(xs)
    torch.cuda.synchronize()
    for x, b in zip(xs, bufs):
        assert torch.equal(x, b)
    del junk
    plan.close()


def test_tuned_schedule_restored(gl):
    """plan_set(n, []) after an explicit schedule, and clear_cache, leave the
    tune file's schedule in force (lengths with a tune-file line)."""
    n, ncols = 1 << 20, 1 << 21  # tune lines "1048576 ..." and "c2097152 ..."
    tuned = gl.plan_ids(n)
    tuned_c = gl.plan_ids(ncols, cols=True)
    vs = {v["name"]: v["id"] for v in gl.variants()}
    alt = [vs["ntt_pass_kernel<16, 16, 256, 16>"], vs["ntt_pass_kernel<16, 16, 256, 16>"],
           vs["ntt_pass_kernel<16, 1, 256, 256>"]]
    gl.plan_set(n, alt)
    assert gl.plan_ids(n) == alt
    gl.plan_set(n, [])
    assert gl.plan_ids(n) == tuned
    gl.clear_cache()
    assert gl.plan_ids(n) == tuned
    assert gl.plan_ids(ncols, cols=True) == tuned_c


def test_mixed_batches_from_twelve_threads(gl):
    """Twelve threads run grouped mixed batches at once, more than the pool
    of pinned descriptor slabs holds: every batch equals its transforms run
    one by one."""
    import random
    import torch
    sizes_per_thread = []
    for t in range(12):
        rng = random.Random(700 + t)
        sizes_per_thread.append([(1 << rng.randint(4, 14)) * rng.choice((1, 3, 5, 17)) for _ in range(24)])
    refs = []
    for t, sizes in enumerate(sizes_per_thread):
        xs = [torch.empty(n, dtype=torch.int64, device="cuda") for n in sizes]
        for i, x in enumerate(xs):
            gl.fill_random(x, 1000 * t + i)
        ys = [x.clone() for x in xs]
        for y in ys:
            gl.ntt(y)
        refs.append((xs, ys))
    torch.cuda.synchronize()
    note("grouped mixed batches from twelve threads")

    def work(t):
        def run():
            xs, ys = refs[t]
            for _ in range(4):
                bufs = [x.clone() for x in xs]
                gl.ntt_batch_mixed(bufs)
                torch.cuda.current_stream().synchronize()
                for b, y in zip(bufs, ys):
                    assert torch.equal(b, y), (t, b.numel())
        return run

    _run_threads([work(t) for t in range(12)])
