"""Serving (continuous batching) and captured batch plans: every result in
the mix equals the same job run alone, and the oracle."""
import random
import time

import numpy as np
import pytest

from gpu_util import from_torch, note, to_torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gl():
    import goldilocks_amd as g
    assert g.lib().gl_init(-1) == 0, g.lib().gl_last_error()
    return g


def test_batch_plan_graph_replays(gl, oracle):
    import torch
    sizes = [1 << 12, 3 << 10, 1 << 16, 17 << 8, 257 << 4, 15 << 6, 1 << 20]
    note(f"batch plan {sizes}")
    xs = [torch.empty(n, dtype=torch.int64, device="cuda") for n in sizes]
    for i, x in enumerate(xs):
        gl.fill_random(x, 50 + i)
    bufs = [x.clone() for x in xs]
    plan = gl.BatchPlan(bufs)
    expect = [x.clone() for x in xs]
    for rep in range(3):  # each replay applies one more forward transform
        gl.ntt_batch_mixed(expect)
        plan.launch()
        torch.cuda.synchronize()
        for e, b in zip(expect, bufs):
            assert torch.equal(e, b), rep
    plan.close()
    assert np.array_equal(from_torch(xs[1]), oracle.intt(oracle.intt(oracle.intt(from_torch(bufs[1]).copy()))))


def test_server_mix_equals_alone_host(gl, oracle):
    rng = random.Random(3)
    jobs = []
    for i in range(40):
        n = (1 << rng.randint(4, 13)) * rng.choice([1, 3, 5, 15, 17, 257])
        count = rng.choice([1, 1, 2, 5])
        jobs.append((n, count, rng.random() < 0.3))
    note(f"server host {len(jobs)} jobs")
    datas = [oracle.fill_random(n * c, 900 + i) for i, (n, c, inv) in enumerate(jobs)]
    work = [d.copy() for d in datas]
    with gl.Server(hbm_budget_bytes=64 << 20, slots=4) as srv:
        tickets = []
        for (n, c, inv), w in zip(jobs, work):
            tickets.append(srv.submit_ntt(w, n=n, inverse=inv))
            time.sleep(rng.random() * 0.002)  # arrivals while others run
        lat = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert st["completed"] == len(jobs) and st["failed"] == 0
    assert all(l >= 0 for l in lat)
    for (n, c, inv), d, w in zip(jobs, datas, work):
        alone = d.copy()
        gl.strategy(n).ntt(alone, inverse=inv)  # the same job alone
        assert np.array_equal(alone, w), (n, c, inv)
        assert np.array_equal(w, oracle.ntt_batch(d, n, inverse=inv)), (n, c, inv)


def test_server_device_jobs_and_commit(gl, oracle):
    import torch
    note("server device jobs + commit")
    sizes = [1 << 20, 1 << 14, 5 << 12, 1 << 18, 257 << 8]
    xs = [torch.empty(n, dtype=torch.int64, device="cuda") for n in sizes]
    for i, x in enumerate(xs):
        gl.fill_random(x, 300 + i)
    ys = [x.clone() for x in xs]
    rows, cols = 1024, 8
    m = oracle.fill_random(rows * cols, 77).reshape(rows, cols)
    lde = np.zeros((2 * rows, cols), dtype=np.uint64)
    layers = np.zeros(gl.merkle_tree_size(lde.size * 8), dtype=np.uint8)
    with gl.Server(slots=3) as srv:
        tickets = [srv.submit_ntt(y) for y in ys]
        tc = srv.submit_commit(m, 2, lde, layers)
        for t in tickets + [tc]:
            srv.wait(t)
        st = srv.stats()
    assert st["completed"] == len(sizes) + 1
    for x, y in zip(xs, ys):
        alone = x.clone()
        gl.ntt(alone)
        assert torch.equal(alone, y)
    lde_o, layers_o = oracle.commit(m, 2)
    assert np.array_equal(lde, lde_o)
    assert layers.tobytes() == b"".join(layers_o)


def test_server_budget_rejects_and_admits(gl):
    note("server budget")
    with gl.Server(hbm_budget_bytes=1 << 20, slots=2) as srv:
        big = np.zeros(1 << 18, dtype=np.uint64)  # 2 MiB + scratch > budget
        with pytest.raises(gl.GoldilocksError):
            srv.submit_ntt(big)
        small = [np.ones(1 << 12, dtype=np.uint64) for _ in range(8)]
        ts = [srv.submit_ntt(s) for s in small]
        for t in ts:
            srv.wait(t)
        st = srv.stats()
        assert st["rejected"] == 1 and st["completed"] == 8
        assert st["peak_bytes"] <= 1 << 20
    for s in small:
        assert s[0] == (1 << 12) and np.all(s[1:] == 0)
