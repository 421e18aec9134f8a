"""GPU parity at the workloads' own sizes, through the C ABI.

- 2^24, 2^27, 2^28 and the mixed batch's largest lengths (257 * 2^20,
  15 * 2^22) forward, compared with the oracle on every word
  (the contract of ntt/src/ntt/mod.rs:128-139: every output equals the exact
  DFT), then the exact roundtrip;
- every schedule of the shipped tune file (goldilocks_amd/tune_gfx950.txt),
  contiguous and 64-column, against the oracle;
- the Merkle commit of 2^20 rows x 64 columns with blowup 2: every LDE word
  and every tree node against the oracle;
- the batch of 1024 mixed transforms (n = 2^k * f, k in [10, 20],
  f in {1, 3, 5, 15, 17, 257}): each transform of the batch equals the same
  transform run alone, the oracle on every transform up to 2^18 words, the
  directly evaluated DFT at sampled outputs of larger ones, X[0] = sum x and
  the exact roundtrip for all;
- served jobs at the serving sizes (2^12..2^22 * f words, 1-64 columns,
  commits of 2^12..2^18 rows) arriving as an open-loop Poisson stream: each
  result equals the same job run alone, the oracle on a sample.
"""
import os
import random
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from gpu_util import from_torch, note, to_torch

pytestmark = pytest.mark.gpu
P = 0xFFFFFFFF00000001
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_THREADS = max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="module")
def gl():
    import goldilocks_amd as g
    rc = g.lib().gl_init(-1)
    assert rc == 0, g.lib().gl_last_error()
    return g


@pytest.fixture(autouse=True)
def oracle_threads(oracle):
    oracle.set_threads(HOST_THREADS)
    yield
    oracle.set_threads(1)


@pytest.fixture(autouse=True)
def free_device_memory():
    yield
    import torch
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def _sum_mod_p(t):
    """Exact sum of a 1-D int64 GPU tensor's u64 words, mod p."""
    import torch
    total = 0
    for c in range(0, t.numel(), 1 << 28):
        u = t[c:c + (1 << 28)]
        lo = (u & 0xFFFFFFFF).sum().item()
        hi = torch.bitwise_right_shift(u, 32).bitwise_and(0xFFFFFFFF).sum().item()
        total += hi * (1 << 32) + lo
    return total % P


FULL = [(1 << 24, "2^24"), (1 << 27, "2^27"), (1 << 28, "2^28"), (257 << 20, "257*2^20"), (15 << 22, "15*2^22")]


@pytest.mark.parametrize("n", [n for n, _ in FULL], ids=[i for _, i in FULL])
def test_full_elementwise(gl, oracle, n):
    """Every output word of the forward transform equals the oracle's at
    2^24, 2^27 and 2^28 (three schedules) and at the mixed batch's largest
    lengths, 257 * 2^20 (two 1024-point passes and the radix-257 Rader pass)
    and 15 * 2^22; the inverse restores the input exactly."""
    import torch
    note(f"full element-wise n={n}")
    x = torch.empty(n, dtype=torch.int64, device="cuda")
    gl.fill_random(x, 2000 + n.bit_length())
    host_x = from_torch(x)
    y = x.clone()
    gl.ntt(y)
    torch.cuda.synchronize()
    expect = oracle.ntt(host_x)
    got = from_torch(y)
    bad = np.flatnonzero(got != expect)
    assert bad.size == 0, f"{bad.size} of {n} words differ, first at {bad[:8]}"
    del expect, got
    gl.intt(y)
    torch.cuda.synchronize()
    assert torch.equal(y, x)


def _tune_lines():
    path = os.path.join(ROOT, "goldilocks_amd", "tune_gfx950.txt")
    rows, cols = [], []
    with open(path) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            toks = line.split()
            sched = [tuple(int(v) for v in t.replace("x", "/").split("/")) for t in toks[1:]]
            if toks[0].startswith("c"):
                cols.append((int(toks[0][1:]), sched))
            else:
                rows.append((int(toks[0]), sched))
    return rows, cols


def _expect_ids(gl, sched):
    """Variant ids of a tune-file schedule [(R1, R2, T), ...]."""
    vs = gl.variants()
    ids = []
    for r1, r2, t in sched:
        if (r1, r2) == (257, 1):
            ids.append(len(vs) - 1)
            continue
        m = [v["id"] for v in vs if v["R1"] == r1 and v["R2"] == r2 and v["T"] == t]
        assert len(m) == 1, (r1, r2, t)
        ids.append(m[0])
    return ids


ROW_LINES, COL_LINES = _tune_lines()


@pytest.mark.parametrize("n,sched", ROW_LINES, ids=[str(n) for n, _ in ROW_LINES])
def test_tuned_schedule_rows(gl, oracle, n, sched):
    """The library runs the tuned schedule for n, and it is bit-exact:
    every word against the oracle up to 2^25 words; above that the
    directly evaluated DFT at sampled outputs, X[0] = sum x and the exact
    roundtrip."""
    import torch
    assert gl.plan_ids(n) == _expect_ids(gl, sched), f"tune line for {n} not in force"
    note(f"tuned rows n={n} {sched}")
    x = torch.empty(n, dtype=torch.int64, device="cuda")
    gl.fill_random(x, n % 100003)
    y = x.clone()
    gl.ntt(y)
    torch.cuda.synchronize()
    host_x = from_torch(x)
    if n <= 1 << 25:
        assert np.array_equal(from_torch(y), oracle.ntt(host_x))
    else:
        rng = random.Random(n)
        for kk in [1, n - 1] + [rng.randrange(n) for _ in range(2)]:
            assert int(y[kk].item()) % (1 << 64) == oracle.eval_dft(host_x, kk), kk
        assert int(y[0].item()) % (1 << 64) == _sum_mod_p(x)
    gl.intt(y)
    torch.cuda.synchronize()
    assert torch.equal(y, x)


@pytest.mark.parametrize("n,sched", COL_LINES, ids=[f"c{n}" for n, _ in COL_LINES])
def test_tuned_schedule_cols(gl, oracle, n, sched):
    """The tuned 64-column schedule for n is in force and bit-exact on
    every column, forward and inverse."""
    import torch
    assert gl.plan_ids(n, cols=True) == _expect_ids(gl, sched), f"column tune line for {n} not in force"
    cols = 64
    note(f"tuned cols n={n} x {cols} {sched}")
    t = torch.empty((n, cols), dtype=torch.int64, device="cuda")
    gl.fill_random(t, 7 * n + 1)
    host = from_torch(t).reshape(n, cols)
    gl.ntt_cols(t)
    torch.cuda.synchronize()
    got = from_torch(t).reshape(n, cols)
    expect = oracle.ntt_batch(np.ascontiguousarray(host.T), n).reshape(cols, n)
    assert np.array_equal(np.ascontiguousarray(got.T), expect)
    gl.ntt_cols(t, inverse=True)
    torch.cuda.synchronize()
    assert np.array_equal(from_torch(t).reshape(n, cols), host)


def test_commit_2_20_x_64_blowup_2(gl, oracle):
    """The commit workload itself: every LDE word and every node of every
    tree layer equal the oracle's (reference primitives composed as in
    bench/src/main.rs:179-183 and pcs/src/k12/mod.rs:41-60), and every node
    of the row-tree form's 22 layers."""
    import torch
    rows, cols, blowup = 1 << 20, 64, 2
    note(f"commit {rows}x{cols} blowup {blowup}")
    m = torch.empty((rows, cols), dtype=torch.int64, device="cuda")
    gl.fill_random(m, 4242)
    lde, layers = gl.commit(m, blowup)
    torch.cuda.synchronize()
    host_m = from_torch(m).reshape(rows, cols)
    got_lde = from_torch(lde).reshape(rows * blowup, cols)
    got_layers = layers.cpu().numpy().tobytes()
    # the row-tree form of the same commit: the same extension, the binary
    # tree of row hashes over it
    note(f"commit_rows {rows}x{cols} blowup {blowup}")
    lde_r, layers_r = gl.commit_rows(m, blowup)
    assert torch.equal(lde_r, lde)
    got_rows = layers_r.cpu().numpy().tobytes()
    del lde, layers, m, lde_r, layers_r
    lde_o, layers_o = oracle.commit(host_m, blowup)
    bad = np.flatnonzero(got_lde.reshape(-1) != lde_o.reshape(-1))
    assert bad.size == 0, f"{bad.size} LDE words differ, first at {bad[:8]}"
    off = 0
    for depth, layer in enumerate(layers_o):
        seg = got_layers[off:off + len(layer)]
        assert seg == layer, f"tree layer {depth} ({len(layer)} bytes) differs"
        off += len(layer)
    assert off == len(got_layers)
    assert len(layers_o[-1]) == 32
    rows_o = oracle.row_tree(lde_o)
    assert len(rows_o) == 22 and len(rows_o[-1]) == 32
    off = 0
    for depth, layer in enumerate(rows_o):
        assert got_rows[off:off + len(layer)] == layer, f"row tree layer {depth} differs"
        off += len(layer)
    assert off == len(got_rows)


def test_batch_1024_mixed(gl, oracle):
    """1024 transforms of 2^k * f words (k in [10, 20], f in {1, 3, 5, 15,
    17, 257}; the bench's seeded list): every transform of the grouped batch
    equals the same transform alone (all words, on the device), the oracle
    on every transform of at most 2^18 words, the directly evaluated DFT at
    sampled outputs of the larger ones, X[0] = sum x, and the batched
    inverse restores every input."""
    import torch
    from goldilocks_amd.benchmarks import mixed_sizes
    sizes = mixed_sizes(1024, 7, 10, 20)
    note(f"batch 1024 mixed, {sum(sizes)} words")
    xs = [torch.empty(n, dtype=torch.int64, device="cuda") for n in sizes]
    for i, x in enumerate(xs):
        gl.fill_random(x, i)
    gl.ntt_batch_mixed(xs)  # transformed in place, grouped launches
    torch.cuda.synchronize()
    small = []
    rng = random.Random(1024)
    sampled = 0
    for i, n in enumerate(sizes):
        x = torch.empty(n, dtype=torch.int64, device="cuda")
        gl.fill_random(x, i)
        alone = x.clone()
        gl.ntt(alone)
        assert torch.equal(alone, xs[i]), f"transform {i} (n={n}) differs from the same transform alone"
        del alone
        assert int(xs[i][0].item()) % (1 << 64) == _sum_mod_p(x), i
        if n <= 1 << 18:
            small.append((i, from_torch(x), from_torch(xs[i])))
        elif sampled < 24 and n <= 1 << 22:
            host_x = from_torch(x)
            kk = rng.randrange(1, n)
            assert int(xs[i][kk].item()) % (1 << 64) == oracle.eval_dft(host_x, kk), (i, n, kk)
            sampled += 1
        del x

    def check(item):
        i, hx, hy = item
        return i, np.array_equal(oracle.ntt(hx), hy)

    oracle.set_threads(1)  # one transform per worker thread
    with ThreadPoolExecutor(HOST_THREADS) as ex:
        bad = [i for i, ok in ex.map(check, small) if not ok]
    assert not bad, f"transforms {bad[:8]} differ from the oracle"
    assert len(small) > 300 and sampled > 10
    gl.ntt_batch_mixed(xs, inverse=True)
    torch.cuda.synchronize()
    for i, n in enumerate(sizes):
        x = torch.empty(n, dtype=torch.int64, device="cuda")
        gl.fill_random(x, i)
        assert torch.equal(x, xs[i]), f"roundtrip of transform {i} (n={n})"


def test_serving_config_sizes(gl, oracle):
    """Open-loop Poisson arrivals at a fixed rate of NTT jobs (2^12..2^22 *
    f words, 1-64 transforms) and commits (2^12..2^18 rows x 1-64 columns,
    blowup 2) into the continuous-batching server with device-resident
    buffers: every job's result equals the same job run alone, and a sample
    equals the oracle."""
    import torch
    from goldilocks_amd.benchmarks import serve_jobs
    jobs = serve_jobs(160, 23, max_elems=1 << 23)
    note(f"serving {len(jobs)} jobs at config sizes")
    bufs = []
    for i, (kind, n, c) in enumerate(jobs):
        if kind == "ntt":
            x = torch.empty(n * c, dtype=torch.int64, device="cuda")
            gl.fill_random(x, 500 + i)
            bufs.append((x,))
        else:
            m = torch.empty((n, c), dtype=torch.int64, device="cuda")
            gl.fill_random(m, 500 + i)
            lde = torch.empty((2 * n, c), dtype=torch.int64, device="cuda")
            layers = torch.empty(gl.merkle_tree_size(2 * n * c * 8), dtype=torch.uint8, device="cuda")
            bufs.append((m, lde, layers))
    inputs = [b[0].clone() for b in bufs]
    torch.cuda.synchronize()
    rng = random.Random(29)
    qps = 2000.0
    arrivals, t = [], 0.0
    for _ in jobs:
        t += rng.expovariate(qps)
        arrivals.append(t)
    with gl.Server(hbm_budget_bytes=8 << 30, slots=4) as srv:
        tickets = []
        t0 = time.perf_counter()
        for (kind, n, c), b, a in zip(jobs, bufs, arrivals):
            delay = t0 + a - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            if kind == "ntt":
                tickets.append(srv.submit_ntt(b[0], n))
            else:
                tickets.append(srv.submit_commit(b[0], 2, b[1], b[2]))
        lat = [srv.wait(tk) for tk in tickets]
        st = srv.stats()
    assert st["completed"] >= len(jobs) and st["failed"] == 0 and st["rejected"] == 0
    assert all(v >= 0 for v in lat)
    checked = 0
    for i, ((kind, n, c), b, x) in enumerate(zip(jobs, bufs, inputs)):
        if kind == "ntt":
            alone = x.clone()
            gl.strategy(n).ntt(alone)
            assert torch.equal(alone, b[0]), f"job {i}: ntt n={n} x {c} differs in the mix"
            if i % 4 == 0 and n * c <= 1 << 21:
                assert np.array_equal(from_torch(b[0]), oracle.ntt_batch(from_torch(x), n)), i
                checked += 1
        else:
            lde, layers = gl.commit(x, 2)
            assert torch.equal(lde, b[1]) and torch.equal(layers, b[2]), f"job {i}: commit {n}x{c} differs in the mix"
            if n * c <= 1 << 18:
                lde_o, layers_o = oracle.commit(from_torch(x).reshape(n, c), 2)
                assert np.array_equal(from_torch(b[1]).reshape(2 * n, c), lde_o), i
                assert b[2].cpu().numpy().tobytes() == b"".join(layers_o), i
                checked += 1
    assert checked >= 12
