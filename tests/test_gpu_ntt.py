"""GPU NTT parity against the CPU oracle (bit-exact), through the C ABI.

Small and medium lengths are compared element by element with the oracle's
restatement of the reference plans; large lengths use size-independent
checks (sampled X[k] evaluated directly, exact roundtrip, linearity)."""
import random

import numpy as np
import pytest

from gpu_util import from_torch, note, supported_by_planner, to_torch

pytestmark = pytest.mark.gpu
P = 0xFFFFFFFF00000001


@pytest.fixture(scope="module")
def gl():
    import goldilocks_amd as g
    rc = g.lib().gl_init(-1)
    assert rc == 0, g.lib().gl_last_error()
    return g


def small_sizes(oracle):
    return [n for n in oracle.divisors() if n <= 4096 and supported_by_planner(n)]


def test_every_small_size_forward_and_inverse(gl, oracle):
    for n in small_sizes(oracle):
        note(f"ntt/intt n={n}")
        x = oracle.fill_random(n, 1000 + n)
        y = x.copy()
        gl.ntt(y)
        assert np.array_equal(y, oracle.ntt(x)), f"forward n={n}"
        z = x.copy()
        gl.intt(z)
        assert np.array_equal(z, oracle.intt(x)), f"inverse n={n}"


MEDIUM = [2**13, 2**15, 2**16, 2**18, 3 * 2**14, 5 * 2**13, 15 * 2**12, 17 * 2**12, 51 * 2**10,
          85 * 2**9, 255 * 2**8, 17 * 2**16, 3 * 2**17, 257 * 2**10, 257 * 3 * 2**8,
          257 * 17 * 2**4, 257 * 2**14, 65537, 65537 * 4, 65537 * 3 * 8]


@pytest.mark.parametrize("n", MEDIUM)
def test_medium_vs_oracle(gl, oracle, n):
    note(f"medium n={n}")
    x = oracle.fill_random(n, n)
    y = x.copy()
    gl.ntt(y)
    assert np.array_equal(y, oracle.ntt(x))
    gl.intt(y)
    assert np.array_equal(y, x)


def test_2_16_forward_inverse_config(gl, oracle):
    """Plain single forward+inverse NTT, length 2^16 (bit-exact)."""
    n = 2**16
    note("config 2^16 fwd+inv")
    x = oracle.fill_random(n, 2024)
    y = x.copy()
    gl.ntt(y)
    assert np.array_equal(y, oracle.ntt(x))
    gl.intt(y)
    assert np.array_equal(y, x)


@pytest.mark.parametrize("n,count", [(64, 37), (1024, 9), (4096, 5), (15 * 64, 11), (2**16, 3), (17, 100),
                                     (257, 33), (257 * 64, 5), (65537, 3)])
def test_batch(gl, oracle, n, count):
    note(f"batch n={n} count={count}")
    x = oracle.fill_random(n * count, 7)
    y = x.copy()
    gl.strategy(n).ntt(y)
    assert np.array_equal(y, oracle.ntt_batch(x, n))
    gl.strategy(n).ntt(y, inverse=True)
    assert np.array_equal(y, x)


def test_edge_inputs(gl, oracle):
    note("edge inputs")
    for n in (2, 64, 4096, 2**16):
        for fill in (0, 1, P - 1):
            x = np.full(n, fill, dtype=np.uint64)
            y = x.copy()
            gl.ntt(y)
            assert np.array_equal(y, oracle.ntt(x))
        e0 = np.zeros(n, dtype=np.uint64)
        e0[0] = 1
        gl.ntt(e0)
        assert np.all(e0 == 1)
    one = np.array([12345], dtype=np.uint64)
    gl.ntt(one)
    assert one[0] == 12345
    gl.intt(np.zeros(0, dtype=np.uint64))  # intt accepts n <= 1 (mod.rs:105-107)


def test_errors(gl):
    note("error paths")
    with pytest.raises(ValueError):
        gl.ntt(np.zeros(7, dtype=np.uint64))
    with pytest.raises(ValueError):
        gl.ntt(np.zeros(0, dtype=np.uint64))
    with pytest.raises(ValueError):
        gl.strategy(3 * 7)
    # n * count words beyond the address space: rejected before any launch
    # (the pointer is never dereferenced)
    import ctypes
    import torch
    x = torch.empty(4096, dtype=torch.int64, device="cuda")
    rc = gl.lib().gl_ntt_dev(ctypes.c_void_p(x.data_ptr()), 4096, 1 << 62, 0, None)
    assert rc == gl.GL_E_ARG and b"overflow" in gl.lib().gl_last_error()


def test_device_columns(gl, oracle):
    """Column transforms of a row-major matrix (the commit's layout)."""
    import torch
    for rows, cols in ((1024, 64), (4096, 3), (2**14, 8), (240, 5), (2**18, 16), (257 * 256, 4), (3 * 2**15, 7)):
        note(f"columns {rows}x{cols}")
        m = oracle.fill_random(rows * cols, rows + cols).reshape(rows, cols)
        t = to_torch(m)
        gl.ntt_cols(t)
        torch.cuda.synchronize()
        got = from_torch(t).reshape(rows, cols)
        for c in range(cols):
            assert np.array_equal(got[:, c], oracle.ntt(m[:, c].copy())), (rows, cols, c)
        gl.ntt_cols(t, inverse=True)
        torch.cuda.synchronize()
        assert np.array_equal(from_torch(t).reshape(rows, cols), m)


@pytest.mark.parametrize("k", [20, 24, 27])
def test_large_device(gl, oracle, k):
    """2^20 / 2^24 / 2^27 forward then inverse on HBM-resident data:
    sampled outputs equal the directly evaluated DFT, the roundtrip is exact,
    and repeated runs are bitwise identical."""
    import torch
    n = 1 << k
    note(f"large n=2^{k}")
    x = torch.empty(n, dtype=torch.int64, device="cuda")
    gl.fill_random(x, 99)
    torch.cuda.synchronize()
    host_x = from_torch(x)
    assert np.array_equal(host_x[:1000], oracle.fill_random(1000, 99))
    y = x.clone()
    gl.ntt(y)
    torch.cuda.synchronize()
    host_y = from_torch(y)
    rng = random.Random(k)
    ks = [0, 1, n - 1, n // 2] + [rng.randrange(n) for _ in range(2 if k >= 27 else 4)]
    for kk in ks:
        assert int(host_y[kk]) == oracle.eval_dft(host_x, kk), (k, kk)
    y2 = x.clone()
    gl.ntt(y2)
    assert torch.equal(y, y2)
    gl.intt(y)
    torch.cuda.synchronize()
    assert torch.equal(y, x)
    if k <= 20:
        assert np.array_equal(host_y, oracle.ntt(host_x))


def _sum_mod_p(t, chunk=1 << 28):
    """Exact sum of u64 words (a 1-D int64 GPU tensor, any length) mod p:
    32-bit halves of at most 2^28 words are summed in int64 (no overflow),
    chunk sums are added as Python integers."""
    import torch
    total = 0
    for c in range(0, t.numel(), chunk):
        u = t[c:c + chunk]
        lo = (u & 0xFFFFFFFF).sum().item()
        hi = torch.bitwise_right_shift(u, 32).bitwise_and(0xFFFFFFFF).sum().item()
        total += hi * (1 << 32) + lo
    return total % P


@pytest.mark.parametrize("k", [30, 31, 32])
def test_max_size_device(gl, k):
    """Lengths where word positions pass 2^30 / 2^31 and byte offsets pass
    2^33 / 2^34 (8 and 16 GiB), up to 2^32, the largest power of two that
    divides p - 1 (32 GiB, four passes): X[0] = sum x_j and
    X[n/2] = sum (-1)^j x_j exactly, a run is deterministic, and the
    roundtrip is exact."""
    import torch
    n = 1 << k
    free, _ = torch.cuda.mem_get_info()
    if free < 4 * 8 * n + (4 << 30):
        pytest.skip(f"needs {4 * 8 * n >> 30} GiB of device memory")
    note(f"max size n=2^{k}")
    x = torch.empty(n, dtype=torch.int64, device="cuda")
    gl.fill_random(x, 7)
    y = x.clone()
    gl.ntt(y)
    torch.cuda.synchronize()
    assert int(y[0].item()) % (1 << 64) == _sum_mod_p(x)
    even, odd = _sum_mod_p(x[0::2]), _sum_mod_p(x[1::2])
    assert int(y[n // 2].item()) % (1 << 64) == (even - odd) % P
    z = x.clone()
    gl.ntt(z)
    assert torch.equal(y, z)
    del z
    gl.intt(y)
    torch.cuda.synchronize()
    assert torch.equal(y, x)


def test_pass_timing_records_every_pass(gl):
    """gl_timing_enable/collect: one record per pass launch, in launch order,
    named like the schedule, none while disabled."""
    import torch
    n = 1 << 16
    x = torch.empty(n, dtype=torch.int64, device="cuda")
    gl.fill_random(x, 3)
    plan = [p["kernel"] for p in gl.plan_info(n)]
    note(f"pass timing n={n} plan={plan}")
    gl.timing_collect()
    gl.ntt(x)
    torch.cuda.synchronize()
    assert gl.timing_collect() == []
    gl.timing_enable(True)
    try:
        for _ in range(3):
            gl.ntt(x)
        torch.cuda.synchronize()
    finally:
        gl.timing_enable(False)
    recs = gl.timing_collect()
    assert [name for name, _ in recs] == plan * 3
    assert all(ms > 0 for _, ms in recs)
    assert gl.timing_collect() == []


def test_column_schedule_override(gl, oracle):
    """gl_plan_set_cols installs a column-transform schedule of its own (here
    a single-sub-step first pass, as the tuned commit schedules use); the
    columns stay bit-exact and the contiguous schedule is untouched."""
    import torch
    n, cols = 4096, 6
    vs = {v["name"]: v["id"] for v in gl.variants()}
    ids = [vs["ntt_pass_kernel<16, 1, 256, 256>"], vs["ntt_pass_kernel<16, 16, 256, 16>"]]
    before = gl.plan_ids(n)
    gl.plan_set(n, ids, cols=True)
    try:
        assert gl.plan_ids(n, cols=True) == ids
        assert gl.plan_ids(n) == before
        m = oracle.fill_random(n * cols, 17).reshape(n, cols)
        t = to_torch(m.reshape(-1)).reshape(n, cols)
        note(f"column schedule override n={n} cols={cols}")
        gl.ntt_cols(t)
        torch.cuda.synchronize()
        got = from_torch(t.reshape(-1)).reshape(n, cols)
        for c in range(cols):
            assert np.array_equal(got[:, c], oracle.ntt(np.ascontiguousarray(m[:, c])))
    finally:
        gl.plan_set(n, [], cols=True)
