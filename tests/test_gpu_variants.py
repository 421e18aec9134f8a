"""Every compiled pass kernel, in every position a schedule can give it,
bit-exact against the oracle.

A variant runs alone (one pass, no output twiddles), first (stride 1, output
twiddles w_M^(m*k)), last (interleaved sub-problems, S > 1) and in the
middle of a schedule, forward and inverse, with the output twiddles taken
from the full table and from the two-level power table (gl_set_twiddle_full_max).
The same schedules then run as one grouped batch (gl_ntt_mixed_dev)."""
import numpy as np
import pytest

from gpu_util import from_torch, note, to_torch

pytestmark = pytest.mark.gpu
MAX_N = 1 << 18


@pytest.fixture(scope="module")
def gl():
    import goldilocks_amd as g
    rc = g.lib().gl_init(-1)
    assert rc == 0, g.lib().gl_last_error()
    return g


def schedules(vs):
    one = {v["R"]: v["id"] for v in vs if v["R2"] == 1 and v["name"].startswith("ntt_pass_kernel")}
    out = []
    for v in vs:
        i = v["id"]
        for ids in ([i], [i, one[16]], [one[16], i], [one[4], i, one[4]]):
            n = 1
            for j in ids:
                n *= vs[j]["R"]
            if n <= MAX_N:
                out.append((v["name"], ids, n))
    return out


@pytest.mark.parametrize("full_max", [1 << 20, 0], ids=["full-table", "two-level"])
def test_every_variant_every_position(gl, oracle, full_max):
    vs = gl.variants()
    prev = gl.set_twiddle_full_max(full_max)
    cache = {}
    touched = set()
    scheds = schedules(vs)
    assert len(scheds) >= 4 * 40, "every variant in several positions"
    try:
        for name, ids, n in scheds:
            note(f"variant {name} plan {ids} n={n} full_max={full_max}")
            gl.plan_set(n, ids)
            touched.add(n)
            if n not in cache:
                x = oracle.fill_random(n, 7 * n + 1)
                cache[n] = (x, oracle.ntt(x))
            x, want = cache[n]
            y = x.copy()
            gl.ntt(y)
            assert np.array_equal(y, want), f"forward {name} {ids} n={n}"
            gl.intt(y)
            assert np.array_equal(y, x), f"inverse {name} {ids} n={n}"
    finally:
        for n in touched:
            gl.plan_set(n, [])
        gl.set_twiddle_full_max(prev)


@pytest.mark.parametrize("full_max", [0, 1 << 10, 1 << 20], ids=["two-level", "shared-full", "full"])
def test_column_twiddle_tables(gl, oracle, full_max):
    """Column transforms of 2^12 rows: with threshold 2^10 the 64-column passes
    (instance stride >= 32) take the full table though M > threshold, the
    16-column ones the two-level table; every column stays bit-exact."""
    import torch
    n = 4096
    prev = gl.set_twiddle_full_max(full_max)
    try:
        for cols in (64, 16):
            note(f"column twiddle tables n={n} cols={cols} full_max={full_max}")
            m = oracle.fill_random(n * cols, 31 + cols).reshape(n, cols)
            t = to_torch(m.reshape(-1)).reshape(n, cols)
            gl.ntt_cols(t)
            torch.cuda.synchronize()
            got = from_torch(t.reshape(-1)).reshape(n, cols)
            for c in range(cols):
                assert np.array_equal(got[:, c], oracle.ntt(np.ascontiguousarray(m[:, c]))), (cols, c)
            gl.ntt_cols(t, inverse=True)
            torch.cuda.synchronize()
            assert np.array_equal(from_torch(t.reshape(-1)).reshape(n, cols), m)
    finally:
        gl.set_twiddle_full_max(prev)


def test_every_variant_grouped_batch(gl, oracle):
    """The same schedules as one grouped launch sequence over device arrays."""
    vs = gl.variants()
    scheds = schedules(vs)
    chunk = 24
    for c0 in range(0, len(scheds), chunk):
        part, seen = [], set()
        for name, ids, n in scheds[c0:c0 + chunk]:
            if n not in seen:  # one schedule per length at a time
                seen.add(n)
                part.append((name, ids, n))
        note(f"grouped batch of {len(part)} schedules starting at {part[0][0]}")
        try:
            xs = []
            for name, ids, n in part:
                gl.plan_set(n, ids)
                xs.append(oracle.fill_random(n, 3 * n + 5))
            ts = [to_torch(x) for x in xs]
            gl.ntt_batch_mixed(ts, batched=True)
            for (name, ids, n), x, t in zip(part, xs, ts):
                assert np.array_equal(from_torch(t), oracle.ntt(x)), f"grouped {name} {ids} n={n}"
        finally:
            for name, ids, n in part:
                gl.plan_set(n, [])
