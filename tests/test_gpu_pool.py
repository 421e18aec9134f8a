"""Scratch reuse across streams (the runtime's stream-ordered pool).

Blocks freed on the NULL stream carry no event until another stream asks
for them; blocks freed on other streams carry one from the free.  Either
way a block is never handed to a second stream while work on the first may
still use it.  The transforms below alternate between the NULL stream and
two side streams with no synchronization in between, with equal lengths (so
every call asks the pool for the same scratch size), and every result must
equal the oracle's.  Host-slice calls return their blocks idle after their
synchronous copy back; a thread making them runs beside a thread that keeps
a side stream busy with the same length.
"""
import pytest

from gpu_util import from_torch, note

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gl():
    import goldilocks_amd as g
    assert g.lib().gl_init(-1) == 0, g.lib().gl_last_error()
    return g


@pytest.mark.parametrize("log_n", [16, 20, 22])
def test_null_and_side_streams_share_scratch_sizes(gl, oracle, log_n):
    import torch
    n = 1 << log_n
    note(f"pool reuse across streams, 2^{log_n}")
    side = [torch.cuda.Stream(), torch.cuda.Stream()]
    xs = [torch.empty(n, dtype=torch.int64, device="cuda") for _ in range(3)]
    for i, x in enumerate(xs):
        gl.fill_random(x, 700 + 10 * log_n + i)
    torch.cuda.synchronize()
    expect = [oracle.ntt(from_torch(x)) for x in xs]
    for rnd in range(6):
        ys = [x.clone() for x in xs]
        torch.cuda.synchronize()
        # NULL stream, then each side stream, then the NULL stream again,
        # all enqueued before any of them has finished
        gl.ntt(ys[0])
        for k, s in enumerate(side):
            with torch.cuda.stream(s):
                gl.ntt(ys[k + 1])
                gl.intt(ys[k + 1])
                gl.ntt(ys[k + 1])
        gl.intt(ys[0])
        gl.ntt(ys[0])
        for s in side:
            torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        for k in range(3):
            got = from_torch(ys[k])
            assert (got == expect[k]).all(), f"round {rnd}, buffer {k}: result differs from the oracle"


def test_null_stream_blocks_after_side_stream_use(gl, oracle):
    """A side stream takes a scratch size first; then the NULL stream runs
    the same size back to back many times, then the side stream again: all
    bit-exact."""
    import torch
    n = 1 << 20
    note("pool: side stream, NULL stream x 50, side stream, 2^20")
    s = torch.cuda.Stream()
    x = torch.empty(n, dtype=torch.int64, device="cuda")
    gl.fill_random(x, 4242)
    torch.cuda.synchronize()
    expect = oracle.ntt(from_torch(x))
    a, b, c = x.clone(), x.clone(), x.clone()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        gl.ntt(a)
    for _ in range(25):
        gl.ntt(b)
        gl.intt(b)
    gl.ntt(b)
    with torch.cuda.stream(s):
        gl.ntt(c)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    for name, t in (("side", a), ("null", b), ("side again", c)):
        assert (from_torch(t) == expect).all(), name


def test_host_slice_calls_beside_side_stream_work(gl, oracle):
    """Host-slice transforms (their blocks go back to the pool idle after
    the synchronous copy back) in one thread while another thread keeps a
    side stream busy with device transforms of the same length (the same
    scratch size class): every result equals the oracle's."""
    import threading

    import numpy as np
    import torch
    n = 1 << 16
    note("pool: host-slice calls beside side-stream work, 2^16")
    rng = np.random.default_rng(55)
    hx = (rng.integers(0, 2**63, n, dtype=np.uint64) * 2 + 1) % np.uint64(0xFFFFFFFF00000001)
    hexp = oracle.ntt(hx)
    dx = torch.empty(n, dtype=torch.int64, device="cuda")
    gl.fill_random(dx, 66)
    torch.cuda.synchronize()
    dexp = oracle.ntt(from_torch(dx))
    errors = []

    def host_loop():
        try:
            for i in range(60):
                v = hx.copy()
                gl.ntt(v)
                if not np.array_equal(v, hexp):
                    errors.append(f"host call {i} differs")
                    return
        except Exception as e:  # reported below
            errors.append(repr(e))

    def device_loop():
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for i in range(60):
                    y = dx.clone()
                    gl.ntt(y)
                    s.synchronize()
                    if not np.array_equal(from_torch(y), dexp):
                        errors.append(f"device call {i} differs")
                        return
        except Exception as e:  # reported below
            errors.append(repr(e))

    ts = [threading.Thread(target=host_loop), threading.Thread(target=device_loop)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=300)
    assert not any(t.is_alive() for t in ts), "a thread did not finish"
    assert not errors, errors[0]
