"""Error paths under forced device allocation failure (gl_debug_fail_allocs,
a host-side switch: the next allocations of the library's pool fail as if
hipMalloc had).  Every entry point must return GL_E_NOMEM without touching
its inputs, and the library -- the server included -- must keep working
afterwards.  No kernel is made to fail; only the host refuses memory."""
import numpy as np
import pytest

from gpu_util import from_torch, note, to_torch

pytestmark = pytest.mark.gpu
GL_E_NOMEM = 2


@pytest.fixture(scope="module")
def gl():
    import goldilocks_amd as g
    assert g.lib().gl_init(-1) == 0, g.lib().gl_last_error()
    yield g
    g.lib().gl_debug_fail_allocs(0)


def test_device_ntt_fails_cleanly(gl, oracle):
    import torch
    n = 1 << 20  # two passes: needs pool scratch
    x = oracle.fill_random(n, 61)
    t = to_torch(x)
    gl.ntt(t.clone())  # plans and tables built
    torch.cuda.synchronize()
    note("forced allocation failure: gl_ntt_dev")
    gl.lib().gl_debug_fail_allocs(1)
    with pytest.raises(Exception) as e:
        gl.ntt(t)
    assert getattr(e.value, "code", None) == GL_E_NOMEM, e.value
    torch.cuda.synchronize()
    assert np.array_equal(from_torch(t), x), "input untouched after the failure"
    gl.ntt(t)
    torch.cuda.synchronize()
    assert np.array_equal(from_torch(t), oracle.ntt(x))


def test_host_entry_points_fail_cleanly(gl, oracle):
    n = 1 << 16
    x = oracle.fill_random(n, 62)
    for name, call in (("ntt", lambda v: gl.ntt(v)),
                       ("merkle_tree", lambda v: gl.merkle_tree(v)),
                       ("commit", lambda v: gl.commit(v.reshape(1024, 64))),
                       ("transpose", lambda v: gl.transpose(v, 256, 256)),
                       ("bit_reverse", lambda v: gl.bit_reverse(v))):
        note(f"forced allocation failure: {name}")
        call(x.copy())  # plans and tables of the shape built beforehand
        v = x.copy()
        gl.lib().gl_debug_fail_allocs(1)
        with pytest.raises(Exception) as e:
            call(v)
        assert getattr(e.value, "code", None) == GL_E_NOMEM, (name, e.value)
        assert np.array_equal(v, x), f"{name}: input untouched"
        gl.lib().gl_debug_fail_allocs(0)
        call(v)  # works again
    y = x.copy()
    gl.ntt(y)
    assert np.array_equal(y, oracle.ntt(x))


def test_pipelined_host_batch_fails_cleanly(gl, oracle):
    """A host batch large enough for the pipelined path (nine 2^20-word
    transforms, chunks of 4 + 4 + 1 over two device buffers): an allocation
    failure before the first chunk leaves the input untouched; one in the
    middle (the second chunk's transform scratch, after the two buffers and
    the first chunk's scratch) returns GL_E_NOMEM with the helper thread
    joined; the library works afterwards."""
    n, count = 1 << 20, 9
    x = oracle.fill_random(n * count, 64)
    s = gl.strategy(n)
    s.ntt(x.copy())  # plan and tables built
    for skip in (0, 3):
        note(f"forced allocation failure: pipelined host batch, skip {skip}")
        v = x.copy()
        gl.lib().gl_debug_fail_allocs_after(skip, 1)
        with pytest.raises(Exception) as e:
            s.ntt(v)
        assert getattr(e.value, "code", None) == GL_E_NOMEM, (skip, e.value)
        gl.lib().gl_debug_fail_allocs(0)
        if skip == 0:
            assert np.array_equal(v, x), "input untouched"
    y = x.copy()
    s.ntt(y)
    assert np.array_equal(y[:n], oracle.ntt(x[:n])) and np.array_equal(y[-n:], oracle.ntt(x[-n:]))


def test_server_job_fails_and_server_continues(gl, oracle):
    n, cols = 1 << 16, 4
    x = oracle.fill_random(n * cols, 63)
    with gl.Server(hbm_budget_bytes=1 << 30, slots=2) as srv:
        warm = x.copy()
        srv.wait(srv.submit_ntt(warm, n))  # plans and tables built
        note("forced allocation failure: server job")
        bad = x.copy()
        gl.lib().gl_debug_fail_allocs(1)
        ticket = srv.submit_ntt(bad, n)
        with pytest.raises(Exception) as e:
            srv.wait(ticket)
        gl.lib().gl_debug_fail_allocs(0)
        assert getattr(e.value, "code", None) == GL_E_NOMEM, e.value
        good = x.copy()
        srv.wait(srv.submit_ntt(good, n))
        want = x.copy()
        for c in range(cols):
            want[c * n:(c + 1) * n] = oracle.ntt(x[c * n:(c + 1) * n])
        assert np.array_equal(good, want)
        st = srv.stats()
        assert st["failed"] >= 1
