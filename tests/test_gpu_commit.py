"""GPU commit (column coset LDE + K12 Merkle) parity against the oracle."""
import numpy as np
import pytest

from gpu_util import from_torch, note, to_torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gl():
    import goldilocks_amd as g
    assert g.lib().gl_init(-1) == 0, g.lib().gl_last_error()
    return g


# Shapes cover the three ways the extension's input is built: the coset
# shift in the intt's last-pass row scale (most shapes; with the tuned
# column schedules 2^14 x 16 and 2^17 x 4 also run three-stage passes last
# in the intt and first in the extension), the shift in the first extension
# pass when the intt ends with the radix-257 Rader kernel (2^10 * 257 rows),
# and the separate coset kernel when the extension starts with a pass that
# is neither (65537 rows).
@pytest.mark.parametrize("rows,cols,blowup", [(16, 4, 2), (8, 3, 2), (1024, 64, 2), (4096, 7, 4),
                                              (2**14, 16, 2), (240, 5, 2), (3 * 2**10, 2, 2), (1, 4, 2),
                                              (257 * 4, 3, 2), (17 * 16, 5, 2), (65537, 1, 2), (2**18, 1, 2),
                                              (2**10 * 257, 1, 2), (2**17, 4, 2)])
def test_commit_host_vs_oracle(gl, oracle, rows, cols, blowup):
    note(f"commit {rows}x{cols} blowup {blowup}")
    m = oracle.fill_random(rows * cols, rows * 31 + cols).reshape(rows, cols)
    lde, layers = gl.commit(m, blowup)
    lde_o, layers_o = oracle.commit(m, blowup)
    assert np.array_equal(lde, lde_o)
    assert layers == layers_o


def test_commit_device_medium(gl, oracle):
    """2^16 x 64 with blowup 2 (a 1/16-height slice of the 2^20 x 64 config)."""
    import torch
    rows, cols = 2**16, 64
    note(f"commit device {rows}x{cols}")
    m = torch.empty((rows, cols), dtype=torch.int64, device="cuda")
    gl.fill_random(m, 11)
    lde, layers = gl.commit(m, 2)
    torch.cuda.synchronize()
    hm = from_torch(m).reshape(rows, cols)
    lde_o, layers_o = oracle.commit(hm, 2)
    assert np.array_equal(from_torch(lde).reshape(2 * rows, cols), lde_o)
    assert layers.cpu().numpy().tobytes() == b"".join(layers_o)


def test_commit_errors(gl):
    with pytest.raises(ValueError):
        gl.commit(np.zeros((7, 2), dtype=np.uint64), 2)


@pytest.mark.parametrize("rows,cols", [(1024, 5), (4096, 64)])
def test_commit_half_exchange_passes(gl, oracle, rows, cols):
    """Column schedules whose intt ends (reversed reads and row scale when it
    is one pass) and whose extension starts (zero extension) with the
    1024-point pass that exchanges 32-bit halves through LDS."""
    vs = {v["name"]: v["id"] for v in gl.variants()}
    k32 = vs["ntt_pass_kernel<32, 32, 256, 8>"]
    k2, k4 = vs["ntt_pass_kernel<2, 1, 256, 256>"], vs["ntt_pass_kernel<4, 1, 256, 256>"]
    plans = {1024: [k32], 2048: [k32, k2], 4096: [k4, k32], 8192: [k32, k4, k2]}
    try:
        for n in (rows, 2 * rows):
            gl.plan_set(n, plans[n], cols=True)
            assert gl.plan_ids(n, cols=True) == plans[n]
        note(f"commit {rows}x{cols} through half-exchange column passes")
        m = oracle.fill_random(rows * cols, rows + 7 * cols).reshape(rows, cols)
        lde, layers = gl.commit(m, 2)
        lde_o, layers_o = oracle.commit(m, 2)
        assert np.array_equal(lde, lde_o)
        assert layers == layers_o
    finally:
        for n in (rows, 2 * rows):
            gl.plan_set(n, [], cols=True)


@pytest.mark.parametrize("rows,cols,first", [(1024, 5, "32, 1"), (2048, 64, "64, 1"), (1024, 3, "16, 1")])
def test_commit_single_stage_extension_pass(gl, oracle, rows, cols, first):
    """The extension starts with a single-sub-step pass; with blowup 2 its
    inputs r >= R/2 are exactly the zero rows, which the 32- and 64-point
    passes skip (lower half read, first codelet layer without additions)."""
    vs = {v["name"]: v["id"] for v in gl.variants()}
    k1 = vs[f"ntt_pass_kernel<{first}, 256, 256>"]
    r1 = int(first.split(",")[0])
    rest = {2048 // 32: [vs["ntt_pass_kernel<8, 8, 256, 32>"]],
            4096 // 64: [vs["ntt_pass_kernel<8, 8, 256, 32>"]],
            2048 // 16: [vs["ntt_pass_kernel<16, 8, 256, 32>"]]}[2 * rows // r1]
    n = 2 * rows
    try:
        gl.plan_set(n, [k1] + rest, cols=True)
        note(f"commit {rows}x{cols}, extension first pass <{first}>")
        m = oracle.fill_random(rows * cols, 3 * rows + cols).reshape(rows, cols)
        lde, layers = gl.commit(m, 2)
        lde_o, layers_o = oracle.commit(m, 2)
        assert np.array_equal(lde, lde_o)
        assert layers == layers_o
    finally:
        gl.plan_set(n, [], cols=True)
