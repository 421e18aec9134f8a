"""GPU row tree (the binary tree of row hashes) and row-tree commit against
the oracle (oracle/k12.c or_row_tree, pinned by test_oracle_k12.py), plus
openings and their verification."""
import numpy as np
import pytest

from gpu_util import from_torch, note, to_torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gl():
    import goldilocks_amd as g
    assert g.lib().gl_init(-1) == 0, g.lib().gl_last_error()
    return g


# leaf counts cover one leaf, odd layers at every height, the single-workgroup
# tail alone (<= 512 nodes) and the layer kernel above it; column counts cover
# one block, a block boundary (21 words = 168 bytes: the padding word opens a
# block of its own), one word either side of it, the 1023-word single-node
# maximum and K12's tree mode beyond it (1024 words: a last chunk holding
# only the customization byte; 1025, 2048, 3000 and 9000 words: one to eight
# chained chunks)
@pytest.mark.parametrize("nrows,cols", [(1, 1), (1, 64), (2, 3), (3, 20), (5, 21), (6, 22), (7, 41), (9, 42),
                                        (33, 64), (512, 8), (513, 8), (1000, 64), (4097, 2), (100, 1023),
                                        (2**16 + 3, 5), (3, 1024), (5, 1025), (600, 2048), (2, 3000), (3, 9000)])
def test_row_tree_vs_oracle(gl, oracle, nrows, cols):
    note(f"row tree {nrows}x{cols}")
    m = oracle.fill_random(nrows * cols, 7 * nrows + cols).reshape(nrows, cols)
    want = oracle.row_tree(m)
    assert gl.row_tree(m) == want
    got = gl.row_tree(to_torch(m)).cpu().numpy().tobytes()
    assert got == b"".join(want)
    assert len(got) == gl.row_tree_size(nrows) == oracle.row_tree_size(nrows)


def test_row_tree_errors(gl):
    import torch
    x = torch.empty(8, dtype=torch.int64, device="cuda")
    # rows wider than 262143 words (pointer never dereferenced)
    rc = gl.lib().gl_row_tree_dev(_p(x), 2, 262144, _p(x), 1 << 20, None, None)
    assert rc == gl.GL_E_SIZE and b"262143" in gl.lib().gl_last_error()
    with pytest.raises(ValueError):
        gl.row_tree(np.zeros((0, 4), dtype=np.uint64))
    with pytest.raises(ValueError):
        gl.commit_rows(np.zeros((7, 4), dtype=np.uint64), 2)


def _p(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("rows,cols,blowup", [(8, 3, 2), (16, 64, 4), (12, 21, 2), (1, 5, 1), (1024, 64, 2),
                                              (3 * 2**10, 5, 2), (257 * 4, 3, 2), (64, 1500, 2)])
def test_commit_rows_vs_oracle(gl, oracle, rows, cols, blowup):
    note(f"commit_rows {rows}x{cols} blowup {blowup}")
    m = oracle.fill_random(rows * cols, rows * 17 + cols).reshape(rows, cols)
    lde_o, layers_o = oracle.commit_rows(m, blowup)
    lde, layers = gl.commit_rows(m, blowup)
    assert np.array_equal(lde, lde_o)
    assert layers == layers_o
    dl, dt = gl.commit_rows(to_torch(m), blowup)
    assert np.array_equal(from_torch(dl).reshape(lde_o.shape), lde_o)
    assert dt.cpu().numpy().tobytes() == b"".join(layers_o)


@pytest.mark.parametrize("leaves,cols", [(1, 6), (2, 6), (5, 6), (513, 6), (1000, 6), (4099, 6), (77, 2100)])
def test_row_open_verify(gl, oracle, leaves, cols):
    import torch
    note(f"row open/verify {leaves} leaves")
    m = oracle.fill_random(leaves * cols, 3 + leaves).reshape(leaves, cols)
    layers_o = oracle.row_tree(m)
    dm = to_torch(m)
    layers = gl.row_tree(dm)
    depth = gl.row_tree_depth(leaves)
    assert depth == len(layers_o) - 1
    rng = np.random.default_rng(leaves)
    idx = sorted(set([0, leaves - 1, leaves // 2] + list(rng.integers(0, leaves, 20))))
    idx_t = torch.tensor(idx + [leaves, leaves + 5], dtype=torch.int64, device="cuda")
    paths = gl.row_open(layers, leaves, idx_t)
    assert paths.shape == (len(idx) + 2, depth, 32)
    hp = paths.cpu().numpy()
    # siblings read from the oracle's layers; zeros where a node moves up
    for i, leaf in enumerate(idx):
        for l in range(depth):
            n_l = len(layers_o[l]) // 32
            sib = (leaf >> l) ^ 1
            want = layers_o[l][32 * sib:32 * sib + 32] if sib < n_l else bytes(32)
            assert hp[i, l].tobytes() == want, (leaf, l)
    assert not hp[len(idx):].any()
    root = layers[-32:]
    assert root.cpu().numpy().tobytes() == layers_o[-1]
    opened = dm[idx_t[:len(idx)]].contiguous()
    ok = gl.row_verify(opened, leaves, idx_t[:len(idx)], paths[:len(idx)].contiguous(), root)
    assert ok.all().item()
    if leaves > 1:
        # a changed word, a changed sibling or a wrong index fails
        bad_rows = opened.clone()
        bad_rows[0, 0] ^= 1
        assert not gl.row_verify(bad_rows, leaves, idx_t[:len(idx)], paths[:len(idx)].contiguous(), root)[0].item()
        bad_paths = paths[:len(idx)].clone()
        bad_paths[-1, 0, 3] ^= 0x40
        flags = gl.row_verify(opened, leaves, idx_t[:len(idx)], bad_paths, root).cpu().numpy()
        assert not flags[-1] or (idx[-1] ^ 1) >= leaves
        shifted = torch.tensor([(j + 1) % leaves for j in idx], dtype=torch.int64, device="cuda")
        flags = gl.row_verify(opened, leaves, shifted, paths[:len(idx)].contiguous(), root).cpu().numpy()
        assert not flags.any()
    # an index past the leaves never verifies
    out = gl.row_verify(opened[:1].contiguous(), leaves, idx_t[len(idx):len(idx) + 1], paths[len(idx):len(idx) + 1]
                        .contiguous(), root)
    assert not out.any().item()


def test_server_commit_rows(gl, oracle):
    rows, cols, blowup = 256, 16, 2
    note("server commit_rows")
    m = oracle.fill_random(rows * cols, 5).reshape(rows, cols)
    lde_o, layers_o = oracle.commit_rows(m, blowup)
    lde = np.zeros((rows * blowup, cols), dtype=np.uint64)
    layers = np.zeros(gl.row_tree_size(rows * blowup), dtype=np.uint8)
    with gl.Server(0, 2) as srv:
        t = srv.submit_commit(m, blowup, lde, layers, row_tree=True)
        srv.wait(t)
    assert np.array_equal(lde, lde_o)
    assert layers.tobytes() == b"".join(layers_o)
