"""Merkle openings from device-resident trees (gl_merkle_open_dev) and
their verification against the root (gl_merkle_verify_dev), checked
against openings built from the oracle's tree (pcs/src/k12/mod.rs:41-60
layers, restated in oracle/commit.c) and re-hashed with the oracle's K12."""
import numpy as np
import pytest

from gpu_util import note

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gl():
    import goldilocks_amd as g
    assert g.lib().gl_init(-1) == 0, g.lib().gl_last_error()
    return g


def oracle_opening(leaf_bytes, layers, b):
    """depth x 160 bytes: block b // 5^l of layer l (layer 0 = leaves)."""
    levels = [leaf_bytes] + layers[:-1]
    out = []
    for lvl in levels:
        blk = lvl[b * 160:(b + 1) * 160]
        out.append(blk + bytes(160 - len(blk)))
        b //= 5
    return b"".join(out)


@pytest.mark.parametrize("nbytes", [8 * 1024, 8 * 65536, 8 * 65536 + 24, 161, 33, 8 * 3 * 5 * 17 * 257])
def test_openings_match_oracle_and_verify(gl, oracle, nbytes):
    import torch
    note(f"merkle openings over {nbytes} bytes")
    words = oracle.fill_random((nbytes + 7) // 8, nbytes)
    leaf = words.view(np.uint8)[:nbytes].copy()
    layers = oracle.merkle_tree(leaf.tobytes())
    depth = gl.merkle_depth(nbytes)
    assert depth == len(layers)
    nblocks = (nbytes + 159) // 160
    rng = np.random.default_rng(nbytes)
    picks = sorted({0, nblocks - 1, *rng.integers(0, nblocks, size=min(40, nblocks)).tolist()})
    d_leaf = torch.from_numpy(leaf).cuda()
    d_layers = gl.merkle_tree(d_leaf)
    blocks = torch.tensor(picks, dtype=torch.int64, device="cuda")
    paths = gl.merkle_open(d_leaf, d_layers, blocks)
    got = paths.cpu().numpy()
    for i, b in enumerate(picks):
        assert got[i].tobytes() == oracle_opening(leaf.tobytes(), layers, b), (nbytes, b)
    # re-hash one opening on the CPU with the oracle's K12
    b = picks[-1]
    for l in range(depth):
        lvl_len = nbytes if l == 0 else len(layers[l - 1])
        bl = min(160, lvl_len - (b // 5 ** l) * 160)
        h = oracle.k12(got[-1][l][:bl].tobytes())
        parent = got[-1][l + 1][32 * ((b // 5 ** l) % 5):][:32].tobytes() if l + 1 < depth else layers[-1]
        assert h == parent
    root = torch.from_numpy(np.frombuffer(layers[-1], dtype=np.uint8).copy()).cuda()
    ok = gl.merkle_verify(paths, nbytes, blocks, root)
    assert bool(ok.all())
    # a changed leaf byte, a changed hash in the next layer's block and a
    # wrong block index all fail
    bad = paths.clone()
    bad[0, 0, 0] ^= 1
    if depth > 1:
        bad[-1, 1, 3] ^= 0x40  # inside the block's first hash, always hashed
    ok2 = gl.merkle_verify(bad, nbytes, blocks, root).cpu().numpy()
    assert not ok2[0]
    if depth > 1 and len(picks) > 1:
        assert not ok2[-1]
    if nblocks > 1:
        swapped = torch.tensor([picks[-1]] + picks[1:], dtype=torch.int64, device="cuda")
        if picks[0] != picks[-1]:
            assert not bool(gl.merkle_verify(paths, nbytes, swapped, root)[0])
