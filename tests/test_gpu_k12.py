"""GPU K12 / Merkle parity: the reference's known answers and the oracle."""
import random

import numpy as np
import pytest

from gpu_util import note

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gl():
    import goldilocks_amd as g
    assert g.lib().gl_init(-1) == 0, g.lib().gl_last_error()
    return g


def hx(s):
    return bytes.fromhex(s.replace(" ", ""))


def test_k12_kats(gl):
    note("k12 KATs")
    # pcs/src/k12/generic.rs:113-160
    assert gl.k12(b"", 32) == hx(
        "1a c2 d4 50 fc 3b 42 05 d1 9d a7 bf ca 1b 37 51 3c 08 03 57 7a c7 16 7f 06 fe 2c e1 f0 ef 39 e5")
    assert gl.k12(b"", 64)[32:] == hx(
        "42 69 c0 56 b8 c8 2e 48 27 60 38 b6 d2 92 96 6c c0 7a 3d 46 45 27 2e 31 ff 38 50 81 39 eb 0a 71")
    ptn = lambda n: bytes(i % 0xFA for i in range(n))
    assert gl.k12(ptn(1), 32) == hx(
        "2b da 92 45 0e 8b 14 7f 8a 7c b6 29 e7 84 a0 58 ef ca 7c f7 d8 21 8e 02 d3 45 df aa 65 24 4a 1f")
    assert gl.k12(ptn(17), 32) == hx(
        "6b f7 5f a2 23 91 98 db 47 72 e3 64 78 f8 e1 9b 0f 37 12 05 f6 a9 a9 3a 27 3f 51 df 37 12 28 88")
    # k12/aarch64.rs:234-240 (first output word of K12(challenge || nonce))
    ch = bytes([0x55] * 32)
    assert int.from_bytes(gl.k12(ch + (2**64 - 1).to_bytes(8, "little"), 8), "little") == 14608096357653754266
    assert int.from_bytes(gl.k12(ch + (0).to_bytes(8, "little"), 8), "little") == 2237294155800119683
    with pytest.raises(ValueError):
        gl.k12(bytes(167))


def test_k12_all_lengths_vs_oracle(gl, oracle):
    note("k12 lengths 0..166")
    rng = random.Random(5)
    for n in range(0, 167):
        msg = bytes(rng.randrange(256) for _ in range(n))
        for H in (32, 8, 168):
            assert gl.k12(msg, H) == oracle.k12(msg, H), (n, H)


def test_process_layer_vs_oracle(gl, oracle):
    rng = random.Random(6)
    for n in [1, 31, 32, 159, 160, 161, 320, 1600, 1637, 160 * 1000 + 7, 1 << 20]:
        note(f"layer len={n}")
        data = bytes(rng.randrange(256) for _ in range(n)) if n < 100000 else np.random.default_rng(n).integers(0, 256, n, dtype=np.uint8).tobytes()
        assert gl.process_layer(data) == oracle.process_layer(data), n


def test_unaligned_input(gl, oracle):
    note("layer unaligned")
    data = bytes(i % 251 for i in range(160 * 50 + 1))
    assert gl.process_layer(data[1:]) == oracle.process_layer(data[1:])


def test_merkle_tree_vs_oracle(gl, oracle):
    for n in [0, 32, 33, 160, 161, 4096, 160 * 5 ** 4, 8 << 16, 3_000_001]:
        note(f"merkle len={n}")
        data = np.random.default_rng(n).integers(0, 256, n, dtype=np.uint8).tobytes()
        assert gl.merkle_tree(data) == oracle.merkle_tree(data), n


def test_merkle_tree_reference_input(gl, oracle):
    """pcs/src/k12/mod.rs:75-84 input: (0..0xfa).cycle().take(1 << 20)."""
    note("merkle reference pattern")
    data = bytes(i % 0xFA for i in range(1 << 20))
    layers = gl.merkle_tree(data)
    assert all(len(l) % 32 == 0 for l in layers) and len(layers[-1]) == 32
    assert layers == oracle.merkle_tree(data)


def test_merkle_tree_device_tensor(gl, oracle):
    import torch
    note("merkle device tensor 2^16 field elements")
    x = torch.empty(1 << 16, dtype=torch.int64, device="cuda")
    gl.fill_random(x, 3)
    out = gl.merkle_tree(x)
    torch.cuda.synchronize()
    raw = out.cpu().numpy().tobytes()
    host = x.cpu().numpy().tobytes()
    assert raw == b"".join(oracle.merkle_tree(host))
