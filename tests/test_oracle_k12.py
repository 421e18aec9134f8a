"""Pins the oracle's Keccak/K12/Merkle against an independent library (hashlib
SHA3 for the permutation) and the reference's known-answer tests."""
import hashlib
import random

import pytest


def test_permutation_vs_hashlib(oracle):
    rng = random.Random(1)
    for n in [0, 1, 3, 135, 136, 137, 200, 1000] + [rng.randrange(600) for _ in range(20)]:
        msg = bytes(rng.randrange(256) for _ in range(n))
        assert oracle.sha3_256(msg) == hashlib.sha3_256(msg).digest()


def hx(s):
    return bytes.fromhex(s.replace(" ", ""))


def test_xkcp_kats(oracle):
    # pcs/src/k12/generic.rs:113-160 (XKCP KangarooTwelve.txt)
    assert oracle.k12(b"", 32) == hx(
        "1a c2 d4 50 fc 3b 42 05 d1 9d a7 bf ca 1b 37 51 3c 08 03 57 7a c7 16 7f 06 fe 2c e1 f0 ef 39 e5")
    assert oracle.k12(b"", 64) == hx(
        "1a c2 d4 50 fc 3b 42 05 d1 9d a7 bf ca 1b 37 51 3c 08 03 57 7a c7 16 7f 06 fe 2c e1 "
        "f0 ef 39 e5 42 69 c0 56 b8 c8 2e 48 27 60 38 b6 d2 92 96 6c c0 7a 3d 46 45 27 2e 31 "
        "ff 38 50 81 39 eb 0a 71")
    ptn = lambda n: bytes(i % 0xFA for i in range(n))
    assert oracle.k12(ptn(1), 32) == hx(
        "2b da 92 45 0e 8b 14 7f 8a 7c b6 29 e7 84 a0 58 ef ca 7c f7 d8 21 8e 02 d3 45 df aa 65 24 4a 1f")
    assert oracle.k12(ptn(17), 32) == hx(
        "6b f7 5f a2 23 91 98 db 47 72 e3 64 78 f8 e1 9b 0f 37 12 05 f6 a9 a9 3a 27 3f 51 df 37 12 28 88")


def test_k12_limits(oracle):
    oracle.k12(bytes(166), 32)
    with pytest.raises(ValueError):
        oracle.k12(bytes(167), 32)


def test_pow_kats(oracle):
    ch = bytes([0x55] * 32)
    # pow.rs:67-85
    assert all(oracle.pow_verify(ch, 0, i) for i in range(4))
    assert [oracle.pow_verify(ch, 1 << 63, i) for i in range(4)] == [False, False, False, True]
    # pow.rs:87-99
    assert oracle.pow_solve(ch, 1 << 63) == 3
    assert oracle.pow_solve(ch, (2**64 - 1) - 2**48) == 18455

    # k12/aarch64.rs:234-240: k12_pow_2([0x55;32], u64::MAX, 0)
    def first_word(nonce):
        h = oracle.k12(ch + nonce.to_bytes(8, "little"), 8)
        return int.from_bytes(h, "little")
    assert first_word(2**64 - 1) == 14608096357653754266
    assert first_word(0) == 2237294155800119683


def test_layers(oracle):
    assert oracle.next_layer_size(0) == 0
    assert oracle.next_layer_size(1) == 32
    assert oracle.next_layer_size(160) == 32
    assert oracle.next_layer_size(161) == 64
    data = bytes(i % 251 for i in range(1600 + 37))
    out = oracle.process_layer(data)
    assert len(out) == 11 * 32
    for i in range(11):
        assert out[32 * i:32 * i + 32] == oracle.k12(data[160 * i:160 * i + 160], 32)


def test_merkle_tree(oracle):
    # mod.rs:75-84: every layer is a multiple of 32 bytes; stop at <= 32.
    data = bytes(i % 0xFA for i in range(1 << 20))
    layers = oracle.merkle_tree(data)
    assert all(len(l) % 32 == 0 for l in layers)
    assert len(layers[-1]) == 32
    sizes, n = [], len(data)
    while n > 32:
        n = oracle.next_layer_size(n)
        sizes.append(n)
    assert [len(l) for l in layers] == sizes
    prev = data
    for l in layers:
        assert l == oracle.process_layer(prev)
        prev = l
    assert oracle.merkle_tree(b"") == [] and oracle.merkle_tree(bytes(32)) == []
    assert len(oracle.merkle_tree(bytes(33))) == 1


def test_sponge_vs_hashlib(oracle):
    """The byte-granular sponge with FIPS 202's SHAKE128 and SHA3-256
    parameters against hashlib, across the block boundaries."""
    rng = random.Random(3)
    lens = [0, 1, 166, 167, 168, 169, 335, 336, 337, 503, 504, 505, 1000] + [rng.randrange(1200) for _ in range(12)]
    for n in lens:
        msg = bytes(rng.randrange(256) for _ in range(n))
        for outlen in (1, 32, 167, 168):
            assert oracle.keccak_sponge(msg, 168, 0x1F, 24, outlen) == hashlib.shake_128(msg).digest(outlen), (n, outlen)
        assert oracle.keccak_sponge(msg, 136, 0x06, 24, 32) == hashlib.sha3_256(msg).digest(), n
    with pytest.raises(ValueError):
        oracle.keccak_sponge(b"", 168, 0x1F, 24, 169)


def test_k12_long(oracle):
    rng = random.Random(4)
    # one permutation: the reference-faithful single-block K12
    for n in range(167):
        msg = bytes(rng.randrange(256) for _ in range(n))
        assert oracle.k12_long(msg, 32) == oracle.k12(msg, 32), n
    assert oracle.k12_long(b"", 168) == oracle.k12(b"", 168)
    # several blocks, single node (XKCP KangarooTwelve.txt; RFC 9861 section
    # 5): the pattern repeats 0x00..0xFA, 251 values (the same bytes as the
    # reference's (0..0xfa).cycle() only below 250)
    ptn = lambda n: bytes(i % 251 for i in range(n))
    assert oracle.k12_long(ptn(17 ** 2), 32) == hx(
        "0c 31 5e bc de db f6 14 26 de 7d cf 8f b7 25 d1 e7 46 75 d7 f5 32 7a 50 67 f3 67 b1 08 ec b6 7c")
    assert oracle.k12_long(ptn(17 ** 3), 32) == hx(
        "cb 55 2e 2e c7 7d 99 10 70 1d 57 8b 45 7d df 77 2c 12 e3 22 e4 ee 7f e4 17 f9 2c 75 8f 0d 59 d0")
    # the last single-node length and the first tree-mode lengths (RFC 9861
    # section 5): 8191 and 8192 bytes, then 17^4 (11 chunks) and 17^5 (174)
    assert oracle.k12_long(ptn(8191), 32) == hx(
        "1b 57 76 36 f7 23 64 3e 99 0c c7 d6 a6 59 83 74 36 fd 6a 10 36 26 60 0e b8 30 1c d1 db e5 53 d6")
    assert oracle.k12_long(ptn(8192), 32) == hx(
        "48 f2 56 f6 77 2f 9e df b6 a8 b6 61 ec 92 dc 93 b9 5e bd 05 a0 8a 17 b3 9a e3 49 08 70 c9 26 c3")
    assert oracle.k12_long(ptn(17 ** 4), 32) == hx(
        "87 01 04 5e 22 20 53 45 ff 4d da 05 55 5c bb 5c 3a f1 a7 71 c2 b8 9b ae f3 7d b4 3d 99 98 b9 fe")
    assert oracle.k12_long(ptn(17 ** 5), 32) == hx(
        "84 4d 61 09 33 b1 b9 96 3c bd eb 5a e3 b6 b0 5c c7 cb d6 7c ee df 88 3e b6 78 a0 a8 e0 37 16 82")
    with pytest.raises(ValueError):
        oracle.k12_long(b"", 169)


def row_tree_py(oracle, rows):
    """The row tree assembled here from the pinned hashes: leaf r =
    k12_long(row bytes), parent = k12(left || right), odd last node moved up."""
    layer = [oracle.k12_long(r.astype("<u8").tobytes(), 32) for r in rows]
    layers = [b"".join(layer)]
    while len(layer) > 1:
        nxt = [oracle.k12(layer[2 * j] + layer[2 * j + 1], 32) if 2 * j + 1 < len(layer) else layer[2 * j]
               for j in range((len(layer) + 1) // 2)]
        layers.append(b"".join(nxt))
        layer = nxt
    return layers


@pytest.mark.parametrize("shape", [(1, 1), (1, 64), (2, 3), (3, 20), (5, 21), (6, 22), (7, 41), (9, 42),
                                   (33, 64), (100, 1023), (257, 5), (3, 1024), (2, 1025), (2, 2048), (1, 3000)])
def test_row_tree(oracle, shape):
    import numpy as np
    rows = oracle.fill_random(shape[0] * shape[1], 11 + shape[0]).reshape(shape)
    got = oracle.row_tree(rows)
    assert got == row_tree_py(oracle, rows)
    assert sum(len(l) for l in got) == oracle.row_tree_size(shape[0])
    assert len(got[-1]) == 32
    with pytest.raises(ValueError):
        oracle.row_tree(np.zeros((0, 4), dtype=np.uint64))
