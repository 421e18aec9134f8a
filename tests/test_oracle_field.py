"""Pins the oracle's field arithmetic and divisor helpers against closed forms
(Python big integers) and the reference's own known answers."""
import random

import pytest

P = 0xFFFFFFFF00000001
EDGE = [0, 1, 2, P - 1, P - 2, 2**32 - 1, 2**32, 2**32 + 1, 2**63, 2**63 - 1,
        0xFFFFFFFF00000000, 0x8000000080000000, 12345678901234567]


def operands(count=300, seed=1):
    rng = random.Random(seed)
    vals = EDGE + [rng.randrange(P) for _ in range(count)]
    return vals


def test_add_sub_mul(oracle):
    vals = operands()
    rng = random.Random(2)
    for a in vals:
        for b in (rng.choice(vals), rng.randrange(P), a):
            assert oracle.add(a, b) == (a + b) % P
            assert oracle.sub(a, b) == (a - b) % P
            assert oracle.mul(a, b) == (a * b) % P


def test_mont_mul(oracle):
    rinv = pow(2**64, -1, P)
    for a in operands(100):
        b = (a * 7919 + 3) % P
        assert oracle.mont_mul(a, b) == (a * b * rinv) % P


def test_inv_pow(oracle):
    assert oracle.inv(0) == 0
    for a in operands(200):
        if a == 0:
            continue
        assert oracle.inv(a) == pow(a, -1, P)
        e = (a * 31) % (2**64)
        assert oracle.pow_(a, e) == pow(a, e, P)


def test_shift_and_root_384(oracle):
    rng = random.Random(3)
    for a in operands(40):
        for s in list(range(0, 200)) + [rng.randrange(2**20) for _ in range(5)]:
            assert oracle.shift(a, s) == (a * pow(2, s, P)) % P
    w384 = oracle.root(384)
    for a in operands(20):
        for i in range(384):
            assert oracle.root_384(a, i) == (a * pow(w384, i, P)) % P


def test_roots_and_constants(oracle):
    order = P - 1
    # GENERATOR_R = 2717 * 2^64 mod p (ntt/src/field/algo/mod.rs:13)
    assert (2717 << 64) % P == 0xA9CFFFFF563
    # 2717 generates the whole group: no proper-divisor power is 1.
    for q in (2, 3, 5, 17, 257, 65537):
        assert pow(2717, order // q, P) != 1
    for n in (2, 3, 4, 5, 15, 17, 192, 257, 384, 1 << 20, 1 << 32, 65537 * 3):
        w = oracle.root(n)
        assert w == pow(2717, order // n, P)
        assert pow(w, n, P) == 1
    assert oracle.root(7) == 0 and oracle.root(0) == 0
    # special roots that make small twiddles shifts (ntt/src/bin/codegen.rs:10-36)
    assert oracle.root(192) == 2
    assert oracle.root(384) == (2**24 - 2**72) % P
    assert oracle.root(4) == 2**48
    assert oracle.root(3) == 2**64 % P
    assert oracle.root(5) == 1041288259238279555
    # field/mod.rs:343-349: root(5)^5 == 1, root(4)^2 == -1
    assert pow(oracle.root(5), 5, P) == 1
    assert pow(oracle.root(4), 2, P) == P - 1


def test_divisors(oracle):
    # divisors.rs:101-115
    d = oracle.divisors()
    assert len(d) == 1056 and d[0] == 1 and d[-1] == P - 1
    assert d == sorted(d)
    assert all(oracle.is_divisor(x) for x in d)
    assert oracle.next_divisor(0) == 0
    assert oracle.next_divisor(1000) == 1020
    assert oracle.next_divisor(1024) == 1024
    assert oracle.next_divisor(1234) == 1280
    assert not oracle.is_divisor(0) and not oracle.is_divisor(7)
    assert oracle.is_smooth(15 * 2**20) and not oracle.is_smooth(17)


def test_split_bruteforce(oracle):
    # split(n): the largest d with d | n, d <= n/d (divisors.rs:61-89)
    for n in oracle.divisors()[:400]:
        best = max(x for x in range(1, int(n**0.5) + 2) if n % x == 0 and n // x >= x)
        assert oracle.split(n) == best, n
