"""Committed regression vectors (tests/golden/golden.json, written by
tests/golden/make_golden.py from the oracle): the oracle still reproduces
them (CPU) and the GPU library does too (gpu)."""
import hashlib
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).astype("<u8").tobytes()).hexdigest()


def test_oracle_reproduces_golden(oracle):
    for e in GOLD["ntt"]:
        assert digest(oracle.ntt(oracle.fill_random(e["n"], e["seed"]))) == e["sha256"], e["n"]
    for e in GOLD["intt"]:
        assert digest(oracle.intt(oracle.fill_random(e["n"], e["seed"]))) == e["sha256"], e["n"]
    for e in GOLD["k12"]:
        msg = bytes((7 * i + 3) % 256 for i in range(e["len"]))
        assert oracle.k12(msg, 32).hex() == e["h32"]
    for e in GOLD["merkle"]:
        layers = oracle.merkle_tree(oracle.fill_random(e["words"], e["seed"]).astype("<u8").tobytes())
        assert layers[-1].hex() == e["root"] and len(layers) == e["layers"]
    for e in GOLD["commit"]:
        m = oracle.fill_random(e["rows"] * e["cols"], e["seed"]).reshape(e["rows"], e["cols"])
        lde, layers = oracle.commit(m, e["blowup"])
        assert digest(lde.reshape(-1)) == e["lde_sha256"] and layers[-1].hex() == e["root"]


@pytest.mark.gpu
def test_gpu_reproduces_golden(oracle):
    import goldilocks_amd as gl
    assert gl.lib().gl_init(-1) == 0
    for e in GOLD["ntt"]:
        x = oracle.fill_random(e["n"], e["seed"])
        gl.ntt(x)
        assert digest(x) == e["sha256"], e["n"]
    for e in GOLD["intt"]:
        x = oracle.fill_random(e["n"], e["seed"])
        gl.intt(x)
        assert digest(x) == e["sha256"], e["n"]
    for e in GOLD["k12"]:
        msg = bytes((7 * i + 3) % 256 for i in range(e["len"]))
        assert gl.k12(msg, 32).hex() == e["h32"]
    for e in GOLD["merkle"]:
        layers = gl.merkle_tree(oracle.fill_random(e["words"], e["seed"]).astype("<u8").tobytes())
        assert layers[-1].hex() == e["root"] and len(layers) == e["layers"]
    for e in GOLD["commit"]:
        m = oracle.fill_random(e["rows"] * e["cols"], e["seed"]).reshape(e["rows"], e["cols"])
        lde, layers = gl.commit(m, e["blowup"])
        assert digest(lde.reshape(-1)) == e["lde_sha256"] and layers[-1].hex() == e["root"]
