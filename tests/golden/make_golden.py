"""Writes tests/golden/golden.json: regression vectors produced by the oracle
(oracle/, a C restatement of the reference pinned by the tests in
tests/test_oracle_*.py).  Inputs are the seeded splitmix64 words of
oracle.fill_random; outputs are stored as SHA-256 digests plus a few words.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle  # noqa: E402


def digest(a):
    return hashlib.sha256(a.astype("<u8").tobytes()).hexdigest()


def main():
    out = {"ntt": [], "intt": [], "k12": [], "merkle": [], "commit": []}
    sizes = [n for n in oracle.divisors() if 2 <= n <= 128]
    sizes += [256, 1024, 4096, 15 * 16, 8 * 17, 2 * 257, 257, 2**16, 65537]
    for n in sizes:
        x = oracle.fill_random(n, 1000 + n)
        y = oracle.ntt(x)
        out["ntt"].append({"n": n, "seed": 1000 + n, "sha256": digest(y),
                           "head": [int(v) for v in y[:4]], "tail": [int(v) for v in y[-2:]]})
        z = oracle.intt(x)
        out["intt"].append({"n": n, "seed": 1000 + n, "sha256": digest(z), "head": [int(v) for v in z[:4]]})
    for n in (0, 1, 40, 159, 160, 161, 166):
        msg = bytes((7 * i + 3) % 256 for i in range(n))
        out["k12"].append({"len": n, "pattern": "(7*i+3)%256", "h32": oracle.k12(msg, 32).hex()})
    for n_words in (1 << 10, 1 << 16):
        x = oracle.fill_random(n_words, 77)
        layers = oracle.merkle_tree(x.astype("<u8").tobytes())
        out["merkle"].append({"words": n_words, "seed": 77, "layers": len(layers), "root": layers[-1].hex(),
                              "sha256_all": hashlib.sha256(b"".join(layers)).hexdigest()})
    for rows, cols, blowup in ((16, 4, 2), (1024, 8, 2)):
        m = oracle.fill_random(rows * cols, 5).reshape(rows, cols)
        lde, layers = oracle.commit(m, blowup)
        out["commit"].append({"rows": rows, "cols": cols, "blowup": blowup, "seed": 5,
                              "lde_sha256": digest(lde.reshape(-1)), "root": layers[-1].hex()})
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
