"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the product.
The C sources restate the reference (recmo/goldilocks); see oracle.h.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
P = 0xFFFFFFFF00000001
GENERATOR = 2717

_lib = None


def build():
    """Compile liboracle.so from this directory's C sources."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u64, sz, i32 = ctypes.c_uint64, ctypes.c_size_t, ctypes.c_int
        vp, cp = ctypes.c_void_p, ctypes.c_char_p
        sigs = {
            "or_add": (u64, [u64, u64]), "or_sub": (u64, [u64, u64]),
            "or_mul": (u64, [u64, u64]), "or_mont_mul": (u64, [u64, u64]),
            "or_pow": (u64, [u64, u64]), "or_inv": (u64, [u64]),
            "or_shift": (u64, [u64, u64]), "or_root_384": (u64, [u64, u64]),
            "or_root": (u64, [u64]),
            "or_is_divisor": (i32, [u64]), "or_is_smooth": (i32, [u64]),
            "or_divisors": (sz, [vp]), "or_next": (u64, [u64]), "or_split": (u64, [u64]),
            "or_naive_ntt": (i32, [vp, sz]), "or_naive_intt": (i32, [vp, sz]),
            "or_ntt": (i32, [vp, sz]), "or_intt": (i32, [vp, sz]),
            "or_ntt_batch": (i32, [vp, sz, sz, i32]),
            "or_eval_dft": (u64, [vp, sz, u64]),
            "or_convolve": (i32, [vp, vp, sz]),
            "or_set_threads": (None, [i32]),
            "or_keccak_p1600": (None, [vp, i32]),
            "or_k12": (i32, [cp, sz, vp, sz]),
            "or_sha3_256": (None, [cp, sz, vp]),
            "or_keccak_sponge": (i32, [cp, sz, sz, ctypes.c_uint8, i32, vp, sz]),
            "or_k12_long": (i32, [cp, sz, vp, sz]),
            "or_row_tree": (i32, [vp, sz, sz, vp, sz, ctypes.POINTER(sz)]),
            "or_row_tree_size": (sz, [sz]),
            "or_next_layer_size": (sz, [sz]),
            "or_process_layer": (None, [cp, sz, vp]),
            "or_merkle_tree": (i32, [cp, sz, vp, sz, ctypes.POINTER(sz)]),
            "or_merkle_tree_size": (sz, [sz]),
            "or_transpose_index": (sz, [sz, sz, sz]), "or_transpose": (i32, [vp, sz, sz]),
            "or_bit_reverse_index": (sz, [sz, sz]), "or_bit_reverse": (i32, [vp, sz]),
            "or_pow_verify": (i32, [cp, u64, u64]),
            "or_pow_solve": (u64, [cp, u64]),
            "or_commit": (i32, [vp, sz, sz, ctypes.c_uint, vp, vp, sz, ctypes.POINTER(sz)]),
            "or_commit_rows": (i32, [vp, sz, sz, ctypes.c_uint, vp, vp, sz, ctypes.POINTER(sz)]),
            "or_fill_random": (None, [vp, sz, u64]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _u64(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a


# ---- field ----
def add(a, b): return lib().or_add(a, b)
def sub(a, b): return lib().or_sub(a, b)
def mul(a, b): return lib().or_mul(a, b)
def mont_mul(a, b): return lib().or_mont_mul(a, b)
def pow_(a, e): return lib().or_pow(a, e)
def inv(a): return lib().or_inv(a)
def shift(a, s): return lib().or_shift(a, s)
def root_384(a, i): return lib().or_root_384(a, i)
def root(n): return lib().or_root(n)
def is_divisor(n): return bool(lib().or_is_divisor(n))
def is_smooth(n): return bool(lib().or_is_smooth(n))
def next_divisor(n): return lib().or_next(n)
def split(n): return lib().or_split(n)


def divisors():
    out = np.zeros(1056, dtype=np.uint64)
    n = lib().or_divisors(_ptr(out))
    return [int(x) for x in out[:n]]


# ---- NTT ----
def _run(fn, v, *args):
    v = _u64(v).copy()
    rc = fn(_ptr(v), *args)
    if rc != 0:
        raise ValueError("unsupported NTT size")
    return v


def naive_ntt(v): return _run(lib().or_naive_ntt, v, len(v))
def naive_intt(v): return _run(lib().or_naive_intt, v, len(v))
def ntt(v): return _run(lib().or_ntt, v, len(v))
def intt(v): return _run(lib().or_intt, v, len(v))


def ntt_batch(v, n, inverse=False):
    v = _u64(v)
    assert v.size % n == 0 if n else v.size == 0
    return _run(lib().or_ntt_batch, v, n, v.size // n if n else 0, int(bool(inverse)))


# ---- permutations ----
def transpose_index(width, height, i): return lib().or_transpose_index(width, height, i)
def bit_reverse_index(size, i): return lib().or_bit_reverse_index(size, i)


def transpose(m, rows, cols):
    """A rows x cols row-major matrix (flat) -> its cols x rows transpose (flat)."""
    v = _u64(m).reshape(-1).copy()
    assert v.size == rows * cols
    if lib().or_transpose(_ptr(v), rows, cols) != 0:
        raise MemoryError("or_transpose")
    return v


def bit_reverse(v):
    v = _u64(v).copy()
    if lib().or_bit_reverse(_ptr(v), v.size) != 0:
        raise ValueError("bit_reverse: length must be a power of two")
    return v


def eval_dft(x, k):
    x = _u64(x)
    return lib().or_eval_dft(_ptr(x), x.size, k)


def convolve(a, b_ntt):
    """circular(a, b) (convolve.rs:13-20) on copies: intt(ntt(a) * b_ntt)."""
    a = _u64(a).copy()
    b = _u64(b_ntt)
    assert a.size == b.size
    if lib().or_convolve(_ptr(a), _ptr(b), a.size) != 0:
        raise ValueError("unsupported length")
    return a


def set_threads(t): lib().or_set_threads(int(t))


# ---- K12 / Merkle ----
def keccak_p1600(state, rounds=12):
    st = np.array(state, dtype=np.uint64)
    assert st.size == 25
    lib().or_keccak_p1600(_ptr(st), rounds)
    return [int(x) for x in st]


def k12(data: bytes, outlen=32) -> bytes:
    out = ctypes.create_string_buffer(outlen)
    if lib().or_k12(bytes(data), len(data), out, outlen) != 0:
        raise ValueError("k12 input longer than 166 bytes or output longer than 168")
    return out.raw


def sha3_256(data: bytes) -> bytes:
    out = ctypes.create_string_buffer(32)
    lib().or_sha3_256(bytes(data), len(data), out)
    return out.raw


def keccak_sponge(data: bytes, rate, suffix, rounds, outlen) -> bytes:
    out = ctypes.create_string_buffer(max(1, outlen))
    if lib().or_keccak_sponge(bytes(data), len(data), rate, suffix, rounds, out, outlen) != 0:
        raise ValueError("bad sponge parameters")
    return out.raw[:outlen]


def k12_long(data: bytes, outlen=32) -> bytes:
    """KangarooTwelve (empty customization) of any length."""
    out = ctypes.create_string_buffer(outlen)
    if lib().or_k12_long(bytes(data), len(data), out, outlen) != 0:
        raise ValueError("k12_long output longer than 168 bytes")
    return out.raw


def row_tree_size(leaves): return lib().or_row_tree_size(leaves)


def _row_layers(raw, leaves):
    layers, n, off = [], leaves, 0
    while n >= 1:
        layers.append(raw[off:off + 32 * n])
        off += 32 * n
        if n == 1:
            break
        n = (n + 1) // 2
    return layers


def row_tree(rows):
    """Layers (leaves first, root last) of the binary tree of row hashes of a
    row-major matrix of words."""
    rows = _u64(rows)
    nrows, cols = rows.shape
    need = row_tree_size(nrows)
    out = ctypes.create_string_buffer(max(1, need))
    got = ctypes.c_size_t(0)
    if lib().or_row_tree(_ptr(rows), nrows, cols, out, need, ctypes.byref(got)) != 0:
        raise ValueError("unsupported row-tree shape")
    return _row_layers(out.raw[:need], nrows)


def next_layer_size(n): return lib().or_next_layer_size(n)


def process_layer(data: bytes) -> bytes:
    out = ctypes.create_string_buffer(max(1, next_layer_size(len(data))))
    lib().or_process_layer(bytes(data), len(data), out)
    return out.raw[: next_layer_size(len(data))]


def merkle_tree_size(n): return lib().or_merkle_tree_size(n)


def merkle_tree(data: bytes):
    """Layers as a list of bytes objects (pcs/src/k12/mod.rs:41-60)."""
    data = bytes(data)
    need = merkle_tree_size(len(data))
    out = ctypes.create_string_buffer(max(1, need))
    got = ctypes.c_size_t(0)
    rc = lib().or_merkle_tree(data, len(data), out, need, ctypes.byref(got))
    assert rc == 0
    raw = out.raw[:need]
    layers, n, off = [], len(data), 0
    while n > 32:
        n = next_layer_size(n)
        layers.append(raw[off:off + n])
        off += n
    return layers


def pow_verify(challenge: bytes, threshold: int, nonce: int) -> bool:
    return bool(lib().or_pow_verify(bytes(challenge), threshold, nonce))


def pow_solve(challenge: bytes, threshold: int) -> int:
    return lib().or_pow_solve(bytes(challenge), threshold)


# ---- commit ----
def commit(m, blowup):
    """(lde[(blowup*rows), cols], layers list) for a row-major rows x cols matrix."""
    m = _u64(m)
    rows, cols = m.shape
    lde = np.zeros((rows * blowup, cols), dtype=np.uint64)
    need = merkle_tree_size(lde.size * 8)
    out = ctypes.create_string_buffer(max(1, need))
    got = ctypes.c_size_t(0)
    rc = lib().or_commit(_ptr(m), rows, cols, blowup, _ptr(lde), out, need, ctypes.byref(got))
    if rc != 0:
        raise ValueError("unsupported commit shape")
    raw = out.raw[:need]
    layers, n, off = [], lde.size * 8, 0
    while n > 32:
        n = next_layer_size(n)
        layers.append(raw[off:off + n])
        off += n
    return lde, layers


def commit_rows(m, blowup):
    """(lde, row-tree layers) of the row-tree commit."""
    m = _u64(m)
    rows, cols = m.shape
    lde = np.zeros((rows * blowup, cols), dtype=np.uint64)
    need = row_tree_size(rows * blowup)
    out = ctypes.create_string_buffer(max(1, need))
    got = ctypes.c_size_t(0)
    rc = lib().or_commit_rows(_ptr(m), rows, cols, blowup, _ptr(lde), out, need, ctypes.byref(got))
    if rc != 0:
        raise ValueError("unsupported commit shape")
    return lde, _row_layers(out.raw[:need], rows * blowup)


def fill_random(n, seed):
    v = np.zeros(n, dtype=np.uint64)
    lib().or_fill_random(_ptr(v), n, seed)
    return v
