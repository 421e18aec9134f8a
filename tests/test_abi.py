"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "goldilocks_amd", "libgoldilocks_amd.so")


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(gl_[a-z0-9_]+)\s*\(", src):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("gl_ntt", "gl_intt", "gl_ntt_batch", "gl_ntt_dev", "gl_merkle_tree", "gl_commit",
                 "gl_k12_layer", "gl_clear_cache", "gl_ntt_batch_mixed"):
        assert must in names


def test_library_exports_all_symbols():
    assert os.path.exists(LIB), "build the library first (python -m goldilocks_amd.build)"
    L = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_pure_host_helpers():
    import goldilocks_amd as gl
    assert gl.next_layer_size(0) == 0 and gl.next_layer_size(161) == 64
    assert gl.merkle_tree_size(32) == 0 and gl.merkle_tree_size(33) == 32
    assert gl.is_divisor(1 << 32) and not gl.is_divisor(7) and not gl.is_divisor(0)
    with pytest.raises(ValueError):
        gl.strategy(7)


def test_next_size_matches_reference_kats(oracle):
    """divisors::next KATs (ntt/src/divisors.rs:110-115) and agreement with
    the oracle's restatement over a sweep (host-only entry point)."""
    import goldilocks_amd as gl
    assert gl.next_size(0) == 0
    assert gl.next_size(1000) == 1020
    assert gl.next_size(1024) == 1024
    assert gl.next_size(1234) == 1280
    assert gl.next_size(0xFFFFFFFF00000000) == 0xFFFFFFFF00000000
    assert gl.next_size(0xFFFFFFFF00000001) == 0
    for n in list(range(1, 5000, 7)) + [3 ** 7, 10 ** 6, 10 ** 9, 2 ** 40 + 1, 2 ** 62 + 3]:
        assert gl.next_size(n) == oracle.next_divisor(n), n
