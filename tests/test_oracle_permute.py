"""Oracle permutations (ntt/src/permute/) pinned by the reference's own
known answers and by independent computations (numpy's transpose, a
bit-reversal built from string reversal)."""
import numpy as np
import pytest


def test_transpose_kats(oracle):
    # permute/mod.rs test_copy_2_3 and test_copy_3_3
    assert list(oracle.transpose(np.arange(6), 2, 3)) == [0, 3, 1, 4, 2, 5]
    assert list(oracle.transpose(np.arange(9), 3, 3)) == [0, 3, 6, 1, 4, 7, 2, 5, 8]
    # permute/permutation.rs test_transpose: transpose(3, 2)
    assert [oracle.transpose_index(3, 2, i) for i in range(6)] == [0, 3, 1, 4, 2, 5]


@pytest.mark.parametrize("rows,cols", [(1, 1), (1, 7), (7, 1), (5, 3), (16, 16), (33, 65), (100, 37), (256, 1024)])
def test_transpose_vs_numpy(oracle, rows, cols):
    m = oracle.fill_random(rows * cols, rows * 1000 + cols)
    got = oracle.transpose(m, rows, cols)
    assert np.array_equal(got, m.reshape(rows, cols).T.reshape(-1))
    # the index map is the destination -> source permutation of the same move
    src = [oracle.transpose_index(cols, rows, i) for i in range(rows * cols)]
    assert np.array_equal(got, m[src])
    # transposing back restores the matrix
    assert np.array_equal(oracle.transpose(got, cols, rows), m)


def _rev(i, bits):
    return int(format(i, f"0{bits}b")[::-1], 2) if bits else 0


@pytest.mark.parametrize("k", [0, 1, 2, 3, 5, 8, 10, 13])
def test_bit_reverse_vs_string_reversal(oracle, k):
    n = 1 << k
    v = oracle.fill_random(n, 40 + k)
    got = oracle.bit_reverse(v)
    assert np.array_equal(got, v[[_rev(i, k) for i in range(n)]])
    assert [oracle.bit_reverse_index(n, i) for i in range(n)] == [_rev(i, k) for i in range(n)]
    assert np.array_equal(oracle.bit_reverse(got), v)  # an involution


def test_bit_reverse_rejects_non_powers_of_two(oracle):
    for n in (0, 3, 6, 1000):
        with pytest.raises(ValueError):
            oracle.bit_reverse(np.zeros(n, dtype=np.uint64))
