"""Pins the oracle's commit (column coset LDE + Merkle) at tiny sizes by
direct polynomial evaluation with Python big integers."""
import numpy as np

P = 0xFFFFFFFF00000001


def test_commit_small_by_evaluation(oracle):
    rows, cols, blowup = 8, 3, 2
    m = oracle.fill_random(rows * cols, 77).reshape(rows, cols)
    lde, layers = oracle.commit(m, blowup)
    ext = rows * blowup
    w_rows = pow(2717, (P - 1) // rows, P)
    w_ext = pow(2717, (P - 1) // ext, P)
    inv_rows = pow(rows, P - 2, P)
    for c in range(cols):
        vals = [int(v) for v in m[:, c]]
        # coefficients of the interpolant through (w_rows^j, vals[j])
        coef = [sum(vals[j] * pow(w_rows, (P - 1 - 0) * 0 + (rows - i) * j % rows, P)
                    for j in range(rows)) * inv_rows % P for i in range(rows)]
        for j in range(rows):  # the interpolant really interpolates
            assert sum(coef[i] * pow(w_rows, i * j, P) for i in range(rows)) % P == vals[j]
        for k in range(ext):
            x = 2717 * pow(w_ext, k, P) % P
            assert int(lde[k, c]) == sum(coef[i] * pow(x, i, P) for i in range(rows)) % P
    assert layers == oracle.merkle_tree(lde.astype("<u8").tobytes())


def test_commit_rejects(oracle):
    import pytest
    with pytest.raises(ValueError):
        oracle.commit(np.zeros((7, 2), dtype=np.uint64), 2)


def test_commit_rows(oracle):
    """The row-tree commit: the same extension as commit, the row tree over it."""
    import pytest
    for rows, cols, blowup in [(8, 3, 2), (16, 64, 4), (12, 21, 2), (1, 5, 1), (4, 1024, 2)]:
        m = oracle.fill_random(rows * cols, 91 + rows).reshape(rows, cols)
        lde, layers = oracle.commit_rows(m, blowup)
        lde5, _ = oracle.commit(m, blowup)
        assert np.array_equal(lde, lde5)
        assert layers == oracle.row_tree(lde)
    with pytest.raises(ValueError):
        oracle.commit_rows(np.zeros((7, 4), dtype=np.uint64), 2)
