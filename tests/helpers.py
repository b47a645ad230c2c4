"""Shared helpers of the test suite (inputs, metrics)."""
import numpy as np

from krans_amd import workloads as W


def csr_small_2x3():
    # tests/test_cyclic_lanczos.cpp:146-150: [[1,0,1],[0,1,1]]
    return W.CsrMatrix(2, 3, np.array([0, 2, 4], np.int32), np.array([0, 2, 1, 2], np.int32),
                       np.array([1.0, 1.0, 1.0, 1.0]))


def csr_identity(n):
    return W.CsrMatrix(n, n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32),
                       np.ones(n))


def csr_ones_2x2():
    # tests/test_cyclic_lanczos.cpp:434-437
    return W.CsrMatrix(2, 2, np.array([0, 2, 4], np.int32), np.array([0, 1, 0, 1], np.int32),
                       np.ones(4))


def scaled(A, s):
    return W.CsrMatrix(A.m, A.n, A.rowptr.copy(), A.col.copy(), A.val * s)


def residuals(A, U):
    S = A.to_scipy()
    return np.array([np.linalg.norm(S @ U[:, j]) for j in range(U.shape[1])])


def random_sparse(m, n, per_row, seed):
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(m):
        c = np.sort(rng.choice(n, size=min(per_row, n), replace=False))
        rows.append(np.full(c.size, i))
        cols.append(c)
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = rng.standard_normal(r.size)
    rp = np.zeros(m + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    return W.CsrMatrix(m, n, rp, c.astype(np.int32), v)
