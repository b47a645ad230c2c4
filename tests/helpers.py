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


# Two projector passes at cg_maxit 2 on the incidence of two 64-vertex
# graphs: 5 CG solves, all at the limit (the count per pass of
# CyclicLanczosSolver.cpp:522-526; one per candidate would give 4).
TWO_PASS_OPTS = dict(k=2, projector_passes=2, cg_maxit=2, stagnation_restarts=4,
                     max_restarts=4, tol_rel=1e-12, tol_abs=0.0, grad_max_iters=10)


def two_pass_case():
    return W.regular_incidence(2, 64, seed=3)


def gradient_steps_closed_form(B, s, v0, thr, max_steps=100000):
    """Steps of v <- (v - tau A^T A v)/||.|| until ||A v|| <= thr, with
    A = s B and tau = 0.9 / max(1e-30, ||A||_F~^2) = 0.9e30 (the estimate is
    below 1e-30 here): v_t is proportional to (I - tau A^T A)^t v0, evaluated
    in the eigenbasis of B^T B."""
    lam, Q = np.linalg.eigh(B.T @ B)
    c = Q.T @ v0
    f = 1.0 - 0.9e30 * s * s * lam
    for t in range(max_steps + 1):
        ct = c * f ** t
        v = Q @ (ct / np.linalg.norm(ct))
        if s * np.linalg.norm(B @ v) <= thr:
            return t
    return None
