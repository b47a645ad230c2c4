"""The CPU oracle pinned against the reference's own published results,
known-answer values, closed forms and LAPACK (numpy) on small inputs."""
import numpy as np
import pytest

import oracle as O
from krans_amd import workloads as W
from conftest import principal_sine
from helpers import csr_identity, csr_ones_2x2, csr_small_2x3, random_sparse


def test_mt19937_default_seed_first_output():
    # C++ standard [rand.predef]: the 10000th output of mt19937 is 4123659995;
    # the first output for the default seed 5489 is 3499211612.
    assert O.lib().orc_mt19937_first(5489) == 3499211612


def test_normal_stream_moments_and_continuity():
    f = O.Factory(42)
    a = f.randn(100001)
    assert abs(a.mean()) < 0.01 and abs(a.std() - 1) < 0.01
    # one stream: two draws of 5 equal one draw of 10 (the cached polar value
    # persists between calls, StdVectorFactory.hpp:55-57)
    f1, f2 = O.Factory(7), O.Factory(7)
    x = np.concatenate([f1.randn(5), f1.randn(5)])
    assert np.array_equal(x, f2.randn(10))


def test_quickstart_published_results():
    """README.md:76-80 for n=1000, N=3, c=0.5, StdVectorFactory(1000, 42),
    cyclic then LOBPCG on one factory (examples/quickstart.cpp:39-85)."""
    n, N, c = 1000, 3, 0.5
    A = O.Operator.shiftdiff(n, N, c)
    Kx = W.shifted_difference_kernel(n, N, c)
    F = O.Factory(42)
    U, _, r = O.cyclic(A, F, O.cyclic_defaults(k=3))
    assert U.shape[1] == 3
    res = max(np.linalg.norm(A.mv(U[:, j])) for j in range(3))
    dist = max(np.linalg.norm(U[:, j] - Kx @ (Kx.T @ U[:, j])) for j in range(3))
    # published: 6.4e-16 and 7.6e-16
    assert res < 5e-15 and dist < 5e-15
    U2, th, rs, r2 = O.lobpcg(A, F, O.lobpcg_defaults(nev=3))
    assert U2.shape[1] == 3
    res2 = max(np.linalg.norm(A.mv(U2[:, j])) for j in range(3))
    dist2 = max(np.linalg.norm(U2[:, j] - Kx @ (Kx.T @ U2[:, j])) for j in range(3))
    # published: 3.5e-11 (= 0.01 * thr, LOBPCGSolver.cpp:808-811) and 6.4e-11
    assert 1e-11 < res2 <= 0.01 * r2.accept_threshold * 1.0001
    assert 1e-11 < dist2 < 2e-10


def test_stencil_equals_csr_bitwise():
    n, N, c = 1000, 3, 0.5
    A = W.shifted_difference(n, N, c)
    S = O.Operator.shiftdiff(n, N, c)
    C = O.Operator.csr(A.m, A.n, A.rowptr, A.col, A.val)
    x = np.random.default_rng(0).standard_normal(n)
    y = np.random.default_rng(1).standard_normal(n - N)
    assert np.array_equal(np.abs(S.mv(x)), np.abs(C.mv(x)))
    assert np.array_equal(S.mv(x), C.mv(x)) or np.allclose(S.mv(x), C.mv(x), rtol=0, atol=0)
    assert np.array_equal(np.abs(S.mtv(y)), np.abs(C.mtv(y)))


def test_sparse_adapter_known_values():
    # tests/test_sparse_adapter.cpp:33-45: A = [[1,0,2],[0,3,4]]... values
    # Ax = [3,23] for x = [1,2,... ] ; reproduce with the same matrix
    A = W.CsrMatrix(2, 3, np.array([0, 2, 4], np.int32), np.array([0, 2, 1, 2], np.int32),
                    np.array([1.0, 2.0, 3.0, 4.0]))
    op = O.Operator.csr(A.m, A.n, A.rowptr, A.col, A.val)
    x = np.array([1.0, 1.0, 1.0])
    assert np.array_equal(op.mv(x), A.to_scipy() @ x)
    # adjoint identity <Ax, y> = <x, A^T y>
    B = random_sparse(50, 40, 5, seed=9)
    ob = O.Operator.csr(B.m, B.n, B.rowptr, B.col, B.val)
    xx = np.random.default_rng(3).standard_normal(40)
    yy = np.random.default_rng(4).standard_normal(50)
    assert abs(ob.mv(xx) @ yy - xx @ ob.mtv(yy)) <= 1e-12 * np.linalg.norm(xx) * np.linalg.norm(yy) * 10


@pytest.mark.parametrize("shape", [(7, 6), (49, 48), (67, 66)])
def test_svd_matches_lapack(shape):
    rng = np.random.default_rng(shape[0])
    A = rng.standard_normal(shape)
    sv, V = O.svd(A)
    ref = np.linalg.svd(A, compute_uv=False)
    assert np.allclose(sv, ref, rtol=1e-13, atol=1e-13 * ref[0])
    # right singular vectors: A V = U S with orthonormal V
    assert np.abs(V.T @ V - np.eye(shape[1])).max() < 1e-13
    AV = A @ V
    assert np.allclose(np.linalg.norm(AV, axis=0), sv, rtol=1e-12, atol=1e-13)


def test_svd_refined_tbar_smallest_vector():
    # T-bar of a Golub-Kahan run with a zero singular value (odd size: the
    # zero-diagonal tridiagonal T is singular): the smallest right singular
    # vector is unique and must match LAPACK's
    l = 11
    beta = np.linspace(1.0, 2.0, l - 1)
    T = np.diag(beta, 1) + np.diag(beta, -1)
    Tb = np.vstack([T, np.zeros((1, l))])
    Tb[l, l - 1] = 1e-20
    sv, V = O.svd(Tb)
    _, s, vt = np.linalg.svd(Tb)
    assert abs(sv[-1] - s[-1]) < 1e-14
    assert min(np.linalg.norm(V[:, -1] - vt[-1]), np.linalg.norm(V[:, -1] + vt[-1])) < 1e-12


@pytest.mark.parametrize("n", [2, 9, 32, 96])
def test_sym_eig_matches_lapack(n):
    rng = np.random.default_rng(n)
    B = rng.standard_normal((n, n))
    A = B + B.T
    w, Q = O.sym_eig(A)
    ref = np.linalg.eigvalsh(A)
    assert np.allclose(w, ref, atol=1e-12 * np.abs(ref).max())
    assert np.abs(Q.T @ Q - np.eye(n)).max() < 1e-12
    assert np.abs(A @ Q - Q * w).max() < 1e-11 * np.abs(ref).max()


def test_cholesky_and_failure():
    rng = np.random.default_rng(0)
    B = rng.standard_normal((20, 20))
    S = B @ B.T + 20 * np.eye(20)
    L = O.cholesky(S)
    assert np.allclose(L, np.linalg.cholesky(S), rtol=1e-13, atol=1e-13)
    assert O.cholesky(-np.eye(3)) is None


def test_gram_schmidt_known_values():
    # tests/test_gram_schmidt.cpp:70-106: [3,2] normalises by sqrt(13)
    ok, Q = O.mgs(np.array([[3.0], [2.0]]))
    assert ok and np.allclose(Q[:, 0], np.array([3.0, 2.0]) / np.sqrt(13))
    ok, _ = O.mgs(np.array([[1.0, 1.0], [1.0, 1.0 + 1e-15]]), tol=1e-12)
    assert not ok
    r, v = O.orth_against(np.array([[1.0], [0.0]]), np.array([1.0, 1.0]))
    assert abs(r - 1.0) < 1e-15 and np.allclose(v, [0.0, 1.0])


def test_subspace_diff_values():
    # tests/test_subspace_compare.cpp: identical spans -> 0, different -> > 0
    B = np.array([[1.0, 0.0], [0.0, 1.0], [0.0, 0.0]])
    R = B @ np.array([[0.6, -0.8], [0.8, 0.6]])
    assert O.subspace_diff(B, R) < 1e-12
    C = np.array([[1.0, 0.0], [0.0, 0.0], [0.0, 1.0]])
    assert O.subspace_diff(B, C) > 0.5
    assert O.subspace_diff(B, B[:, :1]) == 1e9


def test_oracle_scenarios():
    # tests/test_cyclic_lanczos.cpp scenarios, on the oracle
    A = csr_small_2x3()
    U, _, _ = O.cyclic(O.Operator.csr(A.m, A.n, A.rowptr, A.col, A.val), O.Factory(1),
                       O.cyclic_defaults(k=1, ncv=20, max_restarts=50, nullity_known=1,
                                         rank_known=2, cg_maxit=40))
    e = np.array([-1.0, -1.0, 1.0]) / np.sqrt(3)
    assert U.shape[1] == 1 and principal_sine(U, e[:, None]) < 1e-10
    I = csr_identity(3)
    U, _, _ = O.cyclic(O.Operator.csr(3, 3, I.rowptr, I.col, I.val), O.Factory(1),
                       O.cyclic_defaults(k=1, nullity_known=0, rank_known=3))
    assert U.shape[1] == 0
    J = csr_ones_2x2()
    U, L, _ = O.cyclic(O.Operator.csr(2, 2, J.rowptr, J.col, J.val), O.Factory(1),
                       O.cyclic_defaults(k=1, nullity_known=1, rank_known=1,
                                         return_left_space=1))
    e2 = np.array([-1.0, 1.0]) / np.sqrt(2)
    assert principal_sine(U, e2[:, None]) < 1e-10 and principal_sine(L, e2[:, None]) < 1e-10


def test_oracle_incidence_exact_kernel():
    g, v = 4, 256
    A = W.regular_incidence(g, v, seed=5)
    U, _, r = O.cyclic(O.Operator.csr(A.m, A.n, A.rowptr, A.col, A.val), O.Factory(42),
                       O.cyclic_defaults(k=g, tol_rel=1e-12, tol_abs=0.0, prescreen_factor=1e-3))
    assert U.shape[1] == g
    assert principal_sine(U, W.component_indicators(g, v)) < 5e-9
    # dense SVD brute force agrees
    S = A.to_scipy().toarray()
    _, s, vt = np.linalg.svd(S)
    N = vt[-g:].T
    assert s[-g] < 1e-12 and s[-g - 1] > 0.1
    assert principal_sine(U, N) < 5e-9


def test_oracle_lobpcg_rg_kernel():
    A, G = W.rg_block(400, p=4, offdiag=4, gnnz=4, seed=3)
    U, th, rs, r = O.lobpcg(O.Operator.csr(A.m, A.n, A.rowptr, A.col, A.val), O.Factory(42),
                            O.lobpcg_defaults(nev=4, tol_rel=1e-12, tol_abs=0.0))
    assert U.shape[1] == 4
    assert principal_sine(U, W.rg_kernel(G)) < 1e-8
