"""Null spaces of the reference's data matrices (tests/golden/mat{1,3,6}.ija,
copies of data/mat1.ija, mat3.ija, mat6.ija) loaded through kh_load_ija and
solved on the GPU with the options of the reference's own tests
(tests/test_cyclic_lanczos.cpp:192-213 for mat1; tests/test_all_solvers.cpp:
125-140 for mat3/mat6): the count of the known nullity, residuals under the
threshold, principal-angle sine to the dense-SVD null space, and RREF
subspace distance to the oracle's basis on the same input."""
import os

import numpy as np
import pytest

import krans_amd as K
import oracle
from conftest import principal_sine

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def dense_null(A, k):
    """Brute force: the k right singular vectors of the smallest singular
    values (numpy LAPACK SVD of the dense matrix)."""
    D = A.to_scipy().toarray()
    _, s, Vt = np.linalg.svd(D)
    return Vt[-k:].T, s


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_mat1_cyclic_reference_options(precision):
    A = K.load_ija(os.path.join(GOLDEN, "mat1.ija"))
    print(f"[case] mat1 cyclic {precision}", flush=True)
    kw = dict(k=5, ncv=60, max_restarts=200, tol_abs=1e-12, tol_rel=1e-10,
              nullity_known=5, rank_known=100, cg_maxit=60, cg_rel_tol=1e-8)
    r = K.find_nullspace_cyclic_lanczos_ex(K.HipCsrMatrixAdapter.from_csr(A),
                                           K.HipVectorFactory(A.n, 42),
                                           K.CyclicLanczosOptions(**kw), precision)
    U = r.right_null
    assert U.shape[1] == 5
    thr = max(1e-12, 1e-10 * max(1.0, r.frob_norm_est))
    assert np.linalg.norm(A.to_scipy() @ U, axis=0).max() <= thr
    N, s = dense_null(A, 5)
    assert s[-6] > 0.5 and s[-5] < 1e-11  # nullity 5 (rank 100)
    # residual <= thr and a singular-value gap of 1 bound the angle by ~thr
    assert principal_sine(U, N) <= 10 * thr
    op = oracle.Operator.csr(A.m, A.n, A.rowptr, A.col, A.val)
    Uo, _, ro = oracle.cyclic(op, oracle.Factory(42), oracle.cyclic_defaults(**kw))
    assert Uo.shape[1] == 5
    assert oracle.subspace_diff(U, Uo, 1e-7) <= 1e-5  # the reference test's bound


@pytest.mark.parametrize("name,k", [("mat3", 10), ("mat4", 15), ("mat5", 20), ("mat6", 25)])
@pytest.mark.parametrize("solver", ["cyclic", "lobpcg"])
def test_all_solvers_matrices(name, k, solver):
    A = K.load_ija(os.path.join(GOLDEN, name + ".ija"))
    print(f"[case] {name} {solver}", flush=True)
    H = K.HipCsrMatrixAdapter.from_csr(A)
    if solver == "cyclic":
        o = K.CyclicLanczosOptions(k=k, tol_rel=1e-8, max_restarts=200, ncv=max(60, k + 10))
        r = K.find_nullspace_cyclic_lanczos_ex(H, K.HipVectorFactory(A.n, 42), o)
    else:
        o = K.LOBPCGOptions(nev=k, tol_rel=1e-8, block_size=min(min(A.n, 64), max(16, k)),
                            max_iters=1000)
        r = K.find_nullspace_lobpcg_ex(H, K.HipVectorFactory(A.n, 42), o)
    U = r.right_null
    N, s = dense_null(A, k)
    # the stated nullity: k singular values below 1e-8, a gap above 1
    assert s[-k - 1] > 1.0 and s[-k] < 1e-8
    assert U.shape[1] == k
    thr = (r.accept_threshold if solver == "lobpcg"
           else max(o.tol_abs, o.tol_rel * max(1.0, r.frob_norm_est)))
    assert np.linalg.norm(A.to_scipy() @ U, axis=0).max() <= thr
    assert principal_sine(U, N) <= 10 * thr / s[-k - 1] + 1e-9
