"""Matrix-free device operators: the shifted-difference stencil
(kh_matrix_create_shifted_difference) and a user callback operator
(kh_matrix_create_operator), checked against the oracle's matrix-free
ShiftDiffOp (the reference's StdMatrixAdapter lambdas of
examples/quickstart.cpp:25-37) bit-exactly, and against the CSR path: equal
products give bitwise-equal solver results."""
import ctypes as C

import numpy as np
import pytest

import krans_amd as K
import oracle as O
from krans_amd import workloads as W
from conftest import principal_sine

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,N,c", [(1000, 3, 0.5), (100_003, 16, 0.9)])
def test_stencil_products_bit_exact(n, N, c):
    print(f"[case] stencil products n={n}", flush=True)
    S = K.HipShiftedDifferenceOperator(n, N, c)
    assert (S.m, S.n, S.nnz) == (n - N, n, 2 * (n - N))
    Op = O.Operator.shiftdiff(n, N, c)
    H = K.HipCsrMatrixAdapter.from_csr(W.shifted_difference(n, N, c))
    rng = np.random.default_rng(n)
    x, y = rng.standard_normal(n), rng.standard_normal(n - N)
    for got, ref in ((S.multiply(x), Op.mv(x)), (S.multiply_transpose(y), Op.mtv(y)),
                     (S.multiply(x), H.multiply(x)), (S.multiply_transpose(y),
                                                      H.multiply_transpose(y))):
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    assert S.fro == H.fro  # the same sum as the materialised CSR
    # FP32 products: the same rounding as the FP32 CSR path
    assert np.array_equal(S.multiply(x, "fp32"), H.multiply(x, "fp32"))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_quickstart_matrix_free_same_as_csr(precision):
    """Config C1 exactly as the reference runs it: matrix-free, one seed-42
    factory shared by cyclic (k=3) and then LOBPCG (nev=3)."""
    n, N, c = 1000, 3, 0.5
    print(f"[case] quickstart matrix-free {precision}", flush=True)
    Kx = W.shifted_difference_kernel(n, N, c)
    outs = []
    for A in (K.HipShiftedDifferenceOperator(n, N, c),
              K.HipCsrMatrixAdapter.from_csr(W.shifted_difference(n, N, c))):
        vf = K.HipVectorFactory(n, 42)
        rc = K.find_nullspace_cyclic_lanczos_ex(A, vf, K.CyclicLanczosOptions(k=3), precision)
        rl = K.find_nullspace_lobpcg_ex(A, vf, K.LOBPCGOptions(nev=3), precision)
        assert rc.right_null.shape[1] == 3 and rl.right_null.shape[1] == 3
        # acceptance guarantees ||A u|| <= thr = 1e-10 max(1, ||A||_F~), hence
        # a principal-angle sine of at most about thr / sigma_min, sigma_min =
        # 1 - c for the shifted difference
        thr = 1e-10 * max(1.0, rc.frob_norm_est)
        assert principal_sine(rc.right_null, Kx) <= 2 * thr / (1 - c)
        assert principal_sine(rl.right_null, Kx) <= 1e-8
        outs.append((rc.right_null, rl.right_null, rc.frob_norm_est))
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))
    if precision == "fp64":
        # the oracle's matrix-free run: same Hutchinson estimate (same stream)
        _, _, ro = O.cyclic(O.Operator.shiftdiff(n, N, c), O.Factory(42),
                            O.cyclic_defaults(k=3))
        assert abs(ro.frob_norm_est - outs[0][2]) <= 1e-12 * ro.frob_norm_est


def csr_callback(H):
    """A user operator whose device work is the CSR matrix's own product
    (kh_spmm_dev per column, honouring the leading dimensions)."""
    def apply(stream, precision, transpose, s, x, ldx, y, ldy):
        e = 8 if precision == "fp64" else 4
        p = 0 if precision == "fp64" else 1
        for j in range(s):
            K._check(K.lib().kh_spmm_dev(H.h, int(transpose), p, 1,
                                         C.c_void_p(x + j * ldx * e), C.c_void_p(y + j * ldy * e)))
    return K.HipOperatorAdapter(H.m, H.n, apply)


@pytest.mark.parametrize("solver,precision", [("cyclic", "fp64"), ("cyclic", "mixed"),
                                              ("lobpcg", "fp64"), ("lobpcg", "mixed")])
def test_callback_operator_same_as_csr(solver, precision):
    if solver == "cyclic":
        A = W.regular_incidence(4, 1024, seed=3)
        opt = K.CyclicLanczosOptions(k=4, tol_rel=1e-12, tol_abs=0.0, prescreen_factor=1e-3,
                                     max_restarts=200)
        run = K.find_nullspace_cyclic_lanczos_ex
    else:
        A, _ = W.rg_block(3000, p=4, offdiag=8, gnnz=16, seed=5)
        opt = K.LOBPCGOptions(nev=4, block_size=4, tol_rel=1e-12, tol_abs=0.0, max_iters=3000)
        run = K.find_nullspace_lobpcg_ex
    print(f"[case] callback operator {solver} {precision}", flush=True)
    H = K.HipCsrMatrixAdapter.from_csr(A)
    F = csr_callback(H)
    r1 = run(H, K.HipVectorFactory(A.n, 9), opt, precision)
    r2 = run(F, K.HipVectorFactory(A.n, 9), opt, precision)
    assert r1.right_null.shape[1] == 4
    if (solver, precision) == ("lobpcg", "mixed"):
        # the FP64 refinement of FP32 LOBPCG candidates is preconditioned with
        # diag(A A^T) when the operand provides it (a CSR matrix does, a
        # callback operator does not), so the two runs differ in rounding:
        # both must still meet the threshold and span the same null space
        thr = r1.accept_threshold
        S = A.to_scipy()
        for r in (r1, r2):
            assert r.right_null.shape[1] == 4
            assert np.linalg.norm(S @ r.right_null, axis=0).max() <= thr
        assert principal_sine(r1.right_null, r2.right_null) <= 1e-9
        return
    assert np.array_equal(r1.right_null.view(np.uint64), r2.right_null.view(np.uint64))


def test_callback_failure_is_an_error():
    def apply(stream, precision, transpose, s, x, ldx, y, ldy):
        raise RuntimeError("operator failed")
    F = K.HipOperatorAdapter(50, 40, apply)
    print("[case] callback failure", flush=True)
    with pytest.raises(K.KransError):
        K.find_nullspace_cyclic_lanczos_ex(F, K.HipVectorFactory(40, 1),
                                           K.CyclicLanczosOptions(k=1))
    with pytest.raises(ValueError):  # CSR-only entry points reject operators
        K._check(K.lib().kh_matrix_download_csr(F.h, None, None, None))
