"""GPU parity of the LOBPCG hot path (block products, Rayleigh-Ritz, final
polish) against the oracle (restatement of src/Solvers/LOBPCGSolver.cpp) and
exact kernels."""
import numpy as np
import pytest

import oracle as O
import krans_amd as K
from krans_amd import workloads as W
from conftest import principal_sine
from helpers import csr_identity, csr_small_2x3, residuals

pytestmark = pytest.mark.gpu


def hip(A):
    return K.HipCsrMatrixAdapter.from_csr(A)


def test_quickstart_lobpcg_as_csr():
    """README.md:79: 3 vectors, max||Au|| = 0.01 * thr (3.5e-11), dist 6.4e-11.
    The factory is the quickstart's: cyclic runs first on the same stream."""
    n, N, c = 1000, 3, 0.5
    A = W.shifted_difference(n, N, c)
    Kx = W.shifted_difference_kernel(n, N, c)
    print("[case] quickstart lobpcg fp64", flush=True)
    H = hip(A)
    vf = K.HipVectorFactory(n, 42)
    K.find_nullspace_cyclic_lanczos(H, vf, K.CyclicLanczosOptions(k=3))
    r = K.find_nullspace_lobpcg_ex(H, vf, K.LOBPCGOptions(nev=3))
    U = r.right_null
    assert U.shape[1] == 3 and r.converged_to_target
    target = max(1e-14, min(1e-10, 0.01 * r.accept_threshold))
    res = residuals(A, U)
    assert res.max() <= target * 1.01, (res, target)
    assert principal_sine(U, Kx) <= 1e-9
    F = O.Factory(42)
    op = O.Operator.shiftdiff(n, N, c)
    O.cyclic(op, F, O.cyclic_defaults(k=3))
    Uo, th, rs, ro = O.lobpcg(op, F, O.lobpcg_defaults(nev=3))
    assert abs(r.frob_norm_est - ro.frob_norm_est) <= 1e-12 * ro.frob_norm_est
    assert principal_sine(U, Uo) <= 1e-9
    print("gpu iters", r.iters, "oracle iters", ro.iters)


def test_rg_block_kernel_fp64():
    m, p = 3000, 8
    A, G = W.rg_block(m, p=p, offdiag=8, gnnz=16, seed=11)
    Kx = W.rg_kernel(G)
    print("[case] rg block", flush=True)
    r = K.find_nullspace_lobpcg_ex(hip(A), K.HipVectorFactory(A.n, 42),
                                   K.LOBPCGOptions(nev=p, block_size=p, tol_rel=1e-12,
                                                   tol_abs=0.0, max_iters=3000))
    U = r.right_null
    assert U.shape[1] == p
    assert residuals(A, U).max() <= r.accept_threshold
    assert principal_sine(U, Kx) <= 1e-10  # tol_rel 1e-12
    assert np.abs(U.T @ U - np.eye(p)).max() <= 1e-10
    assert r.cholqr_blocks == 0  # FP64 blocks keep the reference's MGS


def test_lobpcg_small_analytic_and_identity():
    A = csr_small_2x3()
    print("[case] lobpcg 2x3", flush=True)
    U = K.find_nullspace_lobpcg(hip(A), K.HipVectorFactory(3, 1),
                                K.LOBPCGOptions(nev=1, nullity_known=1, rank_known=2))
    e = np.array([-1.0, -1.0, 1.0]) / np.sqrt(3)
    assert U.shape[1] == 1 and principal_sine(U, e[:, None]) <= 1e-9
    I = csr_identity(4)
    U = K.find_nullspace_lobpcg(hip(I), K.HipVectorFactory(4, 1),
                                K.LOBPCGOptions(nev=1, nullity_known=0, rank_known=4))
    assert U.shape[1] == 0


def test_lobpcg_inconsistent_hints_raise():
    A = csr_small_2x3()
    with pytest.raises(ValueError):
        K.find_nullspace_lobpcg(hip(A), K.HipVectorFactory(3, 1),
                                K.LOBPCGOptions(nev=1, nullity_known=2, rank_known=2))
    with pytest.raises(ValueError):
        K.find_nullspace_lobpcg(hip(A), K.HipVectorFactory(3, 1),
                                K.LOBPCGOptions(nev=1, nullity_known=5))


def test_cxx_quickstart_drop_in():
    """The reference's quick start, compiled against the unchanged KraNS C++
    API with the HIP adapters swapped in (tests/cxx/quickstart_hip.cpp)."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(K.LIB_PATH), "quickstart_hip")
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    print("[case] c++ quickstart", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "cyclic   3" in r.stdout and "lobpcg   3" in r.stdout


@pytest.mark.parametrize("case", ["rg", "shiftdiff"])
def test_lobpcg_mixed_precision(case):
    """FP32 blocks with FP32-MFMA Gram products; candidates are refined by an
    FP64 projector pass before the reference's FP64 acceptance test."""
    if case == "rg":
        A, G = W.rg_block(3000, p=8, offdiag=8, gnnz=16, seed=11)
        Kx, nev = W.rg_kernel(G), 8
    else:
        A = W.shifted_difference(2000, 3, 0.5)
        Kx, nev = W.shifted_difference_kernel(2000, 3, 0.5), 3
    print(f"[case] lobpcg mixed {case}", flush=True)
    r = K.find_nullspace_lobpcg_ex(hip(A), K.HipVectorFactory(A.n, 42),
                                   K.LOBPCGOptions(nev=nev, block_size=nev, tol_rel=1e-12,
                                                   tol_abs=0.0, max_iters=3000),
                                   precision="mixed")
    U = r.right_null
    assert U.shape[1] == nev and r.converged_to_target
    assert residuals(A, U).max() <= r.accept_threshold
    assert principal_sine(U, Kx) <= 1e-10  # tol_rel 1e-12
    assert np.abs(U.T @ U - np.eye(nev)).max() <= 1e-10
    print("mixed iters", r.iters, "cholesky-qr blocks", r.cholqr_blocks, "fallbacks",
          r.cholqr_fallbacks)
    if nev >= 4:  # FP32 blocks of 4 or more columns take Cholesky QR
        assert r.cholqr_blocks > 0 and r.cholqr_fallbacks <= r.cholqr_blocks
