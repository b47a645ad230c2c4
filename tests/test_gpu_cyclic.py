"""GPU parity of the cyclic Lanczos hot path, through the C ABI.

Checked against the CPU oracle (oracle/, an FP64 restatement of
src/Solvers/CyclicLanczosSolver.cpp) on the same seeded inputs, and against
the exact kernels of the generators. Tolerances are stated per test.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import krans_amd as K
from krans_amd import workloads as W
from conftest import principal_sine
from helpers import csr_identity, csr_ones_2x2, csr_small_2x3, residuals, scaled, random_sparse

pytestmark = pytest.mark.gpu


def hip(A):
    return K.HipCsrMatrixAdapter.from_csr(A)


def ora(A):
    return O.Operator.csr(A.m, A.n, A.rowptr, A.col, A.val)


# ------------------------------------------------------------ structure ----
@pytest.mark.parametrize("case", ["incidence", "random", "shiftdiff"])
def test_csr_csc_twin_bit_exact(case):
    if case == "incidence":
        A = W.regular_incidence(4, 512, seed=7)
    elif case == "random":
        A = random_sparse(300, 200, 7, seed=3)
    else:
        A = W.shifted_difference(1000, 3, 0.5)
    print(f"[case] csc twin {case} m={A.m} n={A.n} nnz={A.nnz}", flush=True)
    H = hip(A)
    rp, ci, va = H.download_csr()
    assert np.array_equal(rp, A.rowptr) and np.array_equal(ci, A.col)
    assert np.array_equal(va.view(np.uint64), A.val.view(np.uint64))
    cp, ri, vt, perm = H.download_csc()
    ocp, ori, ovt, operm = ora(A).csc(A.nnz)
    assert np.array_equal(cp, ocp) and np.array_equal(ri, ori)
    assert np.array_equal(perm, operm)
    assert np.array_equal(vt.view(np.uint64), ovt.view(np.uint64))
    # and against scipy's transpose (rows ascend within each column)
    T = A.to_scipy().T.tocsr()
    T.sort_indices()
    assert np.array_equal(cp, T.indptr) and np.array_equal(ri, T.indices)
    assert np.array_equal(vt, T.data)
    assert np.array_equal(A.val[perm], vt)


# ----------------------------------------------------------------- SpMV ----
@pytest.mark.parametrize("case", ["incidence", "random", "shiftdiff"])
def test_spmv_fp64_bit_exact(case):
    if case == "incidence":
        A = W.regular_incidence(4, 512, seed=7)
    elif case == "random":
        A = random_sparse(300, 200, 7, seed=3)
    else:
        A = W.shifted_difference(1000, 3, 0.5)
    print(f"[case] spmv {case}", flush=True)
    H, Op = hip(A), ora(A)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(A.n)
    y = rng.standard_normal(A.m)
    # one lane per row: the reference's sequential ascending order, no FMA
    assert np.array_equal(H.multiply(x).view(np.uint64), Op.mv(x).view(np.uint64))
    assert np.array_equal(H.multiply_transpose(y).view(np.uint64), Op.mtv(y).view(np.uint64))
    # FP32 operator: within FP32 rounding of the row sums
    S = A.to_scipy()
    bound = 4 * np.finfo(np.float32).eps * (abs(S) @ np.abs(x) + 1e-30)
    assert np.all(np.abs(H.multiply(x, "fp32") - S @ x) <= 8 * bound + 1e-30)


def test_spmv_long_rows_tolerance():
    # rows longer than the one-lane threshold use lane groups: not the
    # sequential order, so compare within 1e-14 relative to sum |a||x|
    A = random_sparse(64, 4000, 300, seed=5)
    print("[case] spmv long rows", flush=True)
    H, Op = hip(A), ora(A)
    x = np.random.default_rng(2).standard_normal(A.n)
    ref = Op.mv(x)
    scale = abs(A.to_scipy()) @ np.abs(x)
    assert np.all(np.abs(H.multiply(x) - ref) <= 1e-14 * scale)


# ----------------------------------------------------- solver scenarios ----
def test_quickstart_as_csr_fp64_matches_oracle():
    """Config C1 with the stencil materialised as CSR (bit-identical products,
    see the stencil/CSR test); same seed 42 factory, k=3."""
    n, N, c = 1000, 3, 0.5
    A = W.shifted_difference(n, N, c)
    Kx = W.shifted_difference_kernel(n, N, c)
    print("[case] quickstart cyclic fp64", flush=True)
    H = hip(A)
    vf = K.HipVectorFactory(n, 42)
    r = K.find_nullspace_cyclic_lanczos_ex(H, vf, K.CyclicLanczosOptions(k=3))
    U = r.right_null
    assert U.shape[1] == 3 and r.converged_to_target
    thr = max(1e-12, 1e-10 * max(1.0, r.frob_norm_est))
    assert residuals(A, U).max() <= thr
    # README.md:76-80: max dist(u, ker A) 7.6e-16 -- allow the 1e-8 of
    # examples/quickstart.cpp:65
    assert principal_sine(U, Kx) <= 1e-8
    assert np.abs(U.T @ U - np.eye(3)).max() <= 1e-12
    Uo, _, ro = O.cyclic(O.Operator.shiftdiff(n, N, c), O.Factory(42), O.cyclic_defaults(k=3))
    # same random stream, same norm estimate to rounding
    assert abs(r.frob_norm_est - ro.frob_norm_est) <= 1e-12 * ro.frob_norm_est
    assert O.subspace_diff(U, Uo, 1e-12) <= 1e-8
    print("gpu", r.restarts, r.krylov_steps_total, r.cg_solves, "oracle", ro.restarts,
          ro.krylov_steps_total, ro.cg_solves)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_incidence_small_matches_exact_kernel(precision):
    g, v, k = 8, 1024, 8
    A = W.regular_incidence(g, v, seed=42)
    Kx = W.component_indicators(g, v)
    print(f"[case] incidence {g}x{v} {precision}", flush=True)
    H = hip(A)
    vf = K.HipVectorFactory(A.n, 42)
    opt = K.CyclicLanczosOptions(k=k, tol_rel=1e-12, tol_abs=0.0, prescreen_factor=1e-3)
    r = K.find_nullspace_cyclic_lanczos_ex(H, vf, opt, precision=precision)
    U = r.right_null
    assert U.shape[1] == k and r.converged_to_target
    thr = 1e-12 * r.frob_norm_est
    res = residuals(A, U)
    assert res.max() <= thr, (res.max(), thr)
    # the north-star bar at tol_rel 1e-12: largest principal-angle sine
    # <= 1e-10 to the exact kernel and to the oracle's basis
    assert principal_sine(U, Kx) <= 1e-10
    assert np.abs(U.T @ U - np.eye(k)).max() <= 1e-10
    Uo, _, ro = O.cyclic(ora(A), O.Factory(42), O.cyclic_defaults(
        k=k, tol_rel=1e-12, tol_abs=0.0, prescreen_factor=1e-3))
    assert Uo.shape[1] == k
    assert principal_sine(U, Uo) <= 1e-10
    print(precision, "gpu restarts", r.restarts, "cg", r.cg_solves, "| oracle", ro.restarts,
          ro.cg_solves)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_incidence_k24_wide_window_products(precision):
    """k = 24: the window-against-basis products grow past 16 columns and
    take the 64-column MFMA tiles (96-row FP32 / 48-row FP64 slabs), in the
    FP32 screen and in the FP64 evaluation."""
    g, v, k = 24, 512, 24
    A = W.regular_incidence(g, v, seed=3)
    Kx = W.component_indicators(g, v)
    print(f"[case] incidence k=24 {precision}", flush=True)
    opt = K.CyclicLanczosOptions(k=k, tol_rel=1e-12, tol_abs=0.0, prescreen_factor=1e-3,
                                 max_restarts=400)
    r = K.find_nullspace_cyclic_lanczos_ex(hip(A), K.HipVectorFactory(A.n, 42), opt, precision)
    U = r.right_null
    assert U.shape[1] == k and r.converged_to_target
    assert residuals(A, U).max() <= 1e-12 * r.frob_norm_est
    assert principal_sine(U, Kx) <= 1e-10
    assert np.abs(U.T @ U - np.eye(k)).max() <= 1e-10
    print(precision, "restarts", r.restarts, "screened", r.screened)


@pytest.mark.parametrize("operand", ["csr", "stencil"])
@pytest.mark.parametrize("precision,second_pass", [("fp64", 0), ("mixed", 0), ("mixed", 1),
                                                   ("fp64", 2)])
def test_selective_second_pass(operand, precision, second_pass):
    """The Golub-Kahan steps' second Gram-Schmidt pass: automatic (0) is the
    reference's double reorthogonalisation for FP64 (every pair runs both
    passes) and the selective DGKS pass for mixed precision (a pair runs its
    second pass only where the first cancelled; after the three-term
    recurrence hardly any do). Every setting returns the exact kernel at
    tol_rel 1e-12 (principal-angle sine <= 1e-10, residual within the
    threshold)."""
    print(f"[case] second pass {operand} {precision} {second_pass}", flush=True)
    if operand == "csr":
        g, v, k = 8, 1024, 8
        A = W.regular_incidence(g, v, seed=42)
        H, Kx, n = hip(A), W.component_indicators(g, v), A.n
    else:
        n, N, c, k = 4096, 8, 0.5, 8  # nonzero singular values in [0.5, 1.5]
        A = W.shifted_difference(n, N, c)
        H, Kx = K.HipShiftedDifferenceOperator(n, N, c), W.shifted_difference_kernel(n, N, c)
    opt = K.CyclicLanczosOptions(k=k, tol_rel=1e-12, tol_abs=0.0, prescreen_factor=1e-3,
                                 max_restarts=400, second_pass=second_pass)
    r = K.find_nullspace_cyclic_lanczos_ex(H, K.HipVectorFactory(n, 42), opt, precision)
    U = r.right_null
    assert U.shape[1] == k and r.converged_to_target
    assert residuals(A, U).max() <= 1e-12 * r.frob_norm_est
    assert principal_sine(U, Kx) <= 1e-10
    assert r.cgs_pairs > 0
    selective = second_pass == 2 or (second_pass == 0 and precision == "mixed")
    if selective:
        assert r.second_passes < r.cgs_pairs / 4, (r.second_passes, r.cgs_pairs)
        assert r.second_passes == r.gk_requeued  # each one forced by one requeue
    else:
        assert r.second_passes == r.cgs_pairs and r.gk_requeued == 0
    print(precision, second_pass, "restarts", r.restarts, "pairs", r.cgs_pairs, "second passes",
          r.second_passes, "requeued", r.gk_requeued)


@pytest.mark.parametrize("graphs", [-1, 1])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_selective_requeue_equals_both_passes(precision, graphs):
    """The selective form queues every step with one pass and queues the
    cycle again from the first step whose DGKS test asks for the second
    pass, with both passes there. With eta = 1 the test asks at every step,
    so every step is requeued once and the cycle ends with both passes at
    every step: the solve must be bitwise the two-pass solve (also with
    cycles replayed from captured HIP graphs, where the requeued steps run
    as direct launches)."""
    g, v, k = 4, 512, 4
    A = W.regular_incidence(g, v, seed=5)
    print(f"[case] requeue {precision} graphs={graphs}", flush=True)
    base = dict(k=k, ncv=10, tol_rel=1e-12, tol_abs=0.0, prescreen_factor=1e-3,
                max_restarts=400, graphs=graphs)
    r1 = K.find_nullspace_cyclic_lanczos_ex(hip(A), K.HipVectorFactory(A.n, 42),
                                            K.CyclicLanczosOptions(second_pass=1, **base),
                                            precision)
    r2 = K.find_nullspace_cyclic_lanczos_ex(
        hip(A), K.HipVectorFactory(A.n, 42),
        K.CyclicLanczosOptions(second_pass=2, second_pass_eta=1.0, **base), precision)
    assert r1.right_null.shape[1] == k and r1.restarts == r2.restarts
    assert np.array_equal(r1.right_null, r2.right_null)
    # every step up to each cycle's stop was requeued once (steps queued
    # after a stop are ignored and keep their single pass)
    assert r2.cgs_pairs == r1.cgs_pairs == r1.second_passes
    assert r2.second_passes == r2.gk_requeued > r2.cgs_pairs / 2
    assert principal_sine(r2.right_null, W.component_indicators(g, v)) <= 1e-10
    if graphs == 1:
        assert r2.graph_replays > 0


def test_small_2x3_analytic():
    # tests/test_cyclic_lanczos.cpp:144-190
    A = csr_small_2x3()
    print("[case] 2x3", flush=True)
    opt = K.CyclicLanczosOptions(k=1, ncv=20, max_restarts=50, nullity_known=1, rank_known=2,
                                 cg_maxit=40)
    U = K.find_nullspace_cyclic_lanczos(hip(A), K.HipVectorFactory(3, 1), opt)
    assert U.shape[1] == 1
    assert abs(np.linalg.norm(U[:, 0]) - 1) <= 1e-12
    assert residuals(A, U)[0] <= 1e-10
    e = np.array([-1.0, -1.0, 1.0]) / np.sqrt(3)
    assert principal_sine(U, e[:, None]) <= 1e-10


def test_identity_has_no_null_space():
    # tests/test_cyclic_lanczos.cpp:268-300
    A = csr_identity(3)
    print("[case] identity", flush=True)
    opt = K.CyclicLanczosOptions(k=1, ncv=20, max_restarts=30, nullity_known=0, rank_known=3)
    U = K.find_nullspace_cyclic_lanczos(hip(A), K.HipVectorFactory(3, 1), opt)
    assert U.shape[1] == 0


def test_rank1_left_and_right():
    # tests/test_cyclic_lanczos.cpp:430-488
    A = csr_ones_2x2()
    print("[case] 2x2 rank-1", flush=True)
    opt = K.CyclicLanczosOptions(k=1, ncv=20, max_restarts=50, nullity_known=1, rank_known=1,
                                 return_left_space=True)
    r = K.find_nullspace_cyclic_lanczos_ex(hip(A), K.HipVectorFactory(2, 1), opt)
    e = np.array([-1.0, 1.0]) / np.sqrt(2)
    assert r.right_null.shape[1] == 1 and r.left_null.shape[1] == 1
    assert principal_sine(r.right_null, e[:, None]) <= 1e-10
    assert principal_sine(r.left_null, e[:, None]) <= 1e-10


def test_scale_invariance_1e6():
    # tests/test_cyclic_lanczos.cpp:490-542: the same subspace for A and 1e6 A
    A = W.regular_incidence(2, 256, seed=3)
    print("[case] scale 1e6", flush=True)
    opt = K.CyclicLanczosOptions(k=2, nullity_known=2)
    U1 = K.find_nullspace_cyclic_lanczos(hip(A), K.HipVectorFactory(A.n, 1), opt)
    U2 = K.find_nullspace_cyclic_lanczos(hip(scaled(A, 1e6)), K.HipVectorFactory(A.n, 1), opt)
    assert U1.shape[1] == 2 and U2.shape[1] == 2
    assert principal_sine(U1, U2) <= 1e-8


def test_backend_mismatch_raises():
    A = csr_small_2x3()
    with pytest.raises(ValueError):
        K.find_nullspace_cyclic_lanczos(object(), K.HipVectorFactory(3, 1),
                                        K.CyclicLanczosOptions())
    with pytest.raises(ValueError):
        hip(A).multiply(np.zeros(2))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_run_twice_bitwise_identical(precision):
    """No atomics and fixed reduction trees: two solves with the same seed
    return bit-identical bases (a race / nondeterminism detector)."""
    A = W.regular_incidence(4, 2048, seed=9)
    print(f"[case] determinism {precision}", flush=True)
    H = hip(A)
    opt = K.CyclicLanczosOptions(k=4, tol_rel=1e-12, tol_abs=0.0, prescreen_factor=1e-3)
    r1 = K.find_nullspace_cyclic_lanczos_ex(H, K.HipVectorFactory(A.n, 5), opt, precision)
    r2 = K.find_nullspace_cyclic_lanczos_ex(H, K.HipVectorFactory(A.n, 5), opt, precision)
    assert r1.right_null.shape == r2.right_null.shape == (A.n, 4)
    assert np.array_equal(r1.right_null.view(np.uint64), r2.right_null.view(np.uint64))
    assert (r1.restarts, r1.cg_solves, r1.krylov_steps_total) == \
        (r2.restarts, r2.cg_solves, r2.krylov_steps_total)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_result_same_as_host_result(precision):
    """kh_cyclic_lanczos_dev leaves the basis in a device block: the same
    bits and counters as kh_cyclic_lanczos with a host basis; a block with
    fewer columns than found receives the first ones."""
    A = W.regular_incidence(4, 2048, seed=9)
    print(f"[case] device result {precision}", flush=True)
    H = hip(A)
    opt = K.CyclicLanczosOptions(k=4, tol_rel=1e-12, tol_abs=0.0, prescreen_factor=1e-3)
    r1 = K.find_nullspace_cyclic_lanczos_ex(H, K.HipVectorFactory(A.n, 5), opt, precision)
    blk = K.DeviceBlock(H.ctx, A.n, 4)
    r2 = K.find_nullspace_cyclic_lanczos_dev(H, K.HipVectorFactory(A.n, 5), opt, blk, precision)
    assert r2.n_right == r1.right_null.shape[1] == 4
    U = blk.to_host(r2.n_right)
    assert np.array_equal(U.view(np.uint64), r1.right_null.view(np.uint64))
    assert (r1.restarts, r1.cg_solves, r1.spmv) == (r2.restarts, r2.cg_solves, r2.spmv)
    small = K.DeviceBlock(H.ctx, A.n, 2)
    r3 = K.find_nullspace_cyclic_lanczos_dev(H, K.HipVectorFactory(A.n, 5), opt, small, precision)
    assert r3.n_right == 4
    assert np.array_equal(small.to_host(2).view(np.uint64), r1.right_null[:, :2].view(np.uint64))
    with pytest.raises(ValueError):
        K.find_nullspace_cyclic_lanczos_dev(H, K.HipVectorFactory(A.n, 5), opt,
                                            K.DeviceBlock(H.ctx, A.n + 1, 4), precision)


@pytest.mark.parametrize("operand,precision",[("csr", "fp64"), ("csr", "mixed"),
                                               ("stencil", "fp64"), ("stencil", "mixed")])
def test_graph_replay_same_as_direct_launches(operand, precision):
    """Golub-Kahan cycles replayed from captured HIP graphs (graphs=1)
    against every kernel launched directly (graphs=-1): bitwise the
    same bases and counters. A graph is captured once per basis size and
    replayed on the following restarts, with the solve's own eager work and
    host synchronisations between the replays."""
    if operand == "csr":
        A = W.regular_incidence(4, 2048, seed=9)
        H, n, k = hip(A), A.n, 4
        kw = dict(k=k, tol_rel=1e-12, tol_abs=0.0, prescreen_factor=1e-3)
    else:
        n, N, c = 1000, 3, 0.5
        H, k = K.HipShiftedDifferenceOperator(n, N, c), 3
        kw = dict(k=k)
    print(f"[case] graph replay {operand} {precision}", flush=True)
    rg = K.find_nullspace_cyclic_lanczos_ex(H, K.HipVectorFactory(n, 5),
                                            K.CyclicLanczosOptions(graphs=1, **kw), precision)
    re = K.find_nullspace_cyclic_lanczos_ex(H, K.HipVectorFactory(n, 5),
                                            K.CyclicLanczosOptions(graphs=-1, **kw), precision)
    print(f"captures={rg.graph_captures} replays={rg.graph_replays} restarts={rg.restarts}",
          flush=True)
    assert re.graph_captures == 0 and re.graph_replays == 0
    assert 1 <= rg.graph_captures <= k + 1
    assert rg.graph_replays >= rg.graph_captures
    assert rg.right_null.shape == re.right_null.shape == (n, k)
    assert np.array_equal(rg.right_null.view(np.uint64), re.right_null.view(np.uint64))
    assert (rg.restarts, rg.cg_solves, rg.krylov_steps_total, rg.spmv, rg.spmtv) == \
        (re.restarts, re.cg_solves, re.krylov_steps_total, re.spmv, re.spmtv)


def test_callback_operator_is_not_captured():
    """A host callback operator synchronises and calls user code, so its
    cycles are launched directly even with graphs=1."""
    A = W.regular_incidence(2, 512, seed=4)
    H = hip(A)

    def apply(stream, precision, transpose, s, x, ldx, y, ldy):
        e = 8 if precision == "fp64" else 4
        p = 0 if precision == "fp64" else 1
        for j in range(s):
            K._check(K.lib().kh_spmm_dev(H.h, int(transpose), p, 1,
                                         C.c_void_p(x + j * ldx * e), C.c_void_p(y + j * ldy * e)))
    F = K.HipOperatorAdapter(A.m, A.n, apply)
    print("[case] callback operator without graphs", flush=True)
    opt = K.CyclicLanczosOptions(k=2, tol_rel=1e-12, tol_abs=0.0, prescreen_factor=1e-3,
                                 graphs=1)
    r = K.find_nullspace_cyclic_lanczos_ex(F, K.HipVectorFactory(A.n, 3), opt, "mixed")
    assert r.graph_captures == 0 and r.graph_replays == 0
    assert r.right_null.shape == (A.n, 2)


@pytest.mark.parametrize("screen", [0, 1])
def test_mixed_candidate_screens(screen):
    """Mixed precision on the incidence of 8 random 4-regular graphs x 16,384
    vertices, prescreen 1e-4 (threshold below sigma_min): candidates are
    screened from the projected problem's singular values (screen 0: with a
    zero target ||A Y c|| equals the even-block singular value of T-bar) or
    from FP32 candidate vectors (screen 1). Both find the 8 indicators to
    the threshold with sine <= 1e-10, screen most candidates, and agree on
    how many within a few (their restart vectors differ in rounding)."""
    g, v = 8, 16384
    A = W.regular_incidence(g, v, seed=42)
    print(f"[case] mixed screen {screen}", flush=True)
    opts = {s: K.CyclicLanczosOptions(k=g, tol_rel=1e-12, tol_abs=0.0, prescreen_factor=1e-4,
                                      max_restarts=400, screen=s) for s in (0, 1)}
    runs = {s: K.find_nullspace_cyclic_lanczos_ex(hip(A), K.HipVectorFactory(A.n, 42), o,
                                                  "mixed") for s, o in opts.items()}
    r = runs[screen]
    U = r.right_null
    assert U.shape[1] == g and r.converged_to_target
    assert residuals(A, U).max() <= 1e-12 * A.fro()
    assert principal_sine(U, W.component_indicators(g, v)) <= 1e-10
    assert r.screened >= 4 * r.restarts
    assert abs(runs[0].screened - runs[1].screened) <= max(8, runs[1].screened // 20)
