"""GPU parity of the paths the reference's difference-operator regression
pins (tests/test_difference_operator.cpp:49-118, README.md:231-237): the
stagnation guard, cg_maxit accounting, the randomized fallback and LOBPCG's
block sizing; and the precision-variant configurations (shifted difference
N=16, c=0.9; weighted grid incidence) at reduced size against the exact
kernels, the FP64 path and the oracle.

Every solve runs through the C ABI (libkrans_hip.so). Tolerances are stated
per test; with tol_rel = 1e-12 the largest principal-angle sine to the exact
kernel and to the oracle's basis is held to 1e-10.
"""
import numpy as np
import pytest

import oracle as O
import krans_amd as K
from krans_amd import workloads as W
from conftest import principal_sine
from helpers import (TWO_PASS_OPTS, csr_small_2x3, gradient_steps_closed_form, residuals,
                     two_pass_case)

pytestmark = pytest.mark.gpu

PRECISIONS = ["fp64", "mixed"]


def diff_operand(kind, n):
    """The 1-D difference operator (D x)_i = x_{i+1} - x_i as the stencil
    (N = 1, c = 1) or as its CSR."""
    if kind == "stencil":
        return K.HipShiftedDifferenceOperator(n, 1, 1.0)
    return K.HipCsrMatrixAdapter.from_csr(W.shifted_difference(n, 1, 1.0))


def constant_unit_err(u):
    return float(np.abs(np.abs(u) - 1.0 / np.sqrt(u.size)).max())


def hip(A):
    return K.HipCsrMatrixAdapter.from_csr(A)


def ora(A):
    return O.Operator.csr(A.m, A.n, A.rowptr, A.col, A.val)


# ---- cyclic Lanczos: stagnation, cg_maxit_hits ----------------------------

@pytest.mark.parametrize("kind", ["stencil", "csr"])
@pytest.mark.parametrize("precision", PRECISIONS)
def test_cyclic_defaults_stagnate(kind, precision):
    """README.md:236 (oracle: the same counts): with the defaults every
    restart's first polish stops at cg_maxit = 80, stagnation ends the loop
    after 30 restarts, the fallback fails: 0 vectors, 30 CG solves, 30 hits."""
    print(f"[case] difference operator defaults {kind} {precision}", flush=True)
    n = 1000
    r = K.find_nullspace_cyclic_lanczos_ex(diff_operand(kind, n), K.HipVectorFactory(n, 42),
                                           K.CyclicLanczosOptions(), precision)
    assert r.right_null.shape[1] == 0 and not r.converged_to_target
    assert r.cg_solves == 30 and r.cg_maxit_hits == 30
    assert r.restarts == 29


@pytest.mark.parametrize("kind", ["stencil", "csr"])
@pytest.mark.parametrize("precision", PRECISIONS)
def test_cyclic_cg_maxit_1000_finds_constant_vector(kind, precision):
    """README.md:237, test_difference_operator.cpp:106-115: with cg_maxit
    1000 one CG solve gives +-(1,...,1)/sqrt(n) (checked to 1e-8 as there)."""
    print(f"[case] difference operator cg_maxit 1000 {kind} {precision}", flush=True)
    n = 1000
    r = K.find_nullspace_cyclic_lanczos_ex(diff_operand(kind, n), K.HipVectorFactory(n, 42),
                                           K.CyclicLanczosOptions(cg_maxit=1000), precision)
    assert r.right_null.shape[1] == 1 and r.converged_to_target
    assert constant_unit_err(r.right_null[:, 0]) <= 1e-8
    if precision == "fp64":
        assert r.cg_solves == 1 and r.cg_maxit_hits == 0
    # mixed: the FP32 projector CG on A A^T (condition number ~4e5 here)
    # refines over several polishes before the FP64 residual passes; the
    # counters then differ from the FP64 path's, the vector does not


def test_cyclic_two_projector_passes_counts_match_oracle():
    """cg_maxit_hits per pass (CyclicLanczosSolver.cpp:522-526): the oracle
    and the FP64 GPU path count 5 solves and 5 hits on the same input."""
    print("[case] two projector passes", flush=True)
    A = two_pass_case()
    Uo, _, ro = O.cyclic(ora(A), O.Factory(42), O.cyclic_defaults(**TWO_PASS_OPTS))
    r = K.find_nullspace_cyclic_lanczos_ex(hip(A), K.HipVectorFactory(A.n, 42),
                                           K.CyclicLanczosOptions(**TWO_PASS_OPTS), "fp64")
    assert (ro.cg_solves, ro.cg_maxit_hits) == (5, 5)
    assert (r.cg_solves, r.cg_maxit_hits) == (ro.cg_solves, ro.cg_maxit_hits)
    assert r.right_null.shape[1] == Uo.shape[1] == 2
    assert principal_sine(r.right_null, Uo) <= 1e-10


# ---- randomized fallback ---------------------------------------------------

@pytest.mark.parametrize("precision", PRECISIONS)
def test_fallback_projector_finds_the_kernel(precision):
    """max_restarts 0 (the fallback alone, FP64 in both precisions): the
    exact projector gives [-1,-1,1]/sqrt(3) with 8 A x and 3 A^T x, as the
    oracle counts them."""
    print(f"[case] fallback projector {precision}", flush=True)
    A = csr_small_2x3()
    r = K.find_nullspace_cyclic_lanczos_ex(hip(A), K.HipVectorFactory(3, 1),
                                           K.CyclicLanczosOptions(k=1, max_restarts=0), precision)
    e = np.array([-1.0, -1.0, 1.0]) / np.sqrt(3)
    assert r.right_null.shape[1] == 1 and r.converged_to_target
    assert principal_sine(r.right_null, e[:, None]) < 1e-14
    # mixed adds the residual product of its final FP64 pass (one A x per
    # locked vector; the vector is below the target, so nothing more)
    assert (r.spmv, r.spmtv) == ((8, 3) if precision == "fp64" else (9, 3))
    assert r.refined == 0


@pytest.mark.parametrize("precision", PRECISIONS)
def test_fallback_when_k_exceeds_the_nullity(precision):
    """k = 2 on a matrix of nullity 1: after 30 restarts without an accept the
    fallback returns the unnormalised remainder (norm < 1e-14) whose residual
    passes the test, as the reference and the oracle do
    (CyclicLanczosSolver.cpp:176-186)."""
    print(f"[case] fallback k > nullity {precision}", flush=True)
    A = csr_small_2x3()
    Uo, _, ro = O.cyclic(ora(A), O.Factory(1), O.cyclic_defaults(k=2))
    r = K.find_nullspace_cyclic_lanczos_ex(hip(A), K.HipVectorFactory(3, 1),
                                           K.CyclicLanczosOptions(k=2), precision)
    U = r.right_null
    e = np.array([-1.0, -1.0, 1.0]) / np.sqrt(3)
    assert U.shape[1] == Uo.shape[1] == 2 and r.converged_to_target
    assert r.restarts == ro.restarts == 30
    assert principal_sine(U[:, :1], e[:, None]) < 1e-14
    assert np.linalg.norm(U[:, 1]) < 1e-14


def test_fallback_gradient_branch_closed_form():
    """A = 2e-16 [[1,0,1],[0,1,1]]: the projector CG breaks down at its first
    iteration and the gradient steps (step 0.9e30 at the L2 clamp) converge;
    the step count agrees with the eigen-decomposition of B^T B and with the
    oracle, the vector is within thr / (s sigma_min) = 5e-10 of the kernel."""
    print("[case] fallback gradient branch", flush=True)
    S = csr_small_2x3()
    s, thr = 2e-16, 1e-25
    As = W.CsrMatrix(S.m, S.n, S.rowptr, S.col, S.val * s)
    opt = dict(k=1, max_restarts=0, tol_abs=thr, tol_rel=0.0)
    r = K.find_nullspace_cyclic_lanczos_ex(hip(As), K.HipVectorFactory(3, 1),
                                           K.CyclicLanczosOptions(**opt), "fp64")
    Uo, _, ro = O.cyclic(ora(As), O.Factory(1), O.cyclic_defaults(**opt))
    e = np.array([-1.0, -1.0, 1.0]) / np.sqrt(3)
    assert r.right_null.shape[1] == 1 and r.converged_to_target
    assert principal_sine(r.right_null, e[:, None]) <= thr / s
    v0 = O.Factory(1).randn(18)[15:]
    v0 /= np.linalg.norm(v0)
    t_pred = gradient_steps_closed_form(S.to_scipy().toarray(), s, v0, thr)
    t = r.spmtv - 1
    assert r.spmv == 11 + 2 * t
    assert abs(t - t_pred) <= 1, (t, t_pred)
    assert abs(t - (ro.spmtv - 1)) <= 1


# ---- LOBPCG block sizing (LOBPCGSolver.cpp:480-490) -----------------------

@pytest.mark.parametrize("kind", ["stencil", "csr"])
def test_lobpcg_auto_block_size_is_one(kind):
    """test_difference_operator.cpp:55-68 (n = 100), FP64: block 1, no vector
    within the default 1000 iterations (README.md:233), the residual gap
    reported."""
    print(f"[case] lobpcg auto block {kind}", flush=True)
    r = K.find_nullspace_lobpcg_ex(diff_operand(kind, 100), K.HipVectorFactory(100, 42),
                                   K.LOBPCGOptions(), "fp64")
    assert r.block_size == 1
    assert r.converged_to_target == (r.right_null.shape[1] == 1)
    assert not r.converged_to_target
    assert r.iters == 1000 and r.best_unaccepted_residual > r.accept_threshold


@pytest.mark.parametrize("kind", ["stencil", "csr"])
@pytest.mark.parametrize("precision", PRECISIONS)
def test_lobpcg_block_16_finds_constant_vector(kind, precision):
    """test_difference_operator.cpp:70-81: block 16 > nev is honoured and
    finds +-(1,...,1)/sqrt(n) to 1e-8."""
    print(f"[case] lobpcg block 16 {kind} {precision}", flush=True)
    r = K.find_nullspace_lobpcg_ex(diff_operand(kind, 100), K.HipVectorFactory(100, 42),
                                   K.LOBPCGOptions(block_size=16), precision)
    assert r.block_size == 16 and r.converged_to_target and r.right_null.shape[1] == 1
    assert constant_unit_err(r.right_null[:, 0]) <= 1e-8


@pytest.mark.parametrize("precision", PRECISIONS)
def test_lobpcg_block_capped_at_cols(precision):
    """test_difference_operator.cpp:83-92: block 16 on a 3 x 4 operator is
    capped at cols(A) = 4 and converges."""
    print(f"[case] lobpcg block cap {precision}", flush=True)
    r = K.find_nullspace_lobpcg_ex(diff_operand("stencil", 4), K.HipVectorFactory(4, 42),
                                   K.LOBPCGOptions(block_size=16), precision)
    assert r.block_size == 4 and r.converged_to_target
    assert constant_unit_err(r.right_null[:, 0]) <= 1e-8


# ---- precision variant (a): shifted difference N = 16, c = 0.9 -----------

C3A_N, C3A_C, C3A_K = 16, 0.9, 16


def c3a_options(A):
    # prescreen threshold 0.05, below sigma_min = 0.1 (as the benchmark)
    return dict(k=C3A_K, tol_rel=1e-12, tol_abs=0.0, prescreen_factor=0.05 / A.fro(),
                max_restarts=400, cg_maxit=2000)


@pytest.fixture(scope="module")
def c3a_bases():
    n = 1 << 18
    A = W.shifted_difference(n, C3A_N, C3A_C)
    opt = K.CyclicLanczosOptions(**c3a_options(A))
    out = {}
    for kind in ("csr", "stencil"):
        H = hip(A) if kind == "csr" else K.HipShiftedDifferenceOperator(n, C3A_N, C3A_C)
        for precision in PRECISIONS:
            print(f"[case] c3a n={n} {kind} {precision}", flush=True)
            out[kind, precision] = K.find_nullspace_cyclic_lanczos_ex(
                H, K.HipVectorFactory(n, 42), opt, precision)
    return A, out


@pytest.mark.parametrize("kind", ["csr", "stencil"])
@pytest.mark.parametrize("precision", PRECISIONS)
def test_c3a_shifted_difference_exact_kernel(c3a_bases, kind, precision):
    """n = 2^18: residual <= 1e-12 ||A||_F (exact norm from the values),
    principal-angle sine to the exact kernel (README.md:24-28) <= 1e-10,
    orthonormal to 1e-12."""
    A, out = c3a_bases
    r = out[kind, precision]
    U = r.right_null
    assert U.shape[1] == C3A_K and r.converged_to_target
    assert residuals(A, U).max() <= 1e-12 * A.fro()
    Q = W.shifted_difference_kernel(A.n, C3A_N, C3A_C)
    assert principal_sine(U, Q) <= 1e-10
    assert np.abs(U.T @ U - np.eye(C3A_K)).max() <= 1e-12


@pytest.mark.parametrize("kind", ["csr", "stencil"])
def test_c3a_mixed_agrees_with_fp64(c3a_bases, kind):
    _, out = c3a_bases
    assert principal_sine(out[kind, "mixed"].right_null, out[kind, "fp64"].right_null) <= 1e-10


@pytest.mark.parametrize("precision", PRECISIONS)
def test_c3a_small_matches_oracle(precision):
    """n = 2^13 on the oracle and the GPU (CSR), same seed and options:
    both bases within 1e-10 of the exact kernel and of each other."""
    n = 1 << 13
    A = W.shifted_difference(n, C3A_N, C3A_C)
    kw = c3a_options(A)
    print(f"[case] c3a oracle n={n} {precision}", flush=True)
    r = K.find_nullspace_cyclic_lanczos_ex(hip(A), K.HipVectorFactory(n, 42),
                                           K.CyclicLanczosOptions(**kw), precision)
    Uo, _, ro = O.cyclic(ora(A), O.Factory(42), O.cyclic_defaults(**kw))
    Q = W.shifted_difference_kernel(n, C3A_N, C3A_C)
    assert r.right_null.shape[1] == Uo.shape[1] == C3A_K
    assert principal_sine(Uo, Q) <= 1e-10
    assert principal_sine(r.right_null, Q) <= 1e-10
    assert principal_sine(r.right_null, Uo) <= 1e-10
    assert residuals(A, r.right_null).max() <= 1e-12 * A.fro()


# ---- precision variant (b): weighted grid incidence -----------------------

@pytest.mark.parametrize("precision", PRECISIONS)
def test_c3b_weighted_grids_match_oracle(precision):
    """4 grids 16 x 16, weights 10^(-3U), the benchmark's options (Jacobi
    preconditioner for A A^T, prescreen 1e-3, cg_maxit 4000): the GPU finds
    as many vectors as the oracle, each within the acceptance threshold and
    inside the span of the grid indicators to 1e-10."""
    grids, side = 4, 16
    A = W.grid_incidence(grids, side, seed=42, weighted=True)
    kw = dict(k=grids, tol_rel=1e-12, tol_abs=0.0, prescreen_factor=1e-3, max_restarts=400,
              cg_maxit=4000)
    print(f"[case] c3b {grids}x{side}x{side} {precision}", flush=True)
    r = K.find_nullspace_cyclic_lanczos_ex(hip(A), K.HipVectorFactory(A.n, 42),
                                           K.CyclicLanczosOptions(precond_AAt="jacobi", **kw),
                                           precision)
    Uo, _, ro = O.cyclic(ora(A), O.Factory(42), O.cyclic_defaults(precond_jacobi=1, **kw))
    U = r.right_null
    assert U.shape[1] == Uo.shape[1] >= 1
    thr = 1e-12 * r.frob_norm_est
    assert residuals(A, U).max() <= thr
    Kx = W.component_indicators(grids, side * side)
    sine = float(np.linalg.norm(U - Kx @ (Kx.T @ U), 2))
    sine_o = float(np.linalg.norm(Uo - Kx @ (Kx.T @ Uo), 2))
    print(f"gpu {U.shape[1]} vectors sine {sine:.2e}, oracle {Uo.shape[1]} sine {sine_o:.2e}",
          flush=True)
    # sigma_min of the weighted incidence is ~1e-5 sigma_max, so a residual
    # at the threshold bounds the angle only to ||A U|| / sigma_min: the
    # reference algorithm's own basis (the oracle's) is ~3e-10 from the
    # indicator span. FP64 is held to the oracle's accuracy; mixed ends with
    # the FP64 projector pass to the final-polish target and meets 1e-10.
    s = np.linalg.svd(A.to_scipy().toarray(), compute_uv=False)
    smin = s[s > 1e-8 * s[0]].min()
    assert sine <= np.linalg.norm(A.to_scipy() @ U, 2) / smin * 1.01
    if precision == "fp64":
        assert sine <= max(1e-10, 2.0 * sine_o)
    else:
        assert sine <= 1e-10


# ---- unweighted grid incidence: the FP32 projector CG's stall exit ---------

def grid_kernel(grids, side):
    """The component indicators of `grids` side x side grids, orthonormal."""
    vpg = side * side
    Q = np.zeros((grids * vpg, grids))
    for g in range(grids):
        Q[g * vpg:(g + 1) * vpg, g] = 1.0 / side
    return Q


def grid_opts(k):
    return K.CyclicLanczosOptions(k=k, tol_rel=1e-12, tol_abs=0.0, prescreen_factor=1e-4,
                                  max_restarts=400, cg_maxit=1000)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_grid_incidence_small_finds_the_kernel(precision):
    """6 unweighted 46 x 46 grids (n = 12,696): both precisions find the 6
    indicators with max ||Au|| <= thr = 1e-12 ||A||_F. The smallest nonzero
    singular value is sigma = 2 sin(pi / 92) ~ 0.068, so the acceptance test
    alone bounds the principal-angle sine by thr / sigma ~ 3e-9: the FP64 path
    (the reference's algorithm) is held to that bound, the mixed path, whose
    locked vectors end with a batched FP64 projector pass to 0.01 thr, to
    1e-10. Before the FP32 projector CG had its stall exit, every mixed
    polish here ran to cg_maxit on candidates whose FP64 residual it then made
    worse."""
    print(f"[case] grid incidence 6 x 46^2 {precision}", flush=True)
    A = W.grid_incidence(6, 46, seed=1202, weighted=False)
    r = K.find_nullspace_cyclic_lanczos_ex(hip(A), K.HipVectorFactory(A.n, 1202), grid_opts(6),
                                           precision)
    U = r.right_null
    thr = 1e-12 * A.fro()
    assert U.shape[1] == 6
    assert residuals(A, U).max() <= thr
    sine = principal_sine(U, grid_kernel(6, 46))
    assert sine <= (1e-10 if precision == "mixed" else thr / (2 * np.sin(np.pi / 92)))
    assert r.cg_maxit_hits == 0


def test_grid_incidence_mixed_escalates_stalled_fp32_polish():
    """5 unweighted 233 x 233 grids (n = 271,445; sigma_min ~ 0.013): the
    candidates' residuals lie on the smallest singular values, the FP32
    projector CG stalls at its rounding level and those columns are solved
    again in FP64 (escalated > 0); all 5 indicators are found to the
    threshold with sine <= 1e-10. (The FP64 path finds 1 of 5 within
    stagnation_restarts here: its candidates' CGs stop at cg_maxit.)"""
    print("[case] grid incidence 5 x 233^2 mixed", flush=True)
    A = W.grid_incidence(5, 233, seed=1197, weighted=False)
    r = K.find_nullspace_cyclic_lanczos_ex(hip(A), K.HipVectorFactory(A.n, 1197), grid_opts(5),
                                           "mixed")
    U = r.right_null
    assert r.escalated > 0
    assert U.shape[1] == 5 and r.converged_to_target
    assert residuals(A, U).max() <= 1e-12 * A.fro()
    assert principal_sine(U, grid_kernel(5, 233)) <= 1e-10
