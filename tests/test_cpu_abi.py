"""The C-ABI library loads and exports every symbol include/krans_hip.h
declares (no compute without a GPU)."""
import os
import re

import krans_amd as K

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "krans_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(kh_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_exported_and_bound():
    L = K.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
        assert name in K.SIGNATURES, name
    assert K.lib().kh_version() == 1


def test_option_defaults_match_reference():
    o = K.CyclicLanczosOptions()
    assert (o.k, o.ncv, o.max_restarts, o.tol_abs, o.tol_rel, o.cg_maxit, o.cg_rel_tol,
            o.projector_passes, o.grad_max_iters, o.frob_samples, o.prescreen_factor,
            o.stagnation_restarts) == (1, 48, 80, 1e-12, 1e-10, 80, 1e-10, 1, 3000, 4, 0.8, 30)
    lo = K.LOBPCGOptions()
    assert (lo.nev, lo.block_size, lo.max_iters, lo.tol_abs, lo.tol_rel, lo.restart,
            lo.gs_tol, lo.frob_samples) == (1, 0, 1000, 1e-12, 1e-10, 0.9, 1e-14, 4)


def test_rng_stream_matches_oracle_without_gpu():
    import numpy as np
    import oracle as O
    f = K.HipVectorFactory(10, 42)
    g = O.Factory(42)
    a = np.concatenate([f.create_random_of_size(7), f.create_random_of_size(9)])
    assert np.array_equal(a, g.randn(16))


def test_rng_stream_long_and_ragged_matches_oracle():
    """The prefetching stream converts polar-method attempts in parallel
    batches of 2^18; draws of ragged sizes that cross several batch and ring
    boundaries must still be libstdc++'s sequential stream bit for bit."""
    import numpy as np
    import oracle as O
    for seed in (1, 42, 2**31 + 5):
        f = K.HipVectorFactory(10, seed)
        a = np.concatenate([f.create_random_of_size(s) for s in (1, 3, 262_143, 777, 1_500_001)])
        b = O.Factory(seed).randn(a.size)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), seed


def _tiny_svd(A, mode):
    import ctypes as C
    import numpy as np
    A = np.asfortranarray(A, dtype=np.float64)
    rows, cols = A.shape
    sv = np.zeros(cols)
    V = np.zeros((cols, cols), order="F")
    K._check(K.lib().kh_tiny_svd_right(A.ctypes.data_as(C.c_void_p), rows, cols, mode,
                                       sv.ctypes.data_as(C.c_void_p),
                                       V.ctypes.data_as(C.c_void_p)))
    return sv, V


def _tbar(l, betas, beta_last, tau=0.0):
    """The (l+1) x l refined Golub-Kahan matrix of solve_tiny_refined
    (src/Solvers/CyclicLanczosSolver.cpp:246-262) with alpha = 0."""
    import numpy as np
    T = np.zeros((l + 1, l))
    for i in range(l):
        T[i, i] = -tau
        if i + 1 < l:
            T[i, i + 1] = T[i + 1, i] = betas[i]
    T[l, l - 1] = beta_last
    return T


def test_tiny_svd_parity_blocks_bitwise_and_against_lapack():
    """The refined-extraction SVD sweeps the even- and odd-column blocks of
    T-bar separately: bitwise the singular values and vectors of the sweeps
    over the whole matrix (including repeated, tiny and zero betas), the even
    block's vectors unchanged when the odd block's are not formed, and
    singular values within 1e-13 * sigma_max of LAPACK's."""
    import numpy as np
    rng = np.random.default_rng(3)
    for l in (1, 2, 3, 8, 47, 48, 66, 67):
        for case in range(4):
            b = rng.uniform(0.1, 3.0, size=max(l - 1, 0))
            if case == 1 and l > 4:
                b[::3] *= 1e-13
            if case == 2:
                b[:] = 1.0
            T = _tbar(l, b, 0.0 if case == 3 else rng.uniform(0.1, 3.0))
            s0, V0 = _tiny_svd(T, 0)
            s1, V1 = _tiny_svd(T, 1)
            s2, V2 = _tiny_svd(T, 2)
            assert np.array_equal(s0.view(np.uint64), s1.view(np.uint64)), (l, case)
            assert np.array_equal(V0.view(np.uint64), V1.view(np.uint64)), (l, case)
            assert np.array_equal(s0.view(np.uint64), s2.view(np.uint64)), (l, case)
            even = (V0[0::2, :] ** 2).sum(axis=0) >= 1e-20
            assert np.array_equal(V0[:, even].view(np.uint64), V2[:, even].view(np.uint64))
            assert np.all(V2[0::2, ~even] == 0.0)
            ref = np.linalg.svd(T, compute_uv=False)
            assert np.abs(s0 - ref).max() <= 1e-13 * max(ref.max(), 1.0), (l, case)
            # right singular vectors: T v_j = s_j u_j with orthonormal V
            assert np.abs(V0.T @ V0 - np.eye(l)).max() <= 1e-13
            Tv = T @ V0
            assert np.allclose(np.linalg.norm(Tv, axis=0), s0, atol=1e-13 * max(ref.max(), 1.0))


def test_tiny_svd_split_falls_back_on_other_matrices():
    """A nonzero target_tau (diagonal) or a dense matrix has no parity
    blocks: every mode gives the whole-matrix sweeps."""
    import numpy as np
    rng = np.random.default_rng(5)
    for T in (_tbar(20, rng.uniform(0.5, 2, 19), 0.7, tau=0.3), rng.standard_normal((9, 7))):
        s0, V0 = _tiny_svd(T, 0)
        for mode in (1, 2):
            s, V = _tiny_svd(T, mode)
            assert np.array_equal(s0.view(np.uint64), s.view(np.uint64))
            assert np.array_equal(V0.view(np.uint64), V.view(np.uint64))
        assert np.abs(s0 - np.linalg.svd(T, compute_uv=False)).max() <= 1e-13 * s0.max()
