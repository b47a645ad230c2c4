"""Kernel-level parity through the C ABI: reductions, CGS2 and the batched
Gram products against numpy FP64 references (the oracle's arithmetic for
these is plain sums; tolerances are stated per test)."""
import ctypes as C

import numpy as np
import pytest

import krans_amd as K

pytestmark = pytest.mark.gpu


class Dev:
    """A device copy of a numpy array (FP64 or FP32)."""

    def __init__(self, ctx, a):
        self.ctx = ctx
        self.a = np.asfortranarray(a)
        self.p = C.c_void_p()
        K._check(K.lib().kh_dev_alloc(ctx.h, max(self.a.nbytes, 8), C.byref(self.p)))
        K._check(K.lib().kh_h2d(ctx.h, self.p, self.a.ctypes.data_as(C.c_void_p), self.a.nbytes))

    def get(self):
        out = np.empty_like(self.a, order="F")
        K._check(K.lib().kh_d2h(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.p, out.nbytes))
        return out

    def __del__(self):
        try:
            K.lib().kh_dev_free(self.ctx.h, self.p)
        except Exception:
            pass


@pytest.fixture(scope="module")
def ctx():
    return K.default_context()


@pytest.mark.parametrize("n", [1, 1000, 2_000_003])
def test_dot_fp64_and_fp32(ctx, n):
    print(f"[case] dot n={n}", flush=True)
    rng = np.random.default_rng(n)
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    da, db = Dev(ctx, a), Dev(ctx, b)
    out = C.c_double()
    K._check(K.lib().kh_dot_dev(ctx.h, 0, da.p, db.p, n, C.byref(out)))
    ref = float(np.dot(a, b))
    assert abs(out.value - ref) <= 1e-14 * np.abs(a) @ np.abs(b) + 1e-300
    a32, b32 = a.astype(np.float32), b.astype(np.float32)
    da32, db32 = Dev(ctx, a32), Dev(ctx, b32)
    K._check(K.lib().kh_dot_dev(ctx.h, 1, da32.p, db32.p, n, C.byref(out)))
    ref32 = float(np.dot(a32.astype(np.float64), b32.astype(np.float64)))
    # FP32 inputs, FP64 accumulation
    assert abs(out.value - ref32) <= 1e-14 * np.abs(a) @ np.abs(b) + 1e-300
    # determinism: same bits twice
    o2 = C.c_double()
    K._check(K.lib().kh_dot_dev(ctx.h, 1, da32.p, db32.p, n, C.byref(o2)))
    assert o2.value == out.value


@pytest.mark.parametrize("s,n", [(1, 5000), (33, 100_000), (96, 300_000)])
def test_cgs2_orthogonalises(ctx, s, n):
    print(f"[case] cgs2 s={s} n={n}", flush=True)
    rng = np.random.default_rng(s)
    V, _ = np.linalg.qr(rng.standard_normal((n, s)))
    w = rng.standard_normal(n) + V @ rng.standard_normal(s)
    dV, dw = Dev(ctx, V), Dev(ctx, w)
    h = np.zeros(s)
    nrm = C.c_double()
    K._check(K.lib().kh_cgs2_dev(ctx.h, 0, dV.p, s, dw.p, n, h.ctypes.data_as(C.c_void_p),
                                 C.byref(nrm)))
    w2 = dw.get()
    ref = w - V @ (V.T @ w)
    assert np.abs(V.T @ w2).max() <= 1e-13 * np.linalg.norm(w)
    assert np.allclose(h, V.T @ w, rtol=1e-12, atol=1e-12 * np.linalg.norm(w))
    assert np.linalg.norm(w2 - ref) <= 1e-12 * np.linalg.norm(w)
    assert abs(nrm.value - np.linalg.norm(w2)) <= 1e-13 * nrm.value


@pytest.mark.parametrize("k,n,prec", [(3, 1000, 0), (40, 50_000, 0), (96, 200_000, 0),
                                      (3, 1000, 1), (17, 4099, 1), (96, 200_000, 1),
                                      (128, 70_001, 1), (130, 5000, 1), (96, 3_000_017, 1)])
def test_gram_blocks(ctx, k, n, prec):
    print(f"[case] gram k={k} n={n} prec={prec}", flush=True)
    rng = np.random.default_rng(k)
    Z = rng.standard_normal((n, k))
    AZ = rng.standard_normal((n, k))
    if prec == 1:
        Z, AZ = Z.astype(np.float32), AZ.astype(np.float32)
    dZ, dA = Dev(ctx, Z), Dev(ctx, AZ)
    S = np.zeros((k, k), order="F")
    G = np.zeros((k, k), order="F")
    K._check(K.lib().kh_gram_dev(ctx.h, prec, dZ.p, dA.p, k, n, S.ctypes.data_as(C.c_void_p),
                                 G.ctypes.data_as(C.c_void_p)))
    Z64, A64 = Z.astype(np.float64), AZ.astype(np.float64)
    Sr, Gr = Z64.T @ Z64, Z64.T @ A64
    scale = np.sqrt(np.outer(np.sum(Z64 * Z64, 0), np.sum(A64 * A64, 0)))
    # FP64: fixed-order FP64 sums; FP32: FP32 MFMA accumulation over row
    # chunks of at most 4096 rows, FP64 across chunks (k <= 128: the fused
    # S/G pass; k = 130: the 64 x 64 tile kernel)
    tol = 1e-13 if prec == 0 else 1e-5
    assert np.abs(S - Sr).max() <= tol * scale.max()
    assert np.abs(G - Gr).max() <= tol * scale.max()


def _sequential_rows(A, X, dt, transpose):
    """op(A) X with every row's sum taken sequentially from 0 in ascending
    column order, each product and sum rounded to dt (no contraction)."""
    S = A.to_scipy()
    if transpose:
        S = S.T.tocsr()
        S.sort_indices()
    rp, ci, va = S.indptr, S.indices, S.data.astype(dt)
    lens = np.diff(rp)
    out = np.zeros((S.shape[0], X.shape[1]), dt)
    for r in range(int(lens.max(initial=0))):
        rows = np.nonzero(lens > r)[0]
        p = rp[rows] + r
        out[rows] = out[rows] + va[p, None] * X[ci[p]]
    return out


@pytest.mark.parametrize("case,transpose", [("incidence", 0), ("incidence", 1),
                                            ("uniform1", 0), ("uniform3", 0),
                                            ("uniform8", 0), ("uniform5", 0),
                                            ("uniform3", 1)])
@pytest.mark.parametrize("prec", [0, 1])
def test_spmm_uniform_rows_bit_exact(case, transpose, prec):
    """Operators whose rows all have the same number of entries R (the
    incidence: 2 per row, 4 per column; random rows of 1, 3, 8) take the
    fixed-length kernels, which read no row pointers; their sums are the
    sequential ones, bitwise, in FP64 and FP32, for one and several columns.
    R = 5 and the ragged transposes take the pointer-based kernels."""
    from helpers import random_sparse
    from krans_amd import workloads as W
    if case == "incidence":
        A = W.regular_incidence(4, 512, seed=7)
    else:
        A = random_sparse(3000, 2000, int(case[-1]), seed=int(case[-1]))
    print(f"[case] spmm uniform {case} t={transpose} prec={prec}", flush=True)
    H = K.HipCsrMatrixAdapter.from_csr(A)
    nin, nout = (A.m, A.n) if transpose else (A.n, A.m)
    dt = np.float64 if prec == 0 else np.float32
    for s in (1, 5):
        X = np.random.default_rng(s).standard_normal((nin, s)).astype(dt)
        dX, dY = Dev(H.ctx, X), Dev(H.ctx, np.zeros((nout, s), dt))
        K._check(K.lib().kh_spmm_dev(H.h, transpose, prec, s, dX.p, dY.p))
        Y, ref = dY.get(), _sequential_rows(A, X, dt, transpose)
        assert Y.dtype == dt and np.array_equal(Y, ref), s


@pytest.mark.parametrize("s,transpose", [(1, 0), (9, 0), (96, 0), (96, 1), (130, 1)])
def test_spmm_rows_bit_exact(s, transpose):
    """Row-major block product (LOBPCG normal products): per entry the same
    sequential ascending sum as the oracle's CSR / CSC product, bit-exact in
    FP64, including rows longer than the one-lane threshold."""
    import oracle as O
    from helpers import random_sparse
    A = random_sparse(700, 500, 40, seed=11)
    print(f"[case] spmm rows s={s} t={transpose}", flush=True)
    H = K.HipCsrMatrixAdapter.from_csr(A)
    Op = O.Operator.csr(A.m, A.n, A.rowptr, A.col, A.val)
    nin, nout = (A.m, A.n) if transpose else (A.n, A.m)
    X = np.random.default_rng(s).standard_normal((nin, s))
    # Dev stores column-major, so the s x nin transpose lands as row-major X
    dX = Dev(H.ctx, X.T)
    dY = Dev(H.ctx, np.zeros((s, nout)))
    K._check(K.lib().kh_spmm_rows_dev(H.h, transpose, 0, s, dX.p, dY.p))
    Y = np.empty((nout, s))
    K._check(K.lib().kh_d2h(H.ctx.h, Y.ctypes.data_as(C.c_void_p), dY.p, Y.nbytes))
    for c in range(s):
        ref = Op.mtv(X[:, c]) if transpose else Op.mv(X[:, c])
        assert np.array_equal(Y[:, c].view(np.uint64), ref.view(np.uint64)), c


@pytest.mark.parametrize("s,n,prec", [(1, 5000, 0), (33, 100_003, 0), (97, 300_000, 1),
                                      (7, 2049, 1), (7, 4096, 1),
                                      (20, 2**21 + 12_345, 1),  # > 16384 tiles: pipelined
                                      (97, 300_000, 0), (128, 70_001, 0)])  # FP64 > 64 KB
def test_cgs2_fused_equals_four_sweeps(ctx, s, n, prec):
    """The fused middle kernel (w -= V h1 and V^T w from one read of V) gives
    the results of the separate update and projection sweeps: the update of
    w is the same arithmetic; the second projection's partial sums are grouped
    per 128-row block in the LDS-tiled form (12 <= s), and the separate FP32
    projection of aligned blocks (len a multiple of 4) groups rows by float4,
    so h2 and the final w agree to rounding there and bitwise otherwise."""
    print(f"[case] cgs2 fused s={s} n={n} prec={prec}", flush=True)
    rng = np.random.default_rng(s + n)
    V, _ = np.linalg.qr(rng.standard_normal((n, s)))
    w = rng.standard_normal(n) + V @ rng.standard_normal(s)
    dt = np.float64 if prec == 0 else np.float32
    V, w = V.astype(dt), w.astype(dt)
    outs = []
    for fused in (0, 1):
        dV, dw = Dev(ctx, V), Dev(ctx, w)
        h = np.zeros(s)
        nrm = np.zeros(1)
        K._check(K.lib().kh_cgs2_dev_ex(ctx.h, prec, dV.p, s, dw.p, n, fused,
                                        h.ctypes.data_as(C.c_void_p),
                                        nrm.ctypes.data_as(C.c_void_p)))
        outs.append((dw.get(), h, nrm))
    if s < 12 and (prec == 0 or n % 4 != 0):
        for a, b in zip(outs[0], outs[1]):
            assert np.array_equal(a, b)
        return
    (w0, h0, n0), (w1, h1, n1) = outs
    eps = np.finfo(dt).eps
    wn = float(np.linalg.norm(w.astype(np.float64)))
    assert np.abs(w0.astype(np.float64) - w1.astype(np.float64)).max() <= 4 * eps * wn
    assert np.abs(h0 - h1).max() <= 1e-13 * wn
    assert abs(n0[0] - n1[0]) <= 16 * eps * n0[0]
    # and both are orthogonal to V to working precision
    for wx in (w0, w1):
        assert np.abs(V.astype(np.float64).T @ wx.astype(np.float64)).max() <= 64 * eps * wn


@pytest.mark.parametrize("s,n,prec", [(1, 5000, 0), (33, 100_003, 1), (64, 2**20, 1),
                                      (97, 300_000, 0), (20, 2**21 + 12_344, 1)])
@pytest.mark.parametrize("cancel", [False, True])
def test_cgs2_selective_pair(ctx, s, n, prec, cancel):
    """Selective second pass (kh_cgs2_dev_select, the form the mixed
    Golub-Kahan steps use). The Daniel-Gragg-Kaufman-Stewart test with
    eta = 1/sqrt(2) decides on the device: w mostly inside span(V) (the first
    pass cancels most of ||w||^2) must take the second pass, and then the
    results are bitwise kh_cgs2_dev's; w mostly outside span(V) must skip it,
    and the single pass must still leave w orthogonal to V to working
    precision (the "twice is enough" bound for an uncancelled pass) and equal
    w - V V^T w to rounding."""
    print(f"[case] cgs2 selective s={s} n={n} prec={prec} cancel={cancel}", flush=True)
    rng = np.random.default_rng(7 * s + n)
    V, _ = np.linalg.qr(rng.standard_normal((n, s)))
    if cancel:  # ||V^T w||^2 > ||w||^2 / 2
        w = V @ rng.standard_normal(s) + 1e-3 * rng.standard_normal(n) / np.sqrt(n)
    else:  # ||V^T w||^2 about s / n of ||w||^2
        w = rng.standard_normal(n) + 1e-2 * (V @ rng.standard_normal(s))
    dt = np.float64 if prec == 0 else np.float32
    V, w = V.astype(dt), w.astype(dt)
    Vd, wd = V.astype(np.float64), w.astype(np.float64)
    h1 = Vd.T @ wd
    assert (h1 @ h1 > 0.5 * (wd @ wd)) == cancel  # the case is what it claims
    dV, dw = Dev(ctx, V), Dev(ctx, w)
    h, nrm, ran = np.zeros(s), np.zeros(1), np.zeros(1, dtype=np.int32)
    K._check(K.lib().kh_cgs2_dev_select(ctx.h, prec, dV.p, s, dw.p, n,
                                        h.ctypes.data_as(C.c_void_p),
                                        nrm.ctypes.data_as(C.c_void_p),
                                        ran.ctypes.data_as(C.c_void_p)))
    wsel = dw.get()
    assert int(ran[0]) == int(cancel)
    eps = np.finfo(dt).eps
    if cancel:
        dV2, dw2 = Dev(ctx, V), Dev(ctx, w)
        h2, n2 = np.zeros(s), np.zeros(1)
        K._check(K.lib().kh_cgs2_dev(ctx.h, prec, dV2.p, s, dw2.p, n,
                                     h2.ctypes.data_as(C.c_void_p),
                                     n2.ctypes.data_as(C.POINTER(C.c_double))))
        assert np.array_equal(wsel, dw2.get())
        assert np.array_equal(h, h2) and np.array_equal(nrm, n2)
    else:
        ref = wd - Vd @ h1
        wn = float(np.linalg.norm(wd))
        assert np.abs(wsel.astype(np.float64) - ref).max() <= 4 * eps * wn
        assert np.abs(h - h1).max() <= 1e-13 * wn
        assert np.abs(Vd.T @ wsel.astype(np.float64)).max() <= 64 * eps * wn
        assert abs(nrm[0] - np.linalg.norm(wsel.astype(np.float64))) <= 16 * eps * nrm[0]


@pytest.mark.parametrize("s,q,n,prec", [(5, 7, 1000, 0), (16, 16, 1_000_003, 0), (11, 15, 257, 0),
                                        (64, 64, 300_001, 0), (17, 33, 70_001, 0),
                                        (64, 5, 1_000_003, 0), (97, 33, 70_000, 0),
                                        (128, 128, 40_000, 0), (130, 9, 5000, 0),
                                        (40, 24, 100_000, 1), (5, 7, 1000, 1),
                                        (16, 16, 1_000_003, 1), (64, 64, 300_001, 1),
                                        (97, 33, 70_000, 1), (128, 128, 40_000, 1),
                                        (130, 9, 5000, 1)])
def test_gemm_tn_blocks(ctx, s, q, n, prec):
    """V^T W of the block Gram-Schmidt steps: FP64 MFMA tiles for s, q <=
    128 (FP32 operands converted exactly; s, q <= 16 through the narrow
    256-row-slab kernel), FP64-accumulated VALU tiles otherwise; within
    1e-13 of numpy in FP64 on the same (rounded) operands."""
    print(f"[case] gemm_tn s={s} q={q} n={n} prec={prec}", flush=True)
    rng = np.random.default_rng(s * 1000 + q)
    V = rng.standard_normal((n, s))
    W = rng.standard_normal((n, q))
    if prec == 1:
        V, W = V.astype(np.float32), W.astype(np.float32)
    dV, dW = Dev(ctx, V), Dev(ctx, W)
    H = np.zeros((s, q), order="F")
    K._check(K.lib().kh_gemm_tn_dev(ctx.h, prec, dV.p, s, dW.p, q, n,
                                    H.ctypes.data_as(C.c_void_p)))
    V64, W64 = V.astype(np.float64), W.astype(np.float64)
    ref = V64.T @ W64
    scale = np.sqrt(np.outer(np.sum(V64 * V64, 0), np.sum(W64 * W64, 0)))
    assert np.all(np.abs(H - ref) <= 1e-13 * scale)


@pytest.mark.parametrize("s,q,n,prec,sub", [(5, 7, 1000, 0, 1), (16, 15, 1_000_003, 0, 1),
                                            (16, 16, 129, 0, 0), (64, 64, 300_001, 0, 1),
                                            (97, 33, 70_000, 0, 0), (64, 128, 20_000, 0, 1),
                                            (3, 1, 5000, 0, 1), (200, 40, 9000, 0, 1),
                                            (40, 24, 100_000, 1, 0), (96, 32, 50_000, 1, 1)])
def test_gemm_nn_blocks(ctx, s, q, n, prec, sub):
    """W = V H / W -= V H (FP64 MFMA for 2 <= q <= 128, the tile kernel
    otherwise); FP64 within 1e-13 of numpy, FP32 outputs within FP32 rounding."""
    print(f"[case] gemm_nn s={s} q={q} n={n} prec={prec} sub={sub}", flush=True)
    rng = np.random.default_rng(s * 7 + q)
    V = rng.standard_normal((n, s))
    H = np.asfortranarray(rng.standard_normal((s, q)))
    W = rng.standard_normal((n, q))
    if prec == 1:
        V, W = V.astype(np.float32), W.astype(np.float32)
    dV, dW = Dev(ctx, V), Dev(ctx, W)
    K._check(K.lib().kh_gemm_nn_dev(ctx.h, prec, dV.p, s, H.ctypes.data_as(C.c_void_p), q,
                                    dW.p, n, sub))
    got = dW.get().astype(np.float64)
    VH = V.astype(np.float64) @ H
    ref = W.astype(np.float64) - VH if sub else VH
    scale = np.abs(V.astype(np.float64)) @ np.abs(H) + (np.abs(W) if sub else 0)
    tol = 1e-13 if prec == 0 else 2 * np.finfo(np.float32).eps
    assert np.all(np.abs(got - ref) <= tol * scale + 1e-300)


def test_trace_shapes_and_vtw_timer(ctx):
    """The roofline plumbing: a traced CGS2 records the (rows, columns) shape
    of its V^T w launches, trace_end clears them, and the back-to-back timer
    returns a positive duration for that shape (and rejects s > 512)."""
    n, s = 100_003, 7
    print(f"[case] trace shapes n={n} s={s}", flush=True)
    rng = np.random.default_rng(5)
    V, _ = np.linalg.qr(rng.standard_normal((n, s)))
    dV, dw = Dev(ctx, V.astype(np.float32)), Dev(ctx, rng.standard_normal(n).astype(np.float32))
    h = np.zeros(2 * s)
    nrm = C.c_double()
    ctx.trace_begin("vtw")
    K._check(K.lib().kh_cgs2_dev(ctx.h, 1, dV.p, s, dw.p, n, h.ctypes.data_as(C.c_void_p),
                                 C.byref(nrm)))
    shapes = ctx.trace_shapes("vtw")
    launches, _, _ = ctx.trace_end()
    assert launches >= 1 and shapes and all(ln == n and sc == s for ln, sc, _ in shapes)
    assert sum(c for _, _, c in shapes) == launches
    assert ctx.trace_shapes("vtw") == []  # cleared by trace_end
    assert 0.0 < ctx.time_vtw(n, s, "fp32", reps=5) < 100.0
    with pytest.raises(ValueError):
        ctx.time_vtw(n, 513, "fp32")


@pytest.mark.parametrize("transpose", [0, 1])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_spmv_lds_window_bitwise(transpose, precision):
    """y = A x and y = A^T x of the incidence of 8 random 4-regular graphs x
    16,384 vertices (262,144 rows of 2 entries; 131,072 columns of 4): the
    LDS-window kernels (x staged per block of rows; pinned one at a time)
    give bitwise the gathers' y, which is the sequential sum; the first
    product after unpinning times the kernels and keeps one. Windows are one
    graph's x (64 KB FP32 / 128 KB FP64 for A, 128 KB FP32 for A^T); FP64
    A^T x needs 256 KB and has no usable window."""
    from krans_amd import workloads as W
    print(f"[case] spmv lds window t={transpose} {precision}", flush=True)
    A = W.regular_incidence(8, 16384, seed=5)
    H = K.HipCsrMatrixAdapter.from_csr(A)
    dt = np.float64 if precision == "fp64" else np.float32
    x = np.random.default_rng(3).standard_normal(A.m if transpose else A.n).astype(dt)
    mul = H.multiply_transpose if transpose else H.multiply
    assert H.spmv_kernel(transpose, precision, pin=1) == 1
    y1 = mul(x.astype(np.float64), precision).astype(dt)
    ref = _sequential_rows(A, x[:, None], dt, transpose)[:, 0]
    assert np.array_equal(y1, ref)
    ran = 0
    for path in (2, 3):
        try:
            H.spmv_kernel(transpose, precision, pin=path)
        except ValueError:  # no table, or a window beyond one workgroup's LDS
            continue
        yp = mul(x.astype(np.float64), precision).astype(dt)
        assert np.array_equal(yp, y1), path
        ran += 1
    assert ran >= (0 if (transpose and precision == "fp64") else 1)
    H.spmv_kernel(transpose, precision, pin=0)
    assert np.array_equal(mul(x.astype(np.float64), precision).astype(dt), y1)
    assert H.spmv_kernel(transpose, precision) in (1, 2, 3)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_same_bits_with_every_gather_kernel(precision):
    """A cyclic solve (y = A x - beta z steps, CG products) on the 8-graph
    incidence returns bitwise the same basis whichever gather kernel its
    products take."""
    from krans_amd import workloads as W
    print(f"[case] solve with pinned gather kernels {precision}", flush=True)
    A = W.regular_incidence(8, 16384, seed=5)
    opt = K.CyclicLanczosOptions(k=8, tol_rel=1e-12, tol_abs=0.0, prescreen_factor=1e-4,
                                 max_restarts=400)
    bases = []
    for path in (1, 2, 3):
        H = K.HipCsrMatrixAdapter.from_csr(A)
        pinned = 0
        for t in (0, 1):
            for p in ("fp64", "fp32"):
                try:
                    H.spmv_kernel(t, p, pin=path)
                    pinned += 1
                except ValueError:
                    H.spmv_kernel(t, p, pin=1)
        if path > 1 and pinned == 0:
            continue
        r = K.find_nullspace_cyclic_lanczos_ex(H, K.HipVectorFactory(A.n, 9), opt, precision)
        assert r.right_null.shape[1] == 8
        bases.append(r.right_null)
    assert len(bases) >= 2
    for U in bases[1:]:
        assert np.array_equal(U.view(np.uint64), bases[0].view(np.uint64))
