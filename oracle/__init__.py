"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure only: imported by tests/, by ``__graft_entry__.smoke()``
as the checker and by the CPU-baseline leg of bench.py. The product library
(krans_amd) never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")


class CyclicOpts(C.Structure):
    """Field for field CyclicLanczosOptions (CyclicLanczosSolver.hpp:84-146)."""
    _fields_ = [
        ("k", C.c_int64), ("ncv", C.c_int64), ("max_restarts", C.c_int64),
        ("tol_abs", C.c_double), ("tol_rel", C.c_double),
        ("reorth", C.c_int32), ("use_harmonic", C.c_int32),
        ("target_tau", C.c_double),
        ("nullity_known", C.c_int32), ("rank_known", C.c_int32),
        ("cg_maxit", C.c_int64),
        ("cg_rel_tol", C.c_double), ("cg_breakdown_tol", C.c_double),
        ("projector_passes", C.c_int32), ("grad_max_iters", C.c_int32),
        ("frob_samples", C.c_int32),
        ("prescreen_factor", C.c_double),
        ("stagnation_restarts", C.c_int64),
        ("return_left_space", C.c_int32), ("verbose", C.c_int32),
        ("precond_jacobi", C.c_int32), ("pad_", C.c_int32),
    ]


def cyclic_defaults(**kw):
    o = CyclicOpts(k=1, ncv=48, max_restarts=80, tol_abs=1e-12, tol_rel=1e-10,
                   reorth=2, use_harmonic=1, target_tau=0.0, nullity_known=-1,
                   rank_known=-1, cg_maxit=80, cg_rel_tol=1e-10,
                   cg_breakdown_tol=1e-30, projector_passes=1,
                   grad_max_iters=3000, frob_samples=4, prescreen_factor=0.8,
                   stagnation_restarts=30, return_left_space=0, verbose=0,
                   precond_jacobi=0)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class CyclicResult(C.Structure):
    _fields_ = [
        ("n_right", C.c_int64), ("n_left", C.c_int64), ("restarts", C.c_int64),
        ("krylov_steps_total", C.c_int64), ("frob_norm_est", C.c_double),
        ("converged_to_target", C.c_int64), ("cg_solves", C.c_int64),
        ("cg_maxit_hits", C.c_int64), ("spmv", C.c_int64), ("spmtv", C.c_int64),
    ]


class LobpcgOpts(C.Structure):
    """Field for field LOBPCGOptions (LOBPCGSolver.hpp:71-115)."""
    _fields_ = [
        ("nev", C.c_int64), ("block_size", C.c_int64), ("max_iters", C.c_int64),
        ("tol_abs", C.c_double), ("tol_rel", C.c_double),
        ("soft_locking", C.c_int32), ("gs_reorth", C.c_int32),
        ("restart", C.c_double), ("gs_tol", C.c_double),
        ("frob_samples", C.c_int32), ("nullity_known", C.c_int32),
        ("rank_known", C.c_int32), ("verbose", C.c_int32),
        ("max_tiny_dim", C.c_int64), ("progress_every", C.c_int64),
    ]


def lobpcg_defaults(**kw):
    o = LobpcgOpts(nev=1, block_size=0, max_iters=1000, tol_abs=1e-12,
                   tol_rel=1e-10, soft_locking=1, gs_reorth=1, restart=0.9,
                   gs_tol=1e-14, frob_samples=4, nullity_known=-1,
                   rank_known=-1, verbose=0, max_tiny_dim=0, progress_every=100)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class LobpcgResult(C.Structure):
    _fields_ = [
        ("n_right", C.c_int64), ("iters", C.c_int64), ("locked", C.c_int64),
        ("frob_norm_est", C.c_double), ("accept_threshold", C.c_double),
        ("best_unaccepted_residual", C.c_double), ("block_size", C.c_int64),
        ("converged_to_target", C.c_int64), ("spmv", C.c_int64),
        ("spmtv", C.c_int64),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        P, I64, I32, D = C.c_void_p, C.c_int64, C.c_int32, C.c_double
        L.orc_last_error.restype = C.c_char_p
        L.orc_op_csr.restype = P
        L.orc_op_csr.argtypes = [I64, I64, P, P, P]
        L.orc_op_shiftdiff.restype = P
        L.orc_op_shiftdiff.argtypes = [I64, I64, D]
        L.orc_op_free.argtypes = [P]
        L.orc_op_rows.restype = I64
        L.orc_op_rows.argtypes = [P]
        L.orc_op_cols.restype = I64
        L.orc_op_cols.argtypes = [P]
        L.orc_op_csc.argtypes = [P, P, P, P, P]
        L.orc_mv.argtypes = [P, P, P]
        L.orc_mtv.argtypes = [P, P, P]
        L.orc_factory.restype = P
        L.orc_factory.argtypes = [C.c_uint32]
        L.orc_factory_free.argtypes = [P]
        L.orc_randn.argtypes = [P, I64, P]
        L.orc_mt19937_first.restype = C.c_uint32
        L.orc_mt19937_first.argtypes = [C.c_uint32]
        L.orc_cyclic.argtypes = [P, P, C.POINTER(CyclicOpts), P, P, I64, P, I64,
                                 C.POINTER(CyclicResult)]
        L.orc_lobpcg.argtypes = [P, P, C.POINTER(LobpcgOpts), P, I64, P, P,
                                 C.POINTER(LobpcgResult)]
        L.orc_svd.argtypes = [I32, I32, P, P, P]
        L.orc_sym_eig.argtypes = [I32, P, P, P]
        L.orc_cholesky.argtypes = [I32, P, P]
        L.orc_mgs.argtypes = [I64, I64, P, D, I32]
        L.orc_orth_against.restype = D
        L.orc_orth_against.argtypes = [I64, I64, P, P, D, I32]
        L.orc_subspace_diff.restype = D
        L.orc_subspace_diff.argtypes = [I64, I64, P, I64, P, D]
        L.orc_load_ija.restype = P
        L.orc_load_ija.argtypes = [C.c_char_p, I32, I32]
        L.orc_load_mtx.restype = P
        L.orc_load_mtx.argtypes = [C.c_char_p]
        L.orc_triplets.restype = P
        L.orc_triplets.argtypes = [I64, I64, I64, P, P, P]
        L.orc_csc_info.argtypes = [P, C.POINTER(I64), C.POINTER(I64), C.POINTER(I64),
                                   C.POINTER(I64)]
        L.orc_csc_copy.argtypes = [P, P, P, P]
        L.orc_csc_free.argtypes = [P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Operator:
    """An oracle operator: CSR (with CSC twin) or the matrix-free stencil."""

    def __init__(self, handle, keep=()):
        self.h = handle
        self._keep = keep
        self.m = lib().orc_op_rows(handle)
        self.n = lib().orc_op_cols(handle)

    @classmethod
    def csr(cls, m, n, rowptr, col, val):
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        col = np.ascontiguousarray(col, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        return cls(lib().orc_op_csr(m, n, _p(rowptr), _p(col), _p(val)))

    @classmethod
    def shiftdiff(cls, n, N, c):
        return cls(lib().orc_op_shiftdiff(n, N, c))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_op_free(self.h)
            self.h = None

    def csc(self, nnz):
        cp = np.empty(self.n + 1, np.int32)
        ri = np.empty(nnz, np.int32)
        vt = np.empty(nnz, np.float64)
        perm = np.empty(nnz, np.int32)
        if lib().orc_op_csc(self.h, _p(cp), _p(ri), _p(vt), _p(perm)) != 0:
            raise ValueError(lib().orc_last_error().decode())
        return cp, ri, vt, perm

    def mv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.m)
        lib().orc_mv(self.h, _p(x), _p(y))
        return y

    def mtv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n)
        lib().orc_mtv(self.h, _p(x), _p(y))
        return y


class Factory:
    """StdVectorFactory(n, seed): mt19937 + normal_distribution<double>."""

    def __init__(self, seed):
        self.h = lib().orc_factory(seed)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_factory_free(self.h)
            self.h = None

    def randn(self, n):
        out = np.empty(n)
        lib().orc_randn(self.h, n, _p(out))
        return out


def _check(rc):
    if rc == 1:
        raise ValueError(lib().orc_last_error().decode())
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())


def cyclic(op, factory, opts, y0=None):
    """Returns (right basis n x k, left basis m x k', CyclicResult)."""
    k = max(int(opts.k), int(opts.nullity_known), 1)
    R = np.zeros((k, op.n))
    L = np.zeros((k, op.m))
    res = CyclicResult()
    y0a = None if y0 is None else np.ascontiguousarray(y0, dtype=np.float64)
    _check(lib().orc_cyclic(op.h, factory.h, C.byref(opts),
                            None if y0a is None else _p(y0a),
                            _p(R), k, _p(L), k, C.byref(res)))
    return R[:res.n_right].T.copy(), L[:res.n_left].T.copy(), res


def lobpcg(op, factory, opts):
    k = max(int(opts.nev), int(opts.nullity_known), 1, op.n if opts.rank_known >= 0 else 0)
    k = min(k, op.n)
    R = np.zeros((k, op.n))
    th = np.zeros(k)
    rs = np.zeros(k)
    res = LobpcgResult()
    _check(lib().orc_lobpcg(op.h, factory.h, C.byref(opts), _p(R), k, _p(th),
                            _p(rs), C.byref(res)))
    n = res.n_right
    return R[:n].T.copy(), th[:n].copy(), rs[:n].copy(), res


def svd(A):
    A = np.asfortranarray(A, dtype=np.float64)
    r, c = A.shape
    sv = np.empty(c)
    V = np.empty((c, c), order="F")
    lib().orc_svd(r, c, _p(A), _p(sv), _p(V))
    return sv, V


def sym_eig(A):
    A = np.asfortranarray(A, dtype=np.float64)
    n = A.shape[0]
    w = np.empty(n)
    Q = np.empty((n, n), order="F")
    lib().orc_sym_eig(n, _p(A), _p(w), _p(Q))
    return w, Q


def cholesky(A):
    A = np.asfortranarray(A, dtype=np.float64)
    n = A.shape[0]
    L = np.zeros((n, n), order="F")
    if lib().orc_cholesky(n, _p(A), _p(L)) != 0:
        return None
    return L


def mgs(V, tol=1e-12, reorth=True):
    """Modified Gram-Schmidt on the columns of V; returns (ok, Q)."""
    Q = np.asfortranarray(V, dtype=np.float64).copy(order="F")
    ok = lib().orc_mgs(Q.shape[0], Q.shape[1], _p(Q), tol, int(reorth))
    return bool(ok), Q


def orth_against(B, v, tol=1e-12, reorth=True):
    B = np.asfortranarray(B, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64).copy()
    r = lib().orc_orth_against(B.shape[0], B.shape[1], _p(B), _p(v), tol, int(reorth))
    return r, v


def subspace_diff(B1, B2, tol=1e-12):
    """RREF Frobenius distance (SubspaceCompare.hpp:86-110); bases as columns."""
    B1 = np.asfortranarray(B1, dtype=np.float64)
    B2 = np.asfortranarray(B2, dtype=np.float64)
    return lib().orc_subspace_diff(B1.shape[0], B1.shape[1], _p(B1),
                                   B2.shape[1], _p(B2), tol)


def _take_csc(h):
    if not h:
        raise RuntimeError(lib().orc_last_error().decode())
    L = lib()
    try:
        r, c, nnz, er = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        L.orc_csc_info(h, C.byref(r), C.byref(c), C.byref(nnz), C.byref(er))
        colptr = np.zeros(c.value + 1, np.int64)
        rowidx = np.zeros(nnz.value, np.int32)
        val = np.zeros(nnz.value)
        L.orc_csc_copy(h, _p(colptr), _p(rowidx), _p(val))
        return (r.value, c.value, colptr, rowidx, val, er.value)
    finally:
        L.orc_csc_free(h)


def load_ija(path, rows=-1, cols=-1):
    """load_ija_triplet_matrix restated (src/IO/MatrixLoader.cpp:21-56):
    returns (rows, cols, colptr, rowidx, val, entries_read), Eigen's CSC."""
    return _take_csc(lib().orc_load_ija(os.fsencode(path), rows, cols))


def load_mtx(path):
    """load_matrix_market restated (MatrixLoader.cpp:12-18, Eigen loadMarket)."""
    return _take_csc(lib().orc_load_mtx(os.fsencode(path)))


def from_triplets(rows, cols, I, J, V):
    """setFromTriplets restated: Eigen's CSC of 0-based triplets."""
    I = np.ascontiguousarray(I, np.int32)
    J = np.ascontiguousarray(J, np.int32)
    V = np.ascontiguousarray(V, np.float64)
    return _take_csc(lib().orc_triplets(rows, cols, I.size, _p(I), _p(J), _p(V)))
