"""KraNS-MI: MI355X-native null-space engine (Python host binding).

Mirrors the reference's plugin API for the null-space hot path:

* ``HipCsrMatrixAdapter``  -- MatrixInterface over a device CSR with CSC twin
  (replaces EigenSparseMatrixAdapter, include/krans/adapters/eigen/
  EigenSparseMatrixAdapter.hpp:12-68);
* ``HipVectorFactory``     -- VectorFactory with the reference's seeded
  std::mt19937 + normal_distribution stream (StdVectorFactory.hpp:16-58);
* ``CyclicLanczosOptions`` / ``LOBPCGOptions`` with the reference's fields and
  defaults (CyclicLanczosSolver.hpp:84-146, LOBPCGSolver.hpp:71-115);
* ``find_nullspace_cyclic_lanczos[_ex]`` / ``find_nullspace_lobpcg[_ex]``.

Everything calls libkrans_hip.so through its C ABI (include/krans_hip.h).
There is no CPU fallback: importing the solvers without the built library
raises.
"""
import ctypes as C
import os
from dataclasses import dataclass, field, fields
from typing import List, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libkrans_hip.so")

KH_OK, KH_EINVAL, KH_ERUNTIME, KH_ENOMEM, KH_EDEVICE = 0, 1, 2, 3, 4
PREC = {"fp64": 0, "mixed": 1, "fp32": 1}


class _CyclicOpts(C.Structure):
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
        ("precond_jacobi", C.c_int32), ("graphs", C.c_int32), ("screen", C.c_int32),
        ("second_pass", C.c_int32), ("second_pass_eta", C.c_double),
    ]


class _CyclicRes(C.Structure):
    _fields_ = [
        ("n_right", C.c_int64), ("n_left", C.c_int64), ("restarts", C.c_int64),
        ("krylov_steps_total", C.c_int64), ("frob_norm_est", C.c_double),
        ("converged_to_target", C.c_int64), ("cg_solves", C.c_int64),
        ("cg_maxit_hits", C.c_int64), ("spmv", C.c_int64), ("spmtv", C.c_int64),
        ("cg_iters", C.c_int64),
        ("t_gk", C.c_double), ("t_tiny", C.c_double), ("t_cand", C.c_double),
        ("t_polish", C.c_double), ("t_fallback", C.c_double), ("t_rng", C.c_double),
        ("screened", C.c_int64), ("graph_captures", C.c_int64), ("graph_replays", C.c_int64),
        ("refined", C.c_int64), ("packed", C.c_int64), ("escalated", C.c_int64),
        ("cgs_pairs", C.c_int64), ("second_passes", C.c_int64), ("gk_requeued", C.c_int64),
    ]


class _LobpcgOpts(C.Structure):
    _fields_ = [
        ("nev", C.c_int64), ("block_size", C.c_int64), ("max_iters", C.c_int64),
        ("tol_abs", C.c_double), ("tol_rel", C.c_double),
        ("soft_locking", C.c_int32), ("gs_reorth", C.c_int32),
        ("restart", C.c_double), ("gs_tol", C.c_double),
        ("frob_samples", C.c_int32), ("nullity_known", C.c_int32),
        ("rank_known", C.c_int32), ("verbose", C.c_int32),
        ("max_tiny_dim", C.c_int64), ("progress_every", C.c_int64),
    ]


class _LobpcgRes(C.Structure):
    _fields_ = [
        ("n_right", C.c_int64), ("iters", C.c_int64), ("locked", C.c_int64),
        ("frob_norm_est", C.c_double), ("accept_threshold", C.c_double),
        ("best_unaccepted_residual", C.c_double), ("block_size", C.c_int64),
        ("converged_to_target", C.c_int64), ("spmv", C.c_int64), ("spmtv", C.c_int64),
        ("t_gram", C.c_double), ("t_sparse", C.c_double), ("t_host", C.c_double),
        ("cholqr_blocks", C.c_int64), ("cholqr_fallbacks", C.c_int64),
    ]


# Every exported symbol of include/krans_hip.h with its ctypes signature.
P, I32, I64, D, SZ = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_size_t
SIGNATURES = {
    "kh_last_error": (C.c_char_p, []),
    "kh_version": (I32, []),
    "kh_device_count": (I32, [C.POINTER(C.c_int)]),
    "kh_context_create": (I32, [I32, C.POINTER(P)]),
    "kh_context_destroy": (I32, [P]),
    "kh_context_sync": (I32, [P]),
    "kh_matrix_create_csr": (I32, [P, I64, I64, P, P, P, C.POINTER(P)]),
    "kh_matrix_create_shifted_difference": (I32, [P, I64, I64, D, C.POINTER(P)]),
    "kh_matrix_create_operator": (I32, [P, I64, I64, P, P, C.POINTER(P)]),
    "kh_load_ija": (I32, [C.c_char_p, I64, I64, C.POINTER(P)]),
    "kh_load_matrix_market": (I32, [C.c_char_p, C.POINTER(P)]),
    "kh_csr_from_triplets": (I32, [I64, I64, I64, P, P, P, C.POINTER(P)]),
    "kh_host_csr_info": (I32, [P, C.POINTER(I64), C.POINTER(I64), C.POINTER(I64),
                               C.POINTER(I64), C.POINTER(I64)]),
    "kh_host_csr_data": (I32, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P)]),
    "kh_host_csr_free": (I32, [P]),
    "kh_matrix_destroy": (I32, [P]),
    "kh_matrix_info": (I32, [P, C.POINTER(I64), C.POINTER(I64), C.POINTER(I64), C.POINTER(D)]),
    "kh_matrix_download_csr": (I32, [P, P, P, P]),
    "kh_matrix_download_csc": (I32, [P, P, P, P, P]),
    "kh_matrix_multiply": (I32, [P, I32, I32, P, P]),
    "kh_rng_create": (I32, [C.c_uint32, C.POINTER(P)]),
    "kh_rng_destroy": (I32, [P]),
    "kh_rng_normal": (I32, [P, I64, P]),
    "kh_tiny_svd_right": (I32, [P, I32, I32, I32, P, P]),
    "kh_cyclic_lanczos": (I32, [P, P, C.POINTER(_CyclicOpts), I32, P, P, I64, P, I64,
                                C.POINTER(_CyclicRes)]),
    "kh_cyclic_lanczos_dev": (I32, [P, P, C.POINTER(_CyclicOpts), I32, P, P, I64, P, I64,
                                    C.POINTER(_CyclicRes)]),
    "kh_lobpcg": (I32, [P, P, C.POINTER(_LobpcgOpts), I32, P, I64, P, I64, P, P,
                        C.POINTER(_LobpcgRes)]),
    "kh_server_create": (I32, [I32, I32, D, C.POINTER(P)]),
    "kh_server_destroy": (I32, [P]),
    "kh_server_submit_cyclic": (I32, [P, I64, I64, P, P, P, C.POINTER(_CyclicOpts), I32,
                                      C.c_uint32, C.POINTER(I64)]),
    "kh_server_submit_resident": (I32, [P, P, C.POINTER(_CyclicOpts), I32, C.c_uint32,
                                        C.POINTER(I64)]),
    "kh_server_wait": (I32, [P, I64, P, I64, C.POINTER(_CyclicRes), C.POINTER(D),
                             C.POINTER(D)]),
    "kh_server_stats": (I32, [P, C.POINTER(I64), C.POINTER(I64), C.POINTER(D), C.POINTER(D),
                              C.POINTER(D)]),
    "kh_server_reset_clock": (I32, [P]),
    "kh_server_set_packing": (I32, [P, I32, I64]),
    "kh_dev_alloc": (I32, [P, SZ, C.POINTER(P)]),
    "kh_dev_free": (I32, [P, P]),
    "kh_h2d": (I32, [P, P, P, SZ]),
    "kh_d2h": (I32, [P, P, P, SZ]),
    "kh_vec_axpby": (I32, [P, I64, D, P, D, P]),
    "kh_matrix_jacobi_AAt": (I32, [P, P, P]),
    "kh_spmm_dev": (I32, [P, I32, I32, I32, P, P]),
    "kh_spmm_rows_dev": (I32, [P, I32, I32, I32, P, P]),
    "kh_gemm_tn_dev": (I32, [P, I32, P, I32, P, I32, I64, P]),
    "kh_gemm_nn_dev": (I32, [P, I32, P, I32, P, I32, P, I64, I32]),
    "kh_dot_dev": (I32, [P, I32, P, P, I64, C.POINTER(D)]),
    "kh_cgs2_dev": (I32, [P, I32, P, I32, P, I64, P, C.POINTER(D)]),
    "kh_cgs2_dev_noalloc": (I32, [P, I32, P, I32, P, I64, P, P]),
    "kh_cgs2_dev_ex": (I32, [P, I32, P, I32, P, I64, I32, P, P]),
    "kh_cgs2_dev_select": (I32, [P, I32, P, I32, P, I64, P, P, P]),
    "kh_gram_dev": (I32, [P, I32, P, P, I32, I64, P, P]),
    "kh_time_spmv": (I32, [P, I32, I32, I32, C.POINTER(D)]),
    "kh_spmv_kernel": (I32, [P, I32, I32, I32, C.POINTER(I32)]),
    "kh_trace_begin": (I32, [P, I32]),
    "kh_trace_end": (I32, [P, C.POINTER(I64), C.POINTER(D), C.POINTER(D)]),
    "kh_trace_begin_mask": (I32, [P, C.c_uint]),
    "kh_trace_class": (I32, [P, I32, C.POINTER(I64), C.POINTER(D), C.POINTER(D)]),
    "kh_trace_shapes": (I32, [P, I32, I32, C.POINTER(I64), C.POINTER(I32), C.POINTER(I64),
                              C.POINTER(I32)]),
    "kh_time_vtw": (I32, [P, I32, I64, I32, I32, C.POINTER(D)]),
}

_lib = None


def lib():
    """Load libkrans_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libkrans_hip.so not built ({LIB_PATH}); run python -m krans_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class KransError(RuntimeError):
    pass


def _check(rc):
    if rc == KH_OK:
        return
    msg = lib().kh_last_error().decode(errors="replace")
    if rc == KH_EINVAL:
        raise ValueError(msg)
    if rc == KH_ENOMEM:
        raise MemoryError(msg)
    raise KransError(msg)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class HipContext:
    """One device, one stream (kh_context)."""

    def __init__(self, device=0):
        h = P()
        _check(lib().kh_context_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().kh_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _check(lib().kh_context_sync(self.h))

    TRACE_CLASSES = {"vtw": 1, "vh_sub": 2, "spmv": 3, "gram": 4, "other": 5}

    def trace_begin(self, kernel_class):
        """Time every launch of a kernel class with HIP events (see kh_trace_begin)."""
        _check(lib().kh_trace_begin(self.h, self.TRACE_CLASSES[kernel_class]))

    def trace_end(self):
        """Returns (launches, algorithmic bytes, device ms) since trace_begin."""
        n, b, ms = I64(), D(), D()
        _check(lib().kh_trace_end(self.h, C.byref(n), C.byref(b), C.byref(ms)))
        return n.value, b.value, ms.value

    def trace_begin_classes(self, classes):
        """Trace several kernel classes at once (kh_trace_begin_mask)."""
        mask = 0
        for k in classes:
            mask |= 1 << self.TRACE_CLASSES[k]
        _check(lib().kh_trace_begin_mask(self.h, mask))

    def trace_shapes(self, kernel_class, max_shapes=4096):
        """[(len, s, launches)] of the traced class's launches (kh_trace_shapes);
        call before trace_end."""
        L = (I64 * max_shapes)()
        S = (I32 * max_shapes)()
        N = (I64 * max_shapes)()
        n = I32()
        _check(lib().kh_trace_shapes(self.h, self.TRACE_CLASSES[kernel_class], max_shapes, L, S,
                                     N, C.byref(n)))
        return [(L[i], S[i], N[i]) for i in range(n.value)]

    def time_vtw(self, length, s, precision="fp32", reps=20):
        """Average device ms of one V^T w partial launch at length x s,
        back-to-back (kh_time_vtw)."""
        out = D()
        _check(lib().kh_time_vtw(self.h, PREC[precision], int(length), int(s), int(reps),
                                 C.byref(out)))
        return out.value

    def trace_classes_end(self, classes):
        """{class: (launches, algorithmic bytes or flops, device ms)}, then
        clears the trace."""
        out = {}
        for k in classes:
            n, b, ms = I64(), D(), D()
            _check(lib().kh_trace_class(self.h, self.TRACE_CLASSES[k], C.byref(n), C.byref(b),
                                        C.byref(ms)))
            out[k] = (n.value, b.value, ms.value)
        self.trace_end()
        return out


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = HipContext(0)
    return _default_ctx


class HipMatrix:
    """MatrixInterface over a kh_matrix handle (a device CSR matrix or a
    matrix-free device operator); the solvers accept any HipMatrix."""

    def _adopt(self, h, ctx):
        self.ctx = ctx
        self.h = h
        mm, nn, nz, fr = I64(), I64(), I64(), D()
        _check(lib().kh_matrix_info(h, C.byref(mm), C.byref(nn), C.byref(nz), C.byref(fr)))
        self.m, self.n, self.nnz, self.fro = mm.value, nn.value, nz.value, fr.value

    def rows(self):
        return self.m

    def cols(self):
        return self.n

    def close(self):
        if getattr(self, "h", None):
            lib().kh_matrix_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

class HipCsrMatrixAdapter(HipMatrix):
    """MatrixInterface over a device CSR matrix with its CSC twin."""

    def __init__(self, m, n, rowptr, col, val, ctx=None):
        ctx = ctx or default_context()
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        col = np.ascontiguousarray(col, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        if rowptr.shape[0] != m + 1:
            raise ValueError("rowptr must have m+1 entries")
        h = P()
        _check(lib().kh_matrix_create_csr(ctx.h, m, n, _ptr(rowptr), _ptr(col),
                                          _ptr(val), C.byref(h)))
        self._adopt(h, ctx)

    @classmethod
    def from_csr(cls, A, ctx=None):
        """From any object with m, n, rowptr, col, val (e.g. workloads.CsrMatrix)."""
        return cls(A.m, A.n, A.rowptr, A.col, A.val, ctx=ctx)

    def multiply(self, x, precision="fp64"):
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.shape != (self.n,):
            raise ValueError("HipMatrix::multiply size mismatch")
        y = np.empty(self.m)
        _check(lib().kh_matrix_multiply(self.h, 0, PREC[precision], _ptr(x), _ptr(y)))
        return y

    def multiply_transpose(self, x, precision="fp64"):
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.shape != (self.m,):
            raise ValueError("HipMatrix::multiply_transpose size mismatch")
        y = np.empty(self.n)
        _check(lib().kh_matrix_multiply(self.h, 1, PREC[precision], _ptr(x), _ptr(y)))
        return y

    def download_csr(self):
        rp = np.empty(self.m + 1, np.int32)
        ci = np.empty(self.nnz, np.int32)
        va = np.empty(self.nnz, np.float64)
        _check(lib().kh_matrix_download_csr(self.h, _ptr(rp), _ptr(ci), _ptr(va)))
        return rp, ci, va

    def download_csc(self):
        cp = np.empty(self.n + 1, np.int32)
        ri = np.empty(self.nnz, np.int32)
        vt = np.empty(self.nnz, np.float64)
        perm = np.empty(self.nnz, np.int32)
        _check(lib().kh_matrix_download_csc(self.h, _ptr(cp), _ptr(ri), _ptr(vt), _ptr(perm)))
        return cp, ri, vt, perm

    def time_spmv(self, transpose=False, precision="fp64", reps=50):
        out = D()
        _check(lib().kh_time_spmv(self.h, int(transpose), PREC[precision], reps, C.byref(out)))
        return out.value

    def spmv_kernel(self, transpose=False, precision="fp64", pin=None):
        """The gather kernel of y = op(A) x (0 not chosen yet, 1 gathers
        through L2, 2/3 x staged in LDS per block of rows); pin= sets it
        first (0: time again at the next product)."""
        v = I32(0 if pin is None else int(pin))
        _check(lib().kh_spmv_kernel(self.h, int(transpose), PREC[precision],
                                    0 if pin is None else 1, C.byref(v)))
        return v.value


for _name in ("multiply", "multiply_transpose", "time_spmv"):
    setattr(HipMatrix, _name, getattr(HipCsrMatrixAdapter, _name))


class HipShiftedDifferenceOperator(HipMatrix):
    """The quick start's matrix-free operator (examples/quickstart.cpp:25-37):
    (A x)_i = x_{i+N} - c x_i, A is (n-N) x n, applied by HIP kernels with no
    matrix in memory (kh_matrix_create_shifted_difference)."""

    def __init__(self, n, N, c, ctx=None):
        ctx = ctx or default_context()
        h = P()
        _check(lib().kh_matrix_create_shifted_difference(ctx.h, int(n), int(N), float(c),
                                                         C.byref(h)))
        self._adopt(h, ctx)


APPLY_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                       C.c_int64, C.c_void_p, C.c_int64)


class HipOperatorAdapter(HipMatrix):
    """A user operator (kh_matrix_create_operator), the device analogue of
    StdMatrixAdapter: apply(stream, precision, transpose, s, x, ldx, y, ldy)
    with device pointers (ints) computes y(:, j) = op(A) x(:, j), j < s, for
    precision "fp64" (double columns) or "mixed" (float columns), finished
    or enqueued on `stream` before it returns. An exception reports failure."""

    def __init__(self, m, n, apply, ctx=None):
        ctx = ctx or default_context()
        self._apply = apply

        def tramp(user, stream, precision, transpose, s, x, ldx, y, ldy):
            try:
                apply(stream, "fp64" if precision == 0 else "mixed", bool(transpose), s, x, ldx,
                      y, ldy)
                return 0
            except Exception as e:  # reported as KH_ERUNTIME by the solver
                self.last_error = e
                return 1
        self._tramp = APPLY_FN(tramp)  # kept alive with the matrix
        h = P()
        _check(lib().kh_matrix_create_operator(ctx.h, int(m), int(n), self._tramp, None,
                                               C.byref(h)))
        self._adopt(h, ctx)


class HipVectorFactory:
    """VectorFactory: the reference's seeded N(0,1) stream; vectors of the
    solvers live on the device, this object owns the random stream."""

    def __init__(self, default_size, seed=0):
        if seed == 0:
            seed = int.from_bytes(os.urandom(4), "little") or 1
        self.default_size = default_size
        self.seed = seed
        h = P()
        _check(lib().kh_rng_create(seed, C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().kh_rng_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def create_random_of_size(self, n):
        out = np.empty(n)
        _check(lib().kh_rng_normal(self.h, n, _ptr(out)))
        return out

    def create_random(self):
        return self.create_random_of_size(self.default_size)

    def create_zero_of_size(self, n):
        return np.zeros(n)

    def create_zero(self):
        return np.zeros(self.default_size)


@dataclass
class CyclicLanczosOptions:
    """CyclicLanczosSolver.hpp:84-146 (same names and defaults)."""
    k: int = 1
    ncv: int = 48
    max_restarts: int = 80
    tol_abs: float = 1e-12
    tol_rel: float = 1e-10
    reorth: int = 2  # ReorthMode: 0 None, 1 Single, 2 Double
    use_harmonic: bool = True
    target_tau: float = 0.0
    nullity_known: int = -1
    rank_known: int = -1
    initial_right_space: List[np.ndarray] = field(default_factory=list)
    precond_AAt: Optional[str] = None  # None or "jacobi" (diag(A A^T)^-1)
    cg_maxit: int = 80
    cg_rel_tol: float = 1e-10
    cg_breakdown_tol: float = 1e-30
    projector_passes: int = 1
    grad_max_iters: int = 3000
    frob_samples: int = 4
    prescreen_factor: float = 0.8
    stagnation_restarts: int = 30
    return_left_space: bool = False
    verbose: bool = False
    # engine choice (not in the reference struct): Golub-Kahan cycles from
    # captured HIP graphs where the operator allows it: 1 always, 0 by problem
    # size, -1 never
    graphs: int = 0
    # engine choice (mixed precision, prescreen_factor < 0.1): candidates
    # screened from the projected problem's singular values (0, when the
    # target is zero), from FP32 candidate vectors (1), or not at all (-1)
    screen: int = 0
    # engine choice (reorth 2): second Gram-Schmidt pass of each Golub-Kahan
    # step always (1), only where the first pass cancelled more than half of
    # ||w||^2 (2, the DGKS test decided on the device), or automatic (0: 2 for
    # mixed precision, 1 for FP64)
    second_pass: int = 0
    # eta of that test (0 < eta <= 1; 1 asks for every second pass)
    second_pass_eta: float = 0.7071067811865476

    def _c(self):
        o = _CyclicOpts()
        for f in fields(self):
            if f.name in ("initial_right_space", "precond_AAt"):
                continue
            setattr(o, f.name, int(getattr(self, f.name)) if isinstance(getattr(self, f.name), bool)
                    else getattr(self, f.name))
        if self.precond_AAt not in (None, "jacobi"):
            raise ValueError("precond_AAt must be None or 'jacobi' on the HIP backend")
        o.precond_jacobi = 1 if self.precond_AAt == "jacobi" else 0
        return o


@dataclass
class CyclicLanczosResult:
    """CyclicLanczosSolver.hpp:149-167 plus counters and phase times."""
    right_null: np.ndarray
    left_null: np.ndarray
    restarts: int
    krylov_steps_total: int
    frob_norm_est: float
    converged_to_target: bool
    cg_solves: int
    cg_maxit_hits: int
    spmv: int = 0
    spmtv: int = 0
    cg_iters: int = 0
    times: dict = field(default_factory=dict)
    screened: int = 0  # mixed precision: candidates rejected by the FP32 screen
    graph_captures: int = 0  # Golub-Kahan cycles captured as HIP graphs
    graph_replays: int = 0   # ... and replayed
    refined: int = 0  # mixed: locked vectors replaced by the final FP64 projector pass
    escalated: int = 0  # mixed: FP32 projector CGs that did not converge, solved again in FP64
    cgs_pairs: int = 0  # Gram-Schmidt pairs of the Golub-Kahan steps
    second_passes: int = 0  # ... that ran their second pass
    gk_requeued: int = 0  # cycles queued again from a step that needed its second pass
    n_right: int = -1  # columns written to the device block (find_nullspace_cyclic_lanczos_dev)


def find_nullspace_cyclic_lanczos_ex(A, vf, opt, precision="fp64"):
    """find_nullspace_cyclic_lanczos_ex (CyclicLanczosSolver.hpp:170-174).

    Returns right_null as an (n x k) array of orthonormal columns.
    """
    if not isinstance(A, HipMatrix):
        raise ValueError("find_nullspace_cyclic_lanczos: backend mismatch (HIP matrix required)")
    o = opt._c()
    kcap = max(int(opt.k), int(opt.nullity_known), 1)
    R = np.zeros((kcap, A.n))
    Lb = np.zeros((kcap, A.m)) if opt.return_left_space else None
    y0 = None
    if opt.initial_right_space:
        y0 = np.ascontiguousarray(opt.initial_right_space[0], dtype=np.float64)
        if y0.shape != (A.n,):
            raise ValueError("initial_right_space[0] size mismatch")
    res = _CyclicRes()
    _check(lib().kh_cyclic_lanczos(A.h, vf.h, C.byref(o), PREC[precision],
                                   None if y0 is None else _ptr(y0), _ptr(R), kcap,
                                   None if Lb is None else _ptr(Lb), kcap, C.byref(res)))
    times = {k: getattr(res, "t_" + k) for k in ("gk", "tiny", "cand", "polish", "fallback", "rng")}
    return CyclicLanczosResult(
        right_null=R[:res.n_right].T.copy(),
        left_null=(Lb[:res.n_left].T.copy() if Lb is not None else np.zeros((A.m, 0))),
        restarts=res.restarts, krylov_steps_total=res.krylov_steps_total,
        frob_norm_est=res.frob_norm_est, converged_to_target=bool(res.converged_to_target),
        cg_solves=res.cg_solves, cg_maxit_hits=res.cg_maxit_hits, spmv=res.spmv,
        spmtv=res.spmtv, cg_iters=res.cg_iters, times=times, screened=res.screened,
        graph_captures=res.graph_captures, graph_replays=res.graph_replays,
        refined=res.refined, escalated=res.escalated, cgs_pairs=res.cgs_pairs,
        second_passes=res.second_passes, gk_requeued=res.gk_requeued)


class DeviceBlock:
    """Column-major block of `cols` FP64 device vectors of length `rows` on a
    context's device (kh_dev_alloc): the result storage of
    find_nullspace_cyclic_lanczos_dev."""

    def __init__(self, ctx, rows, cols):
        self.ctx, self.rows, self.cols = ctx, int(rows), int(cols)
        p = P()
        _check(lib().kh_dev_alloc(ctx.h, self.rows * self.cols * 8, C.byref(p)))
        self.ptr = p

    def to_host(self, count=None):
        """The first `count` columns as an (rows x count) array."""
        count = self.cols if count is None else int(count)
        out = np.empty((count, self.rows))
        if count:
            _check(lib().kh_d2h(self.ctx.h, _ptr(out), self.ptr, count * self.rows * 8))
        return out.T.copy()

    def free(self):
        if getattr(self, "ptr", None):
            lib().kh_dev_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def find_nullspace_cyclic_lanczos_dev(A, vf, opt, out, precision="fp64"):
    """find_nullspace_cyclic_lanczos_ex with the right basis left on the
    device (kh_cyclic_lanczos_dev): its columns are written to the
    DeviceBlock `out` (A.n rows, at least k columns); the result's right_null
    is empty and n_right counts the columns written."""
    if not isinstance(A, HipMatrix):
        raise ValueError("find_nullspace_cyclic_lanczos: backend mismatch (HIP matrix required)")
    if opt.return_left_space:
        raise ValueError("find_nullspace_cyclic_lanczos_dev returns the right space only")
    if out.rows != A.n:
        raise ValueError("output block rows must equal A.n")
    o = opt._c()
    y0 = None
    if opt.initial_right_space:
        y0 = np.ascontiguousarray(opt.initial_right_space[0], dtype=np.float64)
        if y0.shape != (A.n,):
            raise ValueError("initial_right_space[0] size mismatch")
    res = _CyclicRes()
    _check(lib().kh_cyclic_lanczos_dev(A.h, vf.h, C.byref(o), PREC[precision],
                                       None if y0 is None else _ptr(y0), out.ptr, out.cols,
                                       None, 0, C.byref(res)))
    times = {k: getattr(res, "t_" + k) for k in ("gk", "tiny", "cand", "polish", "fallback", "rng")}
    r = CyclicLanczosResult(
        right_null=np.zeros((A.n, 0)), left_null=np.zeros((A.m, 0)),
        restarts=res.restarts, krylov_steps_total=res.krylov_steps_total,
        frob_norm_est=res.frob_norm_est, converged_to_target=bool(res.converged_to_target),
        cg_solves=res.cg_solves, cg_maxit_hits=res.cg_maxit_hits, spmv=res.spmv,
        spmtv=res.spmtv, cg_iters=res.cg_iters, times=times, screened=res.screened,
        graph_captures=res.graph_captures, graph_replays=res.graph_replays,
        refined=res.refined, escalated=res.escalated, cgs_pairs=res.cgs_pairs,
        second_passes=res.second_passes, gk_requeued=res.gk_requeued, n_right=res.n_right)
    return r


def find_nullspace_cyclic_lanczos(A, vf, opt, precision="fp64"):
    """Backwards-compatible wrapper (CyclicLanczosSolver.hpp:176-178)."""
    return find_nullspace_cyclic_lanczos_ex(A, vf, opt, precision).right_null


@dataclass
class LOBPCGOptions:
    """LOBPCGSolver.hpp:71-115 (same names and defaults)."""
    nev: int = 1
    block_size: int = 0
    max_iters: int = 1000
    tol_abs: float = 1e-12
    tol_rel: float = 1e-10
    soft_locking: bool = True
    restart: float = 0.9
    gs_tol: float = 1e-14
    gs_reorth: bool = True
    M_inv: Optional[str] = None
    initial_block: List[np.ndarray] = field(default_factory=list)
    frob_samples: int = 4
    max_tiny_dim: int = 0
    nullity_known: int = -1
    rank_known: int = -1
    verbose: bool = False
    progress_every: int = 100

    def _c(self):
        if self.M_inv is not None:
            raise ValueError("M_inv is not supported on the HIP backend")
        o = _LobpcgOpts()
        for f in fields(self):
            if f.name in ("M_inv", "initial_block"):
                continue
            v = getattr(self, f.name)
            setattr(o, f.name, int(v) if isinstance(v, bool) else v)
        return o


@dataclass
class LOBPCGResult:
    """LOBPCGSolver.hpp:117-139."""
    right_null: np.ndarray
    theta: np.ndarray
    residuals_Av: np.ndarray
    iters: int
    locked: int
    frob_norm_est: float
    accept_threshold: float
    best_unaccepted_residual: float
    block_size: int
    converged_to_target: bool
    spmv: int = 0
    spmtv: int = 0
    times: dict = field(default_factory=dict)
    cholqr_blocks: int = 0     # FP32 blocks orthonormalised by Cholesky QR
    cholqr_fallbacks: int = 0  # ... of which nearly dependent ones went on to MGS


def find_nullspace_lobpcg_ex(A, vf, opt, precision="fp64"):
    """find_nullspace_lobpcg_ex (LOBPCGSolver.hpp:142-144)."""
    if not isinstance(A, HipMatrix):
        raise ValueError("find_nullspace_lobpcg: backend mismatch (HIP matrix required)")
    o = opt._c()
    kcap = max(int(opt.nev), int(opt.nullity_known), 1)
    if opt.rank_known >= 0:
        kcap = max(kcap, A.n - int(opt.rank_known))
    kcap = max(1, min(kcap, A.n))
    R = np.zeros((kcap, A.n))
    th = np.zeros(kcap)
    rs = np.zeros(kcap)
    ib = None
    if opt.initial_block:
        ib = np.ascontiguousarray(np.stack(opt.initial_block, axis=0), dtype=np.float64)
        if ib.shape[1] != A.n:
            raise ValueError("initial_block size mismatch")
    res = _LobpcgRes()
    _check(lib().kh_lobpcg(A.h, vf.h, C.byref(o), PREC[precision],
                           None if ib is None else _ptr(ib), 0 if ib is None else ib.shape[0],
                           _ptr(R), kcap, _ptr(th), _ptr(rs), C.byref(res)))
    n = res.n_right
    return LOBPCGResult(
        right_null=R[:n].T.copy(), theta=th[:n].copy(), residuals_Av=rs[:n].copy(),
        iters=res.iters, locked=res.locked, frob_norm_est=res.frob_norm_est,
        accept_threshold=res.accept_threshold,
        best_unaccepted_residual=res.best_unaccepted_residual, block_size=res.block_size,
        converged_to_target=bool(res.converged_to_target), spmv=res.spmv, spmtv=res.spmtv,
        times={"gram": res.t_gram, "sparse": res.t_sparse, "host": res.t_host},
        cholqr_blocks=res.cholqr_blocks, cholqr_fallbacks=res.cholqr_fallbacks)


def find_nullspace_lobpcg(A, vf, opt, precision="fp64"):
    """Backwards-compatible wrapper (LOBPCGSolver.hpp:147-149)."""
    return find_nullspace_lobpcg_ex(A, vf, opt, precision).right_null


def _take_host_csr(h):
    """Copy a kh_host_csr into a CsrMatrix and free it; returns (A, declared
    entry count, entries read)."""
    from .workloads import CsrMatrix
    try:
        m, n, nnz, dn, er = I64(), I64(), I64(), I64(), I64()
        _check(lib().kh_host_csr_info(h, C.byref(m), C.byref(n), C.byref(nnz), C.byref(dn),
                                      C.byref(er)))
        rp, ci, va = P(), P(), P()
        _check(lib().kh_host_csr_data(h, C.byref(rp), C.byref(ci), C.byref(va)))

        def arr(ptr, ctype, count, dtype):
            if count == 0:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(count,)).copy()
        A = CsrMatrix(m.value, n.value, arr(rp, C.c_int32, m.value + 1, np.int32),
                      arr(ci, C.c_int32, nnz.value, np.int32),
                      arr(va, C.c_double, nnz.value, np.float64))
        return A, dn.value, er.value
    finally:
        lib().kh_host_csr_free(h)


def load_ija(path, rows=-1, cols=-1, counts=False):
    """load_ija_triplet_matrix (src/IO/MatrixLoader.cpp:21-56) as canonical
    CSR: header "rows cols nnz", 0-based "i j value" triplets, repeats summed
    in input order, explicit zeros kept; rows/cols >= 0 override the header.
    With counts=True returns (A, header entry count, triplets read)."""
    h = P()
    _check(lib().kh_load_ija(os.fsencode(path), int(rows), int(cols), C.byref(h)))
    out = _take_host_csr(h)
    return out if counts else out[0]


def load_matrix_market(path):
    """load_matrix_market (src/IO/MatrixLoader.cpp:12-18) as canonical CSR."""
    h = P()
    _check(lib().kh_load_matrix_market(os.fsencode(path), C.byref(h)))
    return _take_host_csr(h)[0]


def csr_from_triplets(m, n, I, J, V):
    """Canonical CSR of 0-based triplets with setFromTriplets semantics."""
    I = np.ascontiguousarray(I, dtype=np.int32)
    J = np.ascontiguousarray(J, dtype=np.int32)
    V = np.ascontiguousarray(V, dtype=np.float64)
    if not (I.shape == J.shape == V.shape and I.ndim == 1):
        raise ValueError("csr_from_triplets: I, J, V must be 1-D arrays of one length")
    h = P()
    _check(lib().kh_csr_from_triplets(int(m), int(n), I.size, _ptr(I), _ptr(J), _ptr(V),
                                      C.byref(h)))
    return _take_host_csr(h)[0]


class NullspaceServer:
    """Continuous-batching null-space server (kh_server_*): requests are
    admitted in arrival order when one of `slots` solver slots (one device
    stream each) is free and their HBM footprint fits the budget."""

    def __init__(self, slots=4, hbm_budget_bytes=0.0, device=0, pack_max=1, pack_len=1 << 19):
        h = P()
        _check(lib().kh_server_create(device, slots, float(hbm_budget_bytes), C.byref(h)))
        self.h = h
        self._pending = {}
        self.set_packing(pack_max, pack_len)

    def set_packing(self, pack_max, pack_len=1 << 19):
        """Small requests (m + n <= pack_len) run in packs of up to pack_max
        whose Golub-Kahan cycles share segmented launches (kh_server_set_packing);
        pack_max=1 (the default) solves every request on its own."""
        _check(lib().kh_server_set_packing(self.h, int(pack_max), int(pack_len)))

    def close(self):
        if getattr(self, "h", None):
            lib().kh_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, A, opt, precision="fp64", seed=42):
        """A: object with m, n, rowptr, col, val (host CSR) or a
        HipCsrMatrixAdapter (resident). Returns a request id."""
        o = opt._c()
        rid = I64()
        if isinstance(A, HipMatrix):
            _check(lib().kh_server_submit_resident(self.h, A.h, C.byref(o), PREC[precision],
                                                   seed, C.byref(rid)))
            n = A.n
        else:
            rp = np.ascontiguousarray(A.rowptr, dtype=np.int32)
            ci = np.ascontiguousarray(A.col, dtype=np.int32)
            va = np.ascontiguousarray(A.val, dtype=np.float64)
            _check(lib().kh_server_submit_cyclic(self.h, A.m, A.n, _ptr(rp), _ptr(ci), _ptr(va),
                                                 C.byref(o), PREC[precision], seed,
                                                 C.byref(rid)))
            n = A.n
        self._pending[rid.value] = (n, max(int(opt.k), int(opt.nullity_known), 1))
        return rid.value

    def wait(self, rid):
        """Returns (basis n x k, CyclicLanczosResult-like dict, latency_s, queued_s)."""
        n, kcap = self._pending.pop(rid)
        R = np.zeros((kcap, n))
        res = _CyclicRes()
        lat, q = D(), D()
        _check(lib().kh_server_wait(self.h, rid, _ptr(R), kcap, C.byref(res), C.byref(lat),
                                    C.byref(q)))
        info = {f[0]: getattr(res, f[0]) for f in _CyclicRes._fields_}
        return R[:res.n_right].T.copy(), info, lat.value, q.value

    def stats(self):
        done, vec = I64(), I64()
        p50, p99, wall = D(), D(), D()
        _check(lib().kh_server_stats(self.h, C.byref(done), C.byref(vec), C.byref(p50),
                                     C.byref(p99), C.byref(wall)))
        return {"done": done.value, "vectors": vec.value, "p50_s": p50.value,
                "p99_s": p99.value, "wall_s": wall.value}

    def reset_clock(self):
        _check(lib().kh_server_reset_clock(self.h))
