"""kh_cgs2_dev_noalloc (the entry point that allocates nothing) against
kh_cgs2_dev on the same inputs: bitwise the same w, h and ||w||, on two
calls in a row (the second smaller, so the shared scratch holds the first
call's contents), in both precisions; oversized calls are refused before any
work."""
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


def inputs(len_, s, dt, seed):
    rng = np.random.default_rng(seed)
    V, _ = np.linalg.qr(rng.standard_normal((len_, s)))
    w = rng.standard_normal(len_) + V @ rng.standard_normal(s)
    return V.astype(dt), w.astype(dt)


def run_pair(ctx, V, w, prec):
    s, n = V.shape[1], V.shape[0]
    # the existing entry point: h and ||w|| come back to the host
    dV, dw = Dev(ctx, V), Dev(ctx, w)
    h_ref = np.zeros(s)
    nrm = C.c_double()
    K._check(K.lib().kh_cgs2_dev(ctx.h, prec, dV.p, s, dw.p, n, h_ref.ctypes.data_as(C.c_void_p),
                                 C.byref(nrm)))
    nrm_ref = np.array([nrm.value])
    w_ref = dw.get()
    # the no-alloc entry point: h and ||w|| stay on the device
    dw2 = Dev(ctx, w)
    dh, dn = Dev(ctx, np.full(max(s, 1), np.nan)), Dev(ctx, np.full(1, np.nan))
    K._check(K.lib().kh_cgs2_dev_noalloc(ctx.h, prec, dV.p, s, dw2.p, n, dh.p, dn.p))
    ctx.sync()
    return (w_ref, h_ref, nrm_ref), (dw2.get(), dh.get()[:s], dn.get())


@pytest.mark.parametrize("prec", [1, 0])
def test_noalloc_equals_cgs2_dev_on_two_calls(prec):
    ctx = K.default_context()
    dt = np.float32 if prec == 1 else np.float64
    # first call: 2^20 x 40 (the LDS-tiled middle sweep; 16-byte loads in
    # FP32); second, smaller: 50,001 x 13 (ragged length)
    for len_, s, seed in ((1 << 20, 40, 1), (50_001, 13, 2)):
        print(f"[case] noalloc cgs2 prec={prec} len={len_} s={s}", flush=True)
        V, w = inputs(len_, s, dt, seed)
        (w0, h0, n0), (w1, h1, n1) = run_pair(ctx, V, w, prec)
        assert np.array_equal(w0.view(np.uint8), w1.view(np.uint8))
        assert np.array_equal(h0.view(np.uint64), h1.view(np.uint64))
        assert np.array_equal(n0.view(np.uint64), n1.view(np.uint64))
        # and the projection itself (the parity bar of the CGS2 kernels)
        w64 = w.astype(np.float64)
        V64 = V.astype(np.float64)
        eps = np.finfo(dt).eps
        assert np.abs(V64.T @ w1.astype(np.float64)).max() <= 64 * eps * np.linalg.norm(w64)
        assert abs(n1[0] - np.linalg.norm(w1.astype(np.float64))) <= 64 * eps * n1[0]


def test_noalloc_refuses_oversized_calls_before_any_work():
    ctx = K.default_context()
    p = C.c_void_p(8)  # never dereferenced: the size check comes first
    rc = K.lib().kh_cgs2_dev_noalloc(ctx.h, 0, p, 1024, p, 1 << 24, p, p)
    assert rc == 1 and b"32 MiB" in K.lib().kh_last_error()
    rc = K.lib().kh_cgs2_dev_noalloc(ctx.h, 0, p, 2048, p, 100, p, p)
    assert rc == 1
