"""Host-side failure handling of the serving loop, run where no device is
present: every solver slot fails to start, and a client already blocked in
wait() must be woken with an error instead of waiting forever; malformed
submissions are rejected at submit."""
import ctypes as C
import threading

import numpy as np
import pytest

import krans_amd as K
from krans_amd import workloads as W


def no_device():
    n = C.c_int(0)
    rc = K.lib().kh_device_count(C.byref(n))
    return rc != 0 or n.value == 0


pytestmark = pytest.mark.skipif(not no_device(), reason="needs a host without a GPU")


@pytest.mark.timeout(60)
def test_waiters_wake_when_no_slot_can_start():
    A = W.regular_incidence(1, 64, seed=1)
    for _ in range(20):  # clients that wait before and after the slots fail
        srv = K.NullspaceServer(slots=3, hbm_budget_bytes=1e9)
        rid = srv.submit(A, K.CyclicLanczosOptions(k=1))
        errs = []

        def client():
            try:
                srv.wait(rid)
            except K.KransError as e:
                errs.append(str(e))
        t = threading.Thread(target=client)
        t.start()
        t.join(30)
        assert not t.is_alive(), "wait() did not return after every slot failed"
        assert errs and "no solver slot could start" in errs[0]
        srv.close()


def test_submit_rejects_malformed_requests():
    srv = K.NullspaceServer(slots=1, hbm_budget_bytes=1e9)
    A = W.regular_incidence(1, 64, seed=1)
    with pytest.raises(ValueError, match="return_left_space"):
        srv.submit(A, K.CyclicLanczosOptions(k=1, return_left_space=True))
    bad = W.CsrMatrix(A.m, A.n, A.rowptr.copy(), A.col, A.val)
    bad.rowptr[-1] = -5
    with pytest.raises(ValueError):
        srv.submit(bad, K.CyclicLanczosOptions(k=1))
    srv.close()


def test_wait_rejects_negative_capacity():
    srv = K.NullspaceServer(slots=1, hbm_budget_bytes=1e9)
    res = K._CyclicRes()
    rc = K.lib().kh_server_wait(srv.h, 1, None, -1, C.byref(res), None, None)
    assert rc != 0 and b"negative basis capacity" in K.lib().kh_last_error()
    srv.close()
