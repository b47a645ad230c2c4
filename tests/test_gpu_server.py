"""Continuous-batching server: each result is the same alone and in the mix
(bitwise), packed small requests the same alone and in any pack (bitwise) and
within 1e-10 of the direct solve, admission by HBM footprint, error
isolation."""
import numpy as np
import pytest

import krans_amd as K
from krans_amd import workloads as W
from conftest import principal_sine

pytestmark = pytest.mark.gpu


def mix():
    """Uneven requests from the three generator families."""
    reqs = []
    reqs.append((W.regular_incidence(8, 2048, seed=1), 8, "mixed"))
    reqs.append((W.shifted_difference(20000, 5, 0.7), 5, "fp64"))
    reqs.append((W.regular_incidence(3, 4096, seed=2), 3, "fp64"))
    reqs.append((W.rg_block(5000, p=4, offdiag=8, gnnz=16, seed=3)[0], 4, "mixed"))
    reqs.append((W.grid_incidence(2, 48, seed=4, weighted=False), 2, "mixed"))
    reqs.append((W.shifted_difference(50000, 2, 0.5), 2, "mixed"))
    reqs.append((W.regular_incidence(16, 1024, seed=5), 16, "fp64"))
    reqs.append((W.regular_incidence(4, 8192, seed=6), 4, "mixed"))
    return reqs


def opts(k, graphs=0):
    return K.CyclicLanczosOptions(k=k, tol_rel=1e-12, tol_abs=0.0, prescreen_factor=1e-3,
                                  max_restarts=200, cg_maxit=400, graphs=graphs)


@pytest.mark.parametrize("graphs", [0, 1])
def test_same_alone_and_in_the_mix(graphs):
    """Solo solves with direct launches against the server's four slots;
    with graphs=1 the slots capture and replay their Golub-Kahan cycles
    concurrently (thread-local capture while other slots allocate)."""
    reqs = mix()
    print(f"[case] server alone (mix graphs={graphs})", flush=True)
    alone = []
    for i, (A, k, prec) in enumerate(reqs):
        r = K.find_nullspace_cyclic_lanczos_ex(K.HipCsrMatrixAdapter.from_csr(A),
                                               K.HipVectorFactory(A.n, 100 + i), opts(k), prec)
        alone.append(r.right_null)
        assert r.right_null.shape[1] == k, (i, r.right_null.shape)
        # every result meets the acceptance threshold on its own
        res = np.linalg.norm(A.to_scipy() @ r.right_null, axis=0).max()
        assert res <= 1e-12 * max(1.0, r.frob_norm_est), (i, res)
        assert np.abs(r.right_null.T @ r.right_null - np.eye(k)).max() <= 1e-10
    print("[case] server mix", flush=True)
    srv = K.NullspaceServer(slots=4, pack_max=1)
    ids = [srv.submit(A, opts(k, graphs), prec, seed=100 + i)
           for i, (A, k, prec) in enumerate(reqs)]
    for i, rid in enumerate(ids):
        U, info, lat, q = srv.wait(rid)
        assert (info["graph_replays"] > 0) == (graphs == 1), (i, info["graph_replays"])
        assert U.shape == alone[i].shape
        assert np.array_equal(U.view(np.uint64), alone[i].view(np.uint64)), i
        assert principal_sine(U, alone[i]) <= 1e-10
        assert lat >= q >= 0
    st = srv.stats()
    assert st["done"] == len(reqs) and st["p99_s"] >= st["p50_s"] > 0
    srv.close()


def small_requests():
    """Small requests of differing sizes and nullities (a pack's members
    carry locked bases of different sizes)."""
    return [(W.regular_incidence(4, 1024, seed=11), 4),
            (W.shifted_difference(6000, 3, 0.7), 3),
            (W.grid_incidence(2, 32, seed=12, weighted=False), 2),
            (W.regular_incidence(8, 512, seed=13), 8),
            (W.rg_block(3000, p=3, offdiag=8, gnnz=16, seed=14)[0], 3),
            (W.regular_incidence(2, 3000, seed=15), 2)]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_packed_same_alone_and_in_a_pack(prec):
    """Each request alone in a pack of one and among the others in a pack
    gives the same bits; both are within 1e-10 (principal sine) of the direct
    solve and meet the residual threshold."""
    reqs = small_requests()
    print(f"[case] packed alone ({prec})", flush=True)
    srv = K.NullspaceServer(slots=1, pack_max=16)
    alone = []
    for i, (A, k) in enumerate(reqs):
        U, info, _, _ = srv.wait(srv.submit(A, opts(k), prec, seed=200 + i))
        assert info["packed"] == 1, (i, info["packed"])
        alone.append(U)
    print(f"[case] packed together ({prec})", flush=True)
    ids = [srv.submit(A, opts(k), prec, seed=200 + i) for i, (A, k) in enumerate(reqs)]
    sizes = []
    for i, rid in enumerate(ids):
        U, info, lat, q = srv.wait(rid)
        sizes.append(info["packed"])
        assert U.shape == alone[i].shape, i
        assert np.array_equal(U.view(np.uint64), alone[i].view(np.uint64)), (i, info["packed"])
        assert lat >= q >= 0
    # the first request starts on its own; the others queue behind it and pack
    assert max(sizes) > 1, sizes
    srv.close()
    print(f"[case] packed vs direct ({prec})", flush=True)
    for i, (A, k) in enumerate(reqs):
        r = K.find_nullspace_cyclic_lanczos_ex(K.HipCsrMatrixAdapter.from_csr(A),
                                               K.HipVectorFactory(A.n, 200 + i), opts(k), prec)
        assert alone[i].shape[1] == k == r.right_null.shape[1], i
        assert principal_sine(alone[i], r.right_null) <= 1e-10, i
        res = np.linalg.norm(A.to_scipy() @ alone[i], axis=0).max()
        assert res <= 1e-12 * max(1.0, r.frob_norm_est), (i, res)
        assert np.abs(alone[i].T @ alone[i] - np.eye(k)).max() <= 1e-10


def test_admission_under_a_small_budget_and_resident_matrices():
    A = W.regular_incidence(4, 4096, seed=7)
    H = K.HipCsrMatrixAdapter.from_csr(A)
    print("[case] server budget", flush=True)
    # room for about one request at a time: the rest wait in the queue
    srv = K.NullspaceServer(slots=4, hbm_budget_bytes=1.6e8)
    ids = [srv.submit(H, opts(4), "mixed", seed=s) for s in (1, 2, 3)]
    outs = [srv.wait(r) for r in ids]
    Kx = W.component_indicators(4, 4096)
    for U, info, lat, q in outs:
        assert U.shape[1] == 4 and principal_sine(U, Kx) <= 1e-10
    assert max(o[3] for o in outs) > 0.0  # somebody queued
    with pytest.raises(ValueError):
        srv.submit(W.regular_incidence(64, 16384, seed=1), opts(64), "fp64")
    srv.close()


@pytest.mark.parametrize("pack_max", [1, 16])
def test_bad_request_fails_alone(pack_max):
    """A malformed matrix fails its own request, solo or inside a pack."""
    srv = K.NullspaceServer(slots=2, pack_max=pack_max)
    bad = W.CsrMatrix(2, 2, np.array([0, 2, 1], np.int32), np.array([0, 1], np.int32),
                      np.ones(2))
    good = W.regular_incidence(2, 512, seed=8)
    print("[case] server bad request", flush=True)
    rb = srv.submit(bad, opts(1), "fp64", seed=1)
    rg = [srv.submit(good, opts(2), "fp64", seed=s) for s in (1, 2, 3)]
    with pytest.raises(ValueError):
        srv.wait(rb)
    Kx = W.component_indicators(2, 512)
    for r in rg:
        U, _, _, _ = srv.wait(r)
        assert U.shape[1] == 2 and principal_sine(U, Kx) <= 1e-10
    srv.close()
