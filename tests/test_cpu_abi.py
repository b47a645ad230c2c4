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
