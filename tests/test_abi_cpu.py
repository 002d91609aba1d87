"""CPU checks of the C ABI: the library loads without a GPU, exports every symbol that
include/pned_gpu.h declares, and reports the missing device as an error."""
import ctypes
import os
import re

import numpy as np

from pned_amd import gpu, lattice, model

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    out = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            out |= set(re.findall(r"\b(pned_[a-z0-9_]+)\s*\(", text))
    return out


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(gpu.LIB_PATH)
    missing = [s for s in sorted(declared_symbols()) if not hasattr(L, s)]
    assert not missing, missing
    assert len(declared_symbols()) > 30


def test_no_gpu_is_an_error_not_a_fallback():
    if gpu.device_count() > 0:
        return
    op = model.build_operator(lattice.reference_cluster(8), 20, 54.0)
    try:
        gpu.Context(op)
    except gpu.PnedError as e:
        assert e.code == gpu.E_HIP
    else:
        raise AssertionError("context created without a GPU")


def test_host_tridiagonal_solvers_match_oracle():
    import oracle
    rng = np.random.default_rng(1)
    a = rng.standard_normal(80)
    b = rng.random(80) + 0.05
    E = gpu.bisec(a, b, 8)
    assert np.array_equal(E, oracle.bisec(a, b, 8))
    V = gpu.vec12(a, b, E, 3)
    assert np.allclose(V, oracle.vec12(a, b, E, 3), atol=1e-14)


def test_ctpq_product_matches_oracle():
    """pned_ctpq_eval (host code of the library) against the oracle's restatement on the
    oracle's own mTPQ output for 3 samples of the 8-site cluster."""
    import oracle
    from pned_amd import lattice, model
    cl = lattice.reference_cluster(8)
    op = model.build_operator(cl, 20, 90.0, tps=True)
    steps, seeds = 120, 3
    A, B, Cf = [], [], []
    for sd in range(seeds):
        a, b, c = oracle.tps(oracle.OracleOp(op), cl.site_vec, np.zeros((0, 2)), op.shift_L, steps,
                             oracle.gaussian(2 ** cl.n, 11 + sd))
        A.append(a)
        B.append(b)
        Cf.append(c.T)
    A, B, Cf = np.array(A), np.array(B), np.array(Cf)
    got = gpu.ctpq(A, B, Cf, 0.05, 5.0, 40)
    want = oracle.ctpq(A, B, Cf, 0.05, 5.0, 40)
    assert np.allclose(got, want, rtol=1e-13, atol=1e-12)
    assert np.all(np.diff(got[:, 0]) > 0)
