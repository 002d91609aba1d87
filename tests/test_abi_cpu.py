"""CPU checks of the C ABI: the library loads without a GPU, exports every symbol that
include/pned_gpu.h declares, and reports the missing device as an error."""
import ctypes
import os
import re

import numpy as np
import pytest

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


@pytest.mark.parametrize("n,seed", [(80, 1), (400, 2), (7, 3), (3000, 5)])
def test_host_tridiagonal_solvers(n, seed):
    """The library's bisection and inverse iteration (its own arithmetic behind the
    reference's bisec / vec12 interface) against LAPACK (scipy.linalg.eigh_tridiagonal):
    eigenvalues within 64 ulps of the spectral radius (a Sturm count resolves them to
    O(n eps ||T||)), also against the oracle's restatement of the reference's TITPACK steps; eigenvectors of unit norm, residual at
    round-off and equal to LAPACK's up to sign."""
    import oracle
    import scipy.linalg as sla
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(n)
    b = rng.random(n) + 0.05
    ne = min(8, n)
    E = gpu.bisec(a, b, ne)
    w, vecs = sla.eigh_tridiagonal(a, b[: n - 1])
    scale = np.max(np.abs(w))
    ulp = np.finfo(float).eps * scale
    assert np.all(np.diff(E) >= 0)
    assert np.max(np.abs(E - w[:ne])) <= 64 * ulp
    assert np.max(np.abs(E - oracle.bisec(a, b, ne))) <= 64 * ulp
    nv = min(3, n)
    V = gpu.vec12(a, b, E, nv)
    T_v = lambda x: a * x + np.concatenate([b[: n - 1] * x[1:], [0.0]]) + np.concatenate([[0.0], b[: n - 1] * x[:-1]])
    for k in range(nv):
        assert abs(np.linalg.norm(V[k]) - 1.0) < 1e-14
        assert np.linalg.norm(T_v(V[k]) - E[k] * V[k]) < 1e-12 * scale
        assert abs(abs(V[k] @ vecs[:, k]) - 1.0) < 1e-12


def test_host_tridiagonal_degenerate_pair():
    """Two decoupled identical blocks: every eigenvalue twice; the two inverse-iteration
    vectors of a pair come out orthogonal eigenvectors."""
    rng = np.random.default_rng(4)
    h = rng.standard_normal(20)
    c = rng.random(19) + 0.1
    a = np.concatenate([h, h])
    b = np.concatenate([c, [0.0], c, [0.0]])
    E = gpu.bisec(a, b, 4)
    assert abs(E[0] - E[1]) < 1e-14 and abs(E[2] - E[3]) < 1e-14
    V = gpu.vec12(a, b, E, 4)
    T = np.diag(a) + np.diag(b[:-1], 1) + np.diag(b[:-1], -1)
    assert abs(V[0] @ V[1]) < 1e-12 and abs(V[2] @ V[3]) < 1e-12
    for k in range(4):
        assert np.linalg.norm(T @ V[k] - E[k] * V[k]) < 1e-12


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
