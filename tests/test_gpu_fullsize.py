"""Full-size checks: the S=1/2 N=32 workload (2^32 states, 64 GiB per complex128 vector) and
the spin-1 N=18 workload (3^18 states) through size-independent properties.

The oracle cannot run these sizes in seconds, so the H.v is checked on product states
|v> = (x)_k |phi_k>, for which the mathematics fixes

    <v|H|v>   = sum_T c_T prod_k <phi_k|A_T,k|phi_k>
    <Hv|Hv>   = sum_{T,T'} c_T c_T' prod_k <phi_k|A_T,k A_T',k|phi_k>

with H written as textbook spin operators (tests/dense.py's site operators and couplings,
not the bond matrices the product consumes): a dropped term, a wrong sign, coefficient,
site order or selector anywhere in the 2^32-state sweep moves one of the two.  The N=32
case also compares the term-major and state-major pass kernels on the same input.
Tolerances: 1e-11 relative (FP64 sums over 2^32 states).
"""
import sys

import numpy as np
import pytest

from pned_amd import gpu, lattice, model
from tests import dense

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if gpu.device_count() < 1:
        pytest.skip("no GPU")


def _vectors(ctx, count):
    try:
        return [ctx.vector() for _ in range(count)]
    except gpu.PnedError as e:
        pytest.skip(f"device memory: {e}")


def test_n32_product_state_closed_forms_and_kernels():
    cl = lattice.torus(4, 4)
    phi, h = 126.0, 0.1
    _say("S=1/2 honeycomb 4x4 N=32, phi=126, h=0.1, product state, both pass kernels")
    op = model.build_operator(cl, 20, phi, field=h)
    ctx = gpu.Context(op)
    v1, v0, v2 = _vectors(ctx, 3)
    phis = dense.random_product_sites(cl.n, 2, 32)
    e1, e2 = dense.product_closed_forms(dense.kh_terms(cl, 2, phi, h), phis)
    dense.upload_product(v1, phis, 2, 24)
    nrm = ctx.calc_vec_inprdct(v1, v1)
    assert abs(nrm - 1.0) <= 1e-12
    pr = ctx.hv_fresh(v1, v0)
    hh = ctx.calc_vec_inprdct(v0, v0)
    vh = ctx.calc_vec_inprdct(v1, v0)
    _say(f"<v|Hv> {pr:.15e} closed form {e1:.15e}; <Hv|Hv> {hh.real:.15e} closed form {e2.real:.15e}")
    scale = np.sqrt(abs(e2))
    assert abs(pr - e1) <= 1e-11 * scale
    assert abs(vh - e1) <= 1e-11 * scale
    assert abs(hh - e2) <= 1e-11 * abs(e2)
    _say("state-major kernel on the same input")
    ctx.s12_kernel(0)
    pr2 = ctx.hv_fresh(v1, v2)
    ctx.axpby(-1.0, v0, 1.0, v2)
    d2 = ctx.calc_vec_inprdct(v2, v2).real
    assert abs(pr2 - pr) <= 1e-12 * scale
    assert np.sqrt(d2) <= 1e-13 * np.sqrt(abs(e2))
    for v in (v1, v0, v2):
        v.free()
    ctx.close()


@pytest.mark.parametrize("phi,h", [(108.0, 0.0), (30.0, 0.2)])
def test_spin1_n18_product_state_closed_forms(phi, h):
    cl = lattice.reference_cluster(18)
    _say(f"spin-1 N=18, phi={phi}, h={h}, product state")
    op = model.build_operator(cl, 20, phi, q=3, field=h)
    ctx = gpu.Context(op)
    v1, v0 = _vectors(ctx, 2)
    phis = dense.random_product_sites(cl.n, 3, 18)
    e1, e2 = dense.product_closed_forms(dense.kh_terms(cl, 3, phi, h), phis)
    dense.upload_product(v1, phis, 3, 12)
    pr = ctx.hv_fresh(v1, v0)
    hh = ctx.calc_vec_inprdct(v0, v0)
    _say(f"<v|Hv> {pr:.15e} closed form {e1:.15e}; <Hv|Hv> {hh.real:.15e} closed form {e2.real:.15e}")
    scale = np.sqrt(abs(e2))
    assert abs(pr - e1) <= 1e-11 * scale
    assert abs(hh - e2) <= 1e-11 * abs(e2)
    for v in (v1, v0):
        v.free()
    ctx.close()


@pytest.mark.parametrize("phi,h", [(108.0, 0.0), (30.0, 0.2)])
def test_spin1_n18_rows_vs_oracle(phi, h):
    """Spin-1 N=18 (3^18 states) on a Gaussian start vector: output rows of the tiled kernel
    against the oracle's reference-structured H.v restricted to the same rows (the oracle
    reads the whole input vector, so every partner is checked): three slices of 2^21 rows
    at the start, the middle and the end of the space.  1e-12 relative."""
    import oracle

    cl = lattice.reference_cluster(18)
    _say(f"spin-1 N=18, phi={phi}, h={h}, rows against the oracle")
    op = model.build_operator(cl, 20, phi, q=3, field=h)
    ctx = gpu.Context(op)
    v1, v0 = _vectors(ctx, 2)
    ctx.set_initial_vec(v1, 20261203)
    ctx.hv_fresh(v1, v0, 0.05)
    X = v1.download()
    Y = v0.download()
    oop = oracle.OracleOp(op)
    span = 1 << 21
    want = np.zeros(op.dim, dtype=np.complex128)   # the oracle writes only rows [lo, lo + span)
    for lo in (0, op.dim // 2, op.dim - span):
        want[lo:lo + span] = 0.0
        oracle.hv_rows(oop, X, want, lo, lo + span, eperbd=0.05)
        got = Y[lo:lo + span]
        err = np.max(np.abs(got - want[lo:lo + span])) / np.max(np.abs(want[lo:lo + span]))
        _say(f"rows [{lo}, {lo + span}): max rel diff {err:.2e}")
        assert err <= 1e-12
    for v in (v1, v0):
        v.free()
    ctx.close()
