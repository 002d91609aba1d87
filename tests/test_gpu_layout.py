"""Memory layouts (site -> memory bit swaps) on the GPU: every entry point that maps basis
states to memory must give the oracle's results in the reference's basis order, whatever
layout the context holds.  Layouts are forced (pned_ctx_set_layout) on small clusters; the
ones the planner chooses by itself are exercised by the H.v parity cases at N=18 and 24."""
import sys

import numpy as np
import pytest

import oracle
from pned_amd import gpu, lattice, model

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if gpu.device_count() < 1:
        pytest.skip("no GPU")


def _rand(d, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(d) + 1j * r.standard_normal(d)


def _table(n, pairs):
    sb = np.arange(n, dtype=np.int32)
    for a, b in pairs:
        sb[a], sb[b] = b, a
    return sb


LAYOUTS = [[(0, 7), (2, 11)], [(1, 5), (3, 10), (0, 9), (2, 4)]]


def _ctx(op, pairs):
    ctx = gpu.Context(op)
    ctx.set_layout(_table(op.n, pairs))
    assert np.array_equal(ctx.layout(), _table(op.n, pairs))
    return ctx


@pytest.mark.parametrize("pairs", LAYOUTS)
def test_layout_hv_start_vector_and_lanczos(pairs):
    _say("layout hv/gauss/lanczos", pairs)
    op = model.build_operator(lattice.reference_cluster(12), 20, 126.0, field=0.1)
    ctx = _ctx(op, pairs)
    d = ctx.dim
    x, y = _rand(d, 1), _rand(d, 2)
    v1, v0 = ctx.vector(), ctx.vector()
    v1.upload(x)
    assert np.array_equal(v1.download(), x)
    assert np.array_equal(v1.download(100, 37), x[100:137])
    v0.upload(y)
    pr = ctx.p_mltply_v3(v1, v0, 0.07)
    want = y.copy()
    pr_ref = oracle.hv(oracle.OracleOp(op), x, want, eperbd=0.07)
    assert np.linalg.norm(v0.download() - want) <= 1e-12 * np.linalg.norm(want)
    assert abs(pr - pr_ref) <= 1e-12 * abs(pr_ref)
    ctx.set_initial_vec(v1, 20261202)
    g = v1.download()
    assert np.linalg.norm(g - oracle.gaussian(d, 20261202)) < 1e-13
    itr, al, be, ene = ctx.lnc1(v1, v0, fixed_steps=60, nvec=2)
    _, a_ref, b_ref, e_ref = oracle.lanczos(oracle.OracleOp(op), g, fixed_steps=60, nvec=2)
    assert abs(ene[0] - e_ref[0]) < 1e-10 * abs(e_ref[0])
    assert np.allclose(al[:20], a_ref[:20], rtol=1e-12)
    # the vectors come back in the reference's loop state, in basis order
    w = ctx.vector()
    ctx.set_initial_vec(w, 20261202)
    ex, ratios = ctx.check1(w, v0)
    hx = np.zeros(d, dtype=complex)
    oracle.hv(oracle.OracleOp(op), g, hx)
    js = np.array(list(range(min(d // 3, 13), d + 1, max(1, d // 20))))
    assert np.allclose(ratios, hx[js - 1] / g[js - 1], rtol=1e-10)


@pytest.mark.parametrize("pairs", LAYOUTS)
def test_layout_excitations(pairs):
    _say("layout S^z(q), S^+-(q)", pairs)
    cl = lattice.reference_cluster(12)
    op = model.build_operator(cl, 20, 54.0, field=0.1)
    ctx = _ctx(op, pairs)
    x = _rand(ctx.dim, 5)
    xv, w = ctx.vector(), ctx.vector()
    xv.upload(x)
    q = cl.q_list[3]
    cn = ctx.intalv_zz(xv, w, q, cl.site_vec)
    vx = oracle.sz_q(cl.n, 2, cl.site_vec, q, x)
    cn_ref = np.vdot(vx, vx).real
    assert abs(cn - cn_ref / cl.n) <= 1e-13 * cn_ref
    assert np.linalg.norm(w.download() - vx / np.sqrt(cn_ref)) <= 1e-13
    # S^+(q) from this layout into a context with the identity layout, and back
    other = gpu.Context(op)
    wo = other.vector()
    cn2 = other.intalv_xy(xv, wo, q, cl.site_vec, 1)
    v_ref, cn_ref2 = oracle.intalv_pm(cl.n, 2, cl.site_vec, q, 1, x)
    assert abs(cn2 - cn_ref2) <= 1e-13 * cn_ref2
    assert np.linalg.norm(wo.download() - v_ref) <= 1e-13
    xo = other.vector()
    xo.upload(x)
    cn3 = ctx.intalv_xy(xo, w, q, cl.site_vec, -1)
    v_ref3, cn_ref3 = oracle.intalv_pm(cl.n, 2, cl.site_vec, q, -1, x)
    assert abs(cn3 - cn_ref3) <= 1e-13 * cn_ref3
    assert np.linalg.norm(w.download() - v_ref3) <= 1e-13
    # plain copies and dots between the two layouts are refused
    with pytest.raises(gpu.PnedError):
        wo.copy_from(w)


@pytest.mark.parametrize("pairs", LAYOUTS)
def test_layout_mtpq_observables(pairs):
    _say("layout mTPQ", pairs)
    cl = lattice.reference_cluster(12)
    op = model.build_operator(cl, 20, 90.0, tps=True)
    ctx = _ctx(op, pairs)
    ctx.tps_set_observables(cl.site_vec, cl.sq_momenta())
    nvec = 2
    v1, v0, vg, vt = (ctx.vector(nvec, gpu.C128) for _ in range(4))
    ctx.set_initial_vec(v1, 11)
    X = v1.download()
    steps = 8
    al, be, co = ctx.tps(v1, v0, vg, vt, op.shift_L, steps)
    ctx.set_initial_vec(v1, 11)
    af, bf, cf = ctx.tps_fused(v1, v0, vg, op.shift_L, steps)
    for c in range(nvec):
        xc = np.ascontiguousarray(X[:, c])
        assert np.linalg.norm(xc - oracle.gaussian(ctx.dim, 11, col=c)) < 1e-13
        a_ref, b_ref, c_ref = oracle.tps(oracle.OracleOp(op), cl.site_vec, cl.sq_momenta(), op.shift_L, steps, xc)
        assert np.allclose(al[:, c], a_ref, rtol=1e-11)
        assert np.allclose(be[:, c], b_ref, rtol=1e-11)
        assert np.allclose(co[:, c, :], c_ref, rtol=1e-9, atol=1e-12)
        assert np.allclose(cf[:, c, :], c_ref, rtol=1e-9, atol=1e-12)


def test_layout_vector_files_are_in_basis_order(tmp_path):
    _say("layout vector files")
    gpu.io_subrecord_limit(1000)    # subrecords that end inside a state
    try:
        op = model.build_operator(lattice.reference_cluster(12), 20, 54.0)
        lay = _ctx(op, LAYOUTS[1])
        ident = gpu.Context(op)
        x = _rand(lay.dim, 9)
        a, b = lay.vector(), ident.vector()
        a.upload(x)
        b.upload(x)
        fa, fb = tmp_path / "vec_a", tmp_path / "vec_b"
        lay.dump_vectors(fa, 3, [a])
        ident.dump_vectors(fb, 3, [b])
        assert fa.read_bytes() == fb.read_bytes()
        c = lay.vector()
        assert lay.read_vectors(fb, [c]) == 3
        assert np.array_equal(c.download(), x)
    finally:
        gpu.io_subrecord_limit(0)


def test_layout_argument_rules():
    _say("layout rules")
    op = model.build_operator(lattice.reference_cluster(12), 20, 54.0)
    ctx = gpu.Context(op)
    with pytest.raises(gpu.PnedError):
        ctx.set_layout(_table(12, [(0, 7)])[::-1].copy())     # not an involution
    v = ctx.vector()
    with pytest.raises(gpu.PnedError):
        ctx.set_layout(_table(12, [(0, 7)]))                   # vectors exist
    del v


def test_chosen_layout_n24_lanczos():
    """The 3x4 torus gets a layout of its own choosing; start vector and Lanczos energies
    agree with the oracle in basis order."""
    _say("chosen layout 3x4")
    cl = lattice.get_cluster("3x4")
    op = model.build_operator(cl, 20, 54.0)
    ctx = gpu.Context(op)
    assert not np.array_equal(ctx.layout(), np.arange(cl.n))
    v1, v0 = ctx.vector(), ctx.vector()
    ctx.set_initial_vec(v1, 4)
    g = v1.download()
    assert np.linalg.norm(g - oracle.gaussian(ctx.dim, 4)) < 1e-13
    _, al, be, ene = ctx.lnc1(v1, v0, fixed_steps=30, nvec=2)
    _, a_ref, b_ref, e_ref = oracle.lanczos(oracle.OracleOp(op), g, fixed_steps=30, nvec=2)
    assert np.allclose(al[:20], a_ref[:20], rtol=1e-12)
    assert np.allclose(be[:20], b_ref[:20], rtol=1e-12)
