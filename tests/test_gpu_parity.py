"""GPU parity: the HIP path through the C ABI against the CPU oracle on the same inputs.

Tolerances (stated by the hot-path contract): H.v <= 1e-12 relative (normwise) for
complex128, energies <= 1e-10 relative; complex64 storage is checked against its own
rounding bound (~1e-6 relative).  Every case prints its name first so a failing GPU
step is identifiable from the log.
"""
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


def _rand(d, seed, nvec=1):
    r = np.random.default_rng(seed)
    z = r.standard_normal(d * nvec) + 1j * r.standard_normal(d * nvec)
    return z if nvec == 1 else z.reshape(d, nvec)


HV_CASES = [
    (8, 54.0, 0.0, 2), (12, 126.0, 0.0, 2), (12, 126.0, 0.1, 2), (16, 200.0, 0.15, 2),
    ("2x4", 252.0, 0.3, 2), (18, 33.0, 0.0, 2), (24, 54.0, 0.0, 2), ("3x4", 126.0, 0.1, 2),
    (8, 108.0, 0.0, 3), (8, 30.0, 0.2, 3), (12, 108.0, 0.0, 3),
]


KERNELS = [1, 0]   # S=1/2 pass kernel: term-major, state-major


@pytest.mark.parametrize("kern", KERNELS)
@pytest.mark.parametrize("spec,phi,h,q", HV_CASES)
def test_hv_parity(spec, phi, h, q, kern):
    _say("hv", spec, phi, h, q, "kernel", kern)
    cl = lattice.get_cluster(spec)
    op = model.build_operator(cl, 20, phi, q=q, field=h)
    ctx = gpu.Context(op)
    ctx.s12_kernel(kern)
    d = ctx.dim
    x = _rand(d, 1)
    y = _rand(d, 2)
    v1, v0 = ctx.vector(), ctx.vector()
    v1.upload(x)
    v0.upload(y)
    ep = 0.07
    pr = ctx.p_mltply_v3(v1, v0, ep)
    got = v0.download()
    want = y.copy()
    pr_ref = oracle.hv(oracle.OracleOp(op), x, want, eperbd=ep)
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
    # <x|H|x> of an unnormalised vector cancels heavily: bound by |x| |Hx|
    hx = want - y
    scale = np.linalg.norm(x) * np.linalg.norm(hx)
    assert abs(pr - pr_ref) <= 1e-13 * scale
    # fresh-output variant
    pr2 = ctx.hv_fresh(v1, v0, ep)
    want2 = np.zeros_like(x)
    oracle.hv(oracle.OracleOp(op), x, want2, eperbd=ep)
    got2 = v0.download()
    assert np.linalg.norm(got2 - want2) <= 1e-12 * np.linalg.norm(want2)
    assert abs(pr2 - pr_ref) <= 1e-13 * scale


@pytest.mark.parametrize("kern", KERNELS)
@pytest.mark.parametrize("tile_bits,threads", [(8, 0), (10, 0), (12, 256), (12, 512), (13, 512), (11, 256),
                                               (13, 1024), (12, 1024)])
def test_hv_tile_shapes(tile_bits, threads, kern):
    _say("hv tile", tile_bits, threads, "kernel", kern)
    op = model.build_operator(lattice.reference_cluster(16), 20, 77.0, field=0.2)
    ctx = gpu.Context(op)
    ctx.s12_kernel(kern)
    if threads == 1024 and kern == 0:
        pytest.skip("1024 threads per block: term-major kernel only")
    ctx.tune(tile_bits=tile_bits, low_bits=3, threads=threads)
    x = _rand(ctx.dim, 3)
    v1, v0 = ctx.vector(), ctx.vector()
    v1.upload(x)
    ctx.hv_fresh(v1, v0)
    want = np.zeros_like(x)
    oracle.hv(oracle.OracleOp(op), x, want)
    assert np.linalg.norm(v0.download() - want) <= 1e-12 * np.linalg.norm(want)


@pytest.mark.parametrize("kern", KERNELS)
@pytest.mark.parametrize("nvec,dtype", [(4, gpu.C128), (8, gpu.C64), (1, gpu.C64), (64, gpu.C64)])
def test_hv_batched_and_c64(nvec, dtype, kern):
    _say("hv batched", nvec, dtype, "kernel", kern)
    op = model.build_operator(lattice.reference_cluster(16), 20, 126.0, field=0.1)
    ctx = gpu.Context(op)
    ctx.s12_kernel(kern)
    X = _rand(ctx.dim, 5, nvec)
    if nvec == 1:
        X = X.reshape(-1, 1)
    v1, v0 = ctx.vector(nvec, dtype), ctx.vector(nvec, dtype)
    v1.upload(X)
    pr = np.atleast_1d(ctx.hv_fresh(v1, v0))
    got = v0.download().reshape(ctx.dim, nvec)
    tol = 1e-12 if dtype == gpu.C128 else 2e-6
    xr = X.astype(np.complex64).astype(np.complex128) if dtype == gpu.C64 else X
    for c in range(nvec):
        want = np.zeros(ctx.dim, complex)
        p_ref = oracle.hv(oracle.OracleOp(op), np.ascontiguousarray(xr[:, c]), want)
        assert np.linalg.norm(got[:, c] - want) <= tol * np.linalg.norm(want), c
        scale = np.linalg.norm(xr[:, c]) * np.linalg.norm(want)
        assert abs(pr[c] - p_ref) <= 0.1 * tol * scale


def test_blas1_and_gaussian():
    _say("blas1")
    op = model.build_operator(lattice.reference_cluster(12), 20, 54.0)
    ctx = gpu.Context(op)
    d = ctx.dim
    v = ctx.vector()
    ctx.set_initial_vec(v, 20261202)
    ref = oracle.gaussian(d, 20261202)
    assert np.linalg.norm(v.download() - ref) <= 1e-14
    x, y = _rand(d, 7), _rand(d, 8)
    a, b = ctx.vector(), ctx.vector()
    a.upload(x)
    b.upload(y)
    assert abs(ctx.calc_vec_inprdct(a, b) - np.vdot(x, y)) <= 1e-13 * abs(np.vdot(x, y))
    ctx.axpby(0.3 - 0.2j, a, 1.5 + 0.1j, b)
    assert np.allclose(b.download(), (0.3 - 0.2j) * x + (1.5 + 0.1j) * y, rtol=1e-14, atol=1e-13)
    b.upload(y)
    ctx.lanczos_update(a, b, 0.7, 1.9)
    assert np.allclose(a.download(), (y - 0.7 * x) / 1.9, atol=1e-14)
    assert np.allclose(b.download(), -1.9 * x, atol=1e-14)


def test_gaussian_batched_columns_independent():
    _say("gaussian batched")
    op = model.build_operator(lattice.reference_cluster(12), 20, 54.0)
    ctx = gpu.Context(op)
    v = ctx.vector(4, gpu.C128)
    ctx.set_initial_vec(v, 99)
    X = v.download()
    for c in range(4):
        ref = oracle.gaussian(ctx.dim, 99, col=c)
        assert np.linalg.norm(X[:, c] - ref) <= 1e-14


def test_lanczos_parity_n16():
    _say("lanczos n16")
    op = model.build_operator(lattice.reference_cluster(16), 20, 126.0, field=0.1)
    ctx = gpu.Context(op)
    v1, v0 = ctx.vector(), ctx.vector()
    ctx.set_initial_vec(v1, 20261202)
    x = v1.download()
    itr, alpha, beta, ene = ctx.lnc1(v1, v0, max_itr=1024, nvec=8)
    i2, a2, b2, e2 = oracle.lanczos(oracle.OracleOp(op), x, max_itr=1024, nvec=8)
    assert abs(ene[0] - e2[0]) <= 1e-10 * abs(e2[0])
    assert np.allclose(alpha[:20], a2[:20], rtol=1e-12, atol=0)
    assert np.allclose(beta[:20], b2[:20], rtol=1e-12, atol=0)


def test_lanczos_n24_300_steps():
    """The fixed-step Lanczos workload at its size: the reference 24-site cluster (2^24
    states), phi = 54 deg, 300 fixed steps.  The first 30 alpha/beta against the oracle's
    Lanczos from the same start vector, and the Ritz vector of the lowest Ritz value of the
    300 steps: residual |Hx - E0 x| / |x| at round-off and Rayleigh quotient equal to E0."""
    _say("lanczos n24 300 steps")
    op = model.build_operator(lattice.reference_cluster(24), 20, 54.0)
    ctx = gpu.Context(op)
    v1, v0, x = ctx.vector(), ctx.vector(), ctx.vector()
    seed = 20261202
    ctx.set_initial_vec(v1, seed)
    xs = v1.download()
    itr, alpha, beta, ene = ctx.lnc1(v1, v0, fixed_steps=300, nvec=1)
    assert itr == 300
    _, a2, b2, _ = oracle.lanczos(oracle.OracleOp(op), xs, fixed_steps=30, nvec=1)
    # alpha_1 is ~4e-4 here: bound by the coefficients' scale, not each one's size
    scale = np.max(np.abs(np.concatenate([a2[:30], b2[:30]])))
    assert np.allclose(alpha[:30], a2[:30], rtol=1e-11, atol=1e-12 * scale)
    assert np.allclose(beta[:30], b2[:30], rtol=1e-11, atol=1e-12 * scale)
    coef = gpu.vec12(alpha, beta, [ene[0]], 1)[0]
    ctx.set_initial_vec(v1, seed)
    ctx.ritz_vector(v1, v0, x, alpha, beta, coef)
    res = ctx.residual(x, ene[0], v0)
    X = x.download()
    HX = np.zeros_like(X)
    oracle.hv(oracle.OracleOp(op), X, HX)
    rq = np.vdot(X, HX).real / np.vdot(X, X).real
    print(f"[result] N=24 300-step Lanczos: E0/N {ene[0] / 24:.13f}, Ritz residual {res:.2e}, "
          f"oracle Rayleigh quotient rel diff {abs(rq - ene[0]) / abs(ene[0]):.2e}", flush=True)
    assert res <= 1e-10
    assert abs(rq - ene[0]) <= 1e-12 * abs(ene[0])


@pytest.mark.parametrize("spec,phi", [(8, 108.0), (12, 108.0)])
def test_lanczos_parity_spin1(spec, phi):
    """Spin-1 Lanczos (tiled spin-S kernel, deferred normalisation) against the oracle."""
    _say("lanczos spin1", spec)
    op = model.build_operator(lattice.reference_cluster(spec), 20, phi, q=3)
    ctx = gpu.Context(op)
    v1, v0 = ctx.vector(), ctx.vector()
    ctx.set_initial_vec(v1, 20261202)
    x = v1.download()
    itr, alpha, beta, ene = ctx.lnc1(v1, v0, max_itr=1024, nvec=4)
    i2, a2, b2, e2 = oracle.lanczos(oracle.OracleOp(op), x, max_itr=1024, nvec=4)
    assert abs(ene[0] - e2[0]) <= 1e-10 * abs(e2[0])
    assert np.allclose(alpha[:20], a2[:20], rtol=1e-12, atol=0)
    assert np.allclose(beta[:20], b2[:20], rtol=1e-12, atol=0)


@pytest.mark.parametrize("dtype", [gpu.C128, gpu.C64])
def test_lanczos_graphs_match_launches(dtype):
    """lnc1 (converging, and fixed steps) and dsflnc1z with their steps replayed from captured
    5-step graphs give the launch-by-launch run's alpha/beta, energies, stop step and final
    vectors bit for bit; complex64 takes the step whose beta kernel is a separate launch."""
    _say("lanczos graphs", dtype)
    op = model.build_operator(lattice.reference_cluster(16), 20, 126.0, field=0.1)
    ctx = gpu.Context(op)
    v1, v0 = ctx.vector(1, dtype), ctx.vector(1, dtype)
    runs = []
    for on in (True, False, True):   # graphs, launches, graphs again
        ctx.graphs(on)
        r = {}
        ctx.set_initial_vec(v1, 20261202)
        r["conv"] = ctx.lnc1(v1, v0, max_itr=1024, nvec=8)
        ctx.set_initial_vec(v1, 11)
        r["fixed"] = ctx.lnc1(v1, v0, fixed_steps=23, nvec=1)
        r["fixed_vecs"] = (v1.download(), v0.download())
        ctx.set_initial_vec(v1, 12)
        r["dsf"] = ctx.dsflnc1z(v1, v0, 0.5, 38)
        r["dsf_vecs"] = (v1.download(), v0.download())
        runs.append(r)
    ref = runs[1]
    for got in (runs[0], runs[2]):
        for key in ("conv", "fixed"):
            assert got[key][0] == ref[key][0], key
            for j in (1, 2, 3):
                assert np.array_equal(got[key][j], ref[key][j]), (key, j)
        for key in ("fixed_vecs", "dsf", "dsf_vecs"):
            for a, b in zip(got[key], ref[key]):
                assert np.array_equal(a, b), key
    assert ref["conv"][0] % 5 == 0 and ref["conv"][0] > 32


@pytest.mark.parametrize("q", [2, 3])
def test_one_column_batch_step_equals_lnc1(q):
    """A one-column lanczos_batch_step sequence (the serving path for large sizes) gives
    lnc1's alpha/beta bit for bit, also after a restart (mode 1) in the middle."""
    _say("one-column batch step", q)
    op = (model.build_operator(lattice.reference_cluster(16), 20, 126.0, field=0.1) if q == 2
          else model.build_operator(lattice.reference_cluster(8), 20, 108.0, q=3))
    ctx = gpu.Context(op)
    v1, v0 = ctx.vector(), ctx.vector()
    ctx.set_initial_vec(v1, 77)
    _, alpha, beta, _ = ctx.lnc1(v1, v0, fixed_steps=25, nvec=1)
    for _round in range(2):
        ctx.set_initial_vec(v1, 77)
        a_prev, b_prev = np.zeros(1), np.zeros(1)
        got_a, got_b = [], []
        for k in range(25):
            a, b = ctx.lanczos_batch_step(v1, v0, np.array([1 if k == 0 else 2], dtype=np.int32), a_prev, b_prev)
            got_a.append(a[0])
            got_b.append(b[0])
            a_prev, b_prev = np.atleast_1d(a).copy(), np.atleast_1d(b).copy()
        assert np.array_equal(np.array(got_a), alpha[:25]) and np.array_equal(np.array(got_b), beta[:25])


@pytest.mark.parametrize("q,steps", [(2, 1), (2, 2), (2, 7), (2, 12), (3, 2), (3, 9)])
def test_lanczos_leaves_the_reference_vectors(q, steps):
    """After lnc1 (fixed steps) and dsflnc1z the two vectors are what the reference's loop
    leaves: v1 the last Lanczos vector (unit norm), v0 = H v1 - beta v_prev, so
    <v1|v0> = alpha_last and |v0 - alpha_last v1| = beta_last -- also when the
    deferred-normalisation steps left the buffers swapped."""
    _say("lanczos vectors", q, steps)
    op = (model.build_operator(lattice.reference_cluster(16), 20, 126.0, field=0.1) if q == 2
          else model.build_operator(lattice.reference_cluster(8), 20, 108.0, q=3))
    ctx = gpu.Context(op)
    v1, v0 = ctx.vector(), ctx.vector()
    ctx.set_initial_vec(v1, 5)
    itr, alpha, beta, ene = ctx.lnc1(v1, v0, fixed_steps=steps, nvec=1)
    a, b = alpha[steps - 1], beta[steps - 1]
    for x1, x0 in ((v1.download(), v0.download()),):
        assert abs(np.linalg.norm(x1) - 1.0) <= 1e-13
        assert abs(np.vdot(x1, x0) - a) <= 1e-12 * np.linalg.norm(x0)
        assert abs(np.linalg.norm(x0 - a * x1) - b) <= 1e-12 * np.linalg.norm(x0)
    if steps >= 2:
        ctx.set_initial_vec(v1, 6)
        sa, sb = ctx.dsflnc1z(v1, v0, 0.5, steps + 1)
        x1, x0 = v1.download(), v0.download()
        a, b = sa[steps - 1], sb[steps]
        assert abs(np.linalg.norm(x1) - 1.0) <= 1e-13
        assert abs(np.vdot(x1, x0) - a) <= 1e-12 * np.linalg.norm(x0)
        assert abs(np.linalg.norm(x0 - a * x1) - b) <= 1e-12 * np.linalg.norm(x0)


def test_inverse_iteration_and_check1_n12():
    _say("inv1 n12")
    op = model.build_operator(lattice.reference_cluster(12), 20, 54.0)
    ctx = gpu.Context(op)
    v1, v0 = ctx.vector(), ctx.vector()
    ctx.set_initial_vec(v1, 5)
    itr, alpha, beta, ene = ctx.lnc1(v1, v0, nvec=8)
    x, b, p, r, y = (ctx.vector() for _ in range(5))
    ctx.set_initial_vec(b, 6)
    b0 = b.download()
    it = ctx.inv1(ene[0], x, b, p, r, y, max_inv=50)
    assert it > 0
    # same start vector and eigenvalue through the oracle's inv1z/cg1 restatement
    it_ref, x_ref = oracle.inverse_iteration(oracle.OracleOp(op), ene[0], b0, max_inv=50)
    assert it == it_ref
    assert abs(abs(np.vdot(x_ref, x.download())) - 1.0) < 1e-10
    # the reference's stopping rule |<x|b>|-1 < 1e-12 leaves a residual ~1e-7
    res = ctx.residual(x, ene[0], y)
    assert res < 1e-6
    ex, ratios = ctx.check1(x, y)
    assert abs(ex.real - ene[0]) < 1e-10 * abs(ene[0])
    # check1 logs (Hx)_j/x_j at j = min(ldim/3,13) : ldim : max(1,ldim/20), 1-based
    xv = x.download()
    hx = np.zeros_like(xv)
    oracle.hv(oracle.OracleOp(op), xv, hx)
    d = ctx.dim
    js = list(range(min(d // 3, 13), d + 1, max(1, d // 20)))
    assert len(ratios) == len(js)
    assert np.allclose(ratios, hx[np.array(js) - 1] / xv[np.array(js) - 1], rtol=1e-10)


def test_tps_parity_n12_batched():
    _say("tps n12")
    cl = lattice.reference_cluster(12)
    op = model.build_operator(cl, 20, 90.0, tps=True)
    ctx = gpu.Context(op)
    ctx.tps_set_observables(cl.site_vec, cl.sq_momenta())
    nvec = 4
    v1, v0, vg, vt = (ctx.vector(nvec, gpu.C128) for _ in range(4))
    ctx.set_initial_vec(v1, 11)
    X = v1.download()
    steps = 15
    al, be, co = ctx.tps(v1, v0, vg, vt, op.shift_L, steps)
    for c in range(nvec):
        a_ref, b_ref, c_ref = oracle.tps(oracle.OracleOp(op), cl.site_vec, cl.sq_momenta(), op.shift_L, steps,
                                         np.ascontiguousarray(X[:, c]))
        assert np.allclose(al[:, c], a_ref, rtol=1e-11)
        assert np.allclose(be[:, c], b_ref, rtol=1e-11)
        assert np.allclose(co[:, c, :], c_ref, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("spec", [8, 12])
def test_tps_parity_spin1(spec):
    """Spin-1 mTPQ (S_arbitrary/Hybrid_TPS9_TPS.f90:45-409): the partially shifted operator
    (S_arbitrary/Hybrid_TPS9_params.f90:1660-1686), the tiled spin-S H.v and the
    observables on the bits of the base-q index, against the oracle for two samples."""
    _say("tps spin1", spec)
    cl = lattice.reference_cluster(spec)
    op = model.build_operator(cl, 20, 108.0, q=3, field=0.1, tps=True)
    ctx = gpu.Context(op)
    ctx.tps_set_observables(cl.site_vec, cl.sq_momenta())
    v1, v0, vg, vt = (ctx.vector() for _ in range(4))
    steps = 12
    for seed in (31, 32):
        ctx.set_initial_vec(v1, seed)
        x = v1.download()
        al, be, co = ctx.tps(v1, v0, vg, vt, op.shift_L, steps)
        a_ref, b_ref, c_ref = oracle.tps(oracle.OracleOp(op), cl.site_vec, cl.sq_momenta(), op.shift_L, steps, x)
        assert np.allclose(np.ravel(al), a_ref, rtol=1e-11)
        assert np.allclose(np.ravel(be), b_ref, rtol=1e-11)
        assert np.allclose(np.asarray(co).reshape(steps, -1), c_ref, rtol=1e-9, atol=1e-12)


def test_dsf_parity_n12():
    _say("dsf n12")
    cl = lattice.reference_cluster(12)
    op = model.build_operator(cl, 20, 54.0)
    ctx = gpu.Context(op)
    v1, v0 = ctx.vector(), ctx.vector()
    ctx.set_initial_vec(v1, 3)
    x = v1.download()
    q = cl.q_list[3]
    w = ctx.vector()
    cn = ctx.intalv_zz(v1, w, q, cl.site_vec)
    vx = oracle.sz_q(cl.n, 2, cl.site_vec, q, x)
    cn_ref = np.vdot(vx, vx).real
    assert abs(cn - cn_ref / cl.n) <= 1e-13 * cn_ref
    assert np.linalg.norm(w.download() - vx / np.sqrt(cn_ref)) <= 1e-13
    sa, sb = ctx.dsflnc1z(w, v0, cn, 40)
    sa_ref, sb_ref = oracle.dsf_lanczos(oracle.OracleOp(op), vx / np.sqrt(cn_ref), cn, 40)
    assert np.allclose(sa[:20], sa_ref[:20], rtol=1e-11)
    assert np.allclose(sb[:20], sb_ref[:20], rtol=1e-11)
    ww, dsf, conv = ctx.renbun(sa, sb, -5.0, 0.01, 2000, 2.5e-3, 40, sq=cn)
    w2, d2, c2 = oracle.renbun(sa, sb, -5.0, 0.01, 2000, 2.5e-3, 40)
    d2 = oracle.normalize_dsf(d2, 2.5e-3, cn)
    assert np.allclose(dsf, d2, rtol=1e-10, atol=1e-14)
    assert np.array_equal(conv, c2)


@pytest.mark.parametrize("spec", [8, 12])
def test_dsf_parity_spin1(spec):
    """The spin-S branch of intalv_zz (weights d - S, S_arbitrary/Hybrid_TPS9_DSF.f90:18-89)
    at every momentum, then dsflnc1z and the continued fraction, against the oracle."""
    _say("dsf spin1", spec)
    cl = lattice.reference_cluster(spec)
    op = model.build_operator(cl, 20, 108.0, q=3)
    ctx = gpu.Context(op)
    v1, v0, w = ctx.vector(), ctx.vector(), ctx.vector()
    ctx.set_initial_vec(v1, 7)
    x = v1.download()
    for k, q in enumerate(cl.q_list):
        cn = ctx.intalv_zz(v1, w, q, cl.site_vec)
        vx = oracle.sz_q(cl.n, 3, cl.site_vec, q, x)
        cn_ref = np.vdot(vx, vx).real
        assert abs(cn - cn_ref / cl.n) <= 1e-13 * cn_ref, k
        assert np.linalg.norm(w.download() - vx / np.sqrt(cn_ref)) <= 1e-13, k
    sa, sb = ctx.dsflnc1z(w, v0, cn, 40)
    sa_ref, sb_ref = oracle.dsf_lanczos(oracle.OracleOp(op), vx / np.sqrt(cn_ref), cn, 40)
    assert np.allclose(sa[:20], sa_ref[:20], rtol=1e-11)
    assert np.allclose(sb[:20], sb_ref[:20], rtol=1e-11)
    ww, dsf, conv = ctx.renbun(sa, sb, -8.0, 0.05, 2000, 2.5e-3, 40, sq=cn)
    w2, d2, c2 = oracle.renbun(sa, sb, -8.0, 0.05, 2000, 2.5e-3, 40)
    d2 = oracle.normalize_dsf(d2, 2.5e-3, cn)
    assert np.allclose(dsf, d2, rtol=1e-10, atol=1e-14)
    assert np.array_equal(conv, c2)


def test_cocg_complex_shift_spin1():
    _say("cocg spin1 n8")
    cl = lattice.reference_cluster(8)
    op = model.build_operator(cl, 20, 108.0, q=3)
    ctx = gpu.Context(op)
    b, x, r, p, y = (ctx.vector() for _ in range(5))
    ctx.set_initial_vec(b, 4)
    bh = b.download()
    e0, eta = -3.0, 0.05
    # a CPU run of the same recurrence needs about 3300 iterations at this shift
    it, res = ctx.cocg(e0, eta, b, x, r, p, y, max_it=8000, tol=1e-10)
    assert res < 1e-10
    xs = x.download()
    # independent check: residual of the oracle's H.v, and a sparse direct solve
    hx = np.zeros_like(xs)
    oracle.hv(oracle.OracleOp(op), xs, hx)
    resid = hx - e0 * xs + 1j * eta * xs - bh
    assert np.linalg.norm(resid) <= 2e-10 * np.linalg.norm(bh)
    import scipy.sparse.linalg as spla
    from tests import dense
    J, K = dense.model20_couplings(108.0)
    H = dense.kh_dense(cl, J, K, K, K, q=3)
    import scipy.sparse as sp
    A = (sp.csr_matrix(H) - (e0 - 1j * eta) * sp.identity(H.shape[0])).tocsc()
    xd = spla.spsolve(A, bh)
    assert np.linalg.norm(xs - xd) <= 1e-8 * np.linalg.norm(xd)


def test_cocg_fused_shift_equals_generic_path():
    """The tiled spin-S kernel folds i eta p and p^T y into the H.v; the gather kernel
    takes the generic path (H.v, axpby, separate p^T y): same recurrence, same answer."""
    _say("cocg spin1 n12 fused vs generic")
    cl = lattice.reference_cluster(12)
    op = model.build_operator(cl, 20, 108.0, q=3)
    out = []
    for tiled in (True, False):
        ctx = gpu.Context(op)
        ctx.spin_kernel(tiled)
        b, x, r, p, y = (ctx.vector() for _ in range(5))
        ctx.set_initial_vec(b, 7)
        it, res = ctx.cocg(-5.0, 0.2, b, x, r, p, y, max_it=4000, tol=1e-10)
        out.append((it, res, x.download()))
        ctx.close()
    (it1, r1, x1), (it2, r2, x2) = out
    assert r1 < 1e-10 and r2 < 1e-10
    assert abs(it1 - it2) <= 2
    assert np.linalg.norm(x1 - x2) <= 1e-8 * np.linalg.norm(x2)


def _fortran_records(path):
    """Records of a Fortran unformatted sequential file (4-byte markers, no subrecords)."""
    raw = open(path, "rb").read()
    out, k = [], 0
    while k < len(raw):
        n = int(np.frombuffer(raw[k:k + 4], dtype=np.int32)[0])
        out.append(raw[k + 4:k + 4 + n])
        k += n + 8
    return out


def test_fortran_reference_call_sequence(tmp_path):
    """The ISO_C_BINDING layer as a drop-in: a Fortran host runs the reference's call
    sequence (S_one_half/Hybrid_TPS_main.f90:89-143) through the same-name wrappers --
    parallelized_lnc1, parallelized_inv1, readwrite_vec_data, check1, calc_DSF, then
    change_to_TPS_hamiltonian and parallelized_TPS -- and every printed value and result
    file is checked against the oracle run on the same start vectors."""
    _say("fortran reference call sequence")
    import os
    import re
    import subprocess
    exe = os.path.join(os.path.dirname(gpu.LIB_PATH), "fortran", "_build", "pned_main_driver")
    assert os.path.exists(exe), "build the Fortran layer first (__graft_entry__.build())"
    cl = lattice.reference_cluster(12)
    op = model.build_operator(cl, 20, 54.0, field=0.1)
    opt = model.build_operator(cl, 20, 54.0, field=0.1, tps=True)
    seed, max_itr, s_maxitr, i_range, first, last, nvec = 20261202, 400, 40, 200, 2, 3, 8
    split, delt = 2.5e-2, 0.05
    sq_q = cl.sq_momenta()
    f = tmp_path / "tables.bin"
    with open(f, "wb") as fh:
        np.array([op.n, op.q, op.ibond], dtype=np.int32).tofile(fh)
        np.asarray(op.ipair, dtype=np.int32).tofile(fh)
        np.asarray(op.ibond_type, dtype=np.int32).tofile(fh)
        op.fortran_xlocal().astype(np.complex128).tofile(fh)
        op.fortran_llocal().astype(np.int32).tofile(fh)
        opt.fortran_xlocal().astype(np.complex128).tofile(fh)
        opt.fortran_llocal().astype(np.int32).tofile(fh)
        np.array([opt.shift_L]).tofile(fh)
        np.array([len(cl.q_list)], dtype=np.int32).tofile(fh)
        np.ascontiguousarray(cl.q_list, dtype=np.float64).tofile(fh)
        np.array([len(sq_q)], dtype=np.int32).tofile(fh)
        np.ascontiguousarray(sq_q, dtype=np.float64).tofile(fh)
        np.ascontiguousarray(cl.site_vec, dtype=np.float64).tofile(fh)
        np.array([max_itr, s_maxitr, i_range, first, last, nvec], dtype=np.int32).tofile(fh)
        np.array([split, delt]).tofile(fh)
        np.array([seed], dtype=np.int64).tofile(fh)
    out = subprocess.run([exe, str(f)], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    assert out.returncode == 0 and "DONE" in out.stdout, out.stdout + out.stderr
    vals = dict(line.split(None, 1) for line in out.stdout.splitlines() if line.split()[:1] in (["E0"], ["ITR"]))
    oop = oracle.OracleOp(op)
    d = op.dim
    g = oracle.gaussian(d, seed)
    itr, _, _, ene = oracle.lanczos(oop, g, max_itr=max_itr, nvec=nvec)
    e0 = float(vals["E0"])
    assert abs(e0 - ene[0]) <= 1e-10 * abs(ene[0])
    # stop step of the 1e-13 change test (checks every 5 steps): bounded, not matched, as
    # rounding at 1e-15 in alpha/beta can move it by a check or two
    assert abs(int(vals["ITR"]) - itr) <= 10, (vals["ITR"], itr)
    # inverse iteration from the same seeded start: basisme00000 holds E0 and x
    rec = _fortran_records(tmp_path / "basisme00000")
    assert np.frombuffer(rec[0], dtype=np.float64)[0] == e0
    x = np.frombuffer(rec[1], dtype=np.complex128)
    _, x_ref = oracle.inverse_iteration(oop, e0, g)
    assert abs(abs(np.vdot(x_ref, x)) - 1.0) < 1e-10
    # check1's log: <x|H|x> and the ratios (Hx)_j / x_j
    log = (tmp_path / "acc_log_me00000.txt").read_text()
    m = re.search(r"<x\*H\*x> =\(\s*([-+.\dEe]+),\s*([-+.\dEe]+)\)", log)
    hx = np.zeros_like(x)
    oracle.hv(oop, x, hx)
    assert abs(float(m.group(1)) - np.vdot(x, hx).real) <= 2e-8 * abs(e0)     # 1pe16.8
    nums = [complex(float(a), float(b)) for a, b in re.findall(r"\(\s*([-+.\dEe]+),\s*([-+.\dEe]+)\)",
                                                                log.split("H*x(j)/x(j)")[1])]
    js = np.array(list(range(min(d // 3, 13), d + 1, max(1, d // 20))))
    assert np.allclose(nums, hx[js - 1] / x[js - 1], rtol=2e-6)                # 1pe16.6
    # calc_DSF at momenta first..last: sq.txt, SalphaSbeta_q*, dsf_res_zz.dat*
    sq_lines = (tmp_path / "sq.txt").read_text().splitlines()
    blocks = (tmp_path / "first_dsf_res_zz.datcl00000lme00000").read_text().split("#")[1:]
    ext_blocks = (tmp_path / "dsf_res_zz.datcl00000lme00000").read_text().split("#")[1:]
    assert len(sq_lines) == len(blocks) == last - first + 1
    pctx = gpu.Context(op)
    for k, iq in enumerate(range(first, last + 1)):
        qv = cl.q_list[iq - 1]
        vx = oracle.sz_q(cl.n, 2, cl.site_vec, qv, x)
        cn = np.vdot(vx, vx).real / cl.n
        assert abs(float(sq_lines[k].split()[-1]) - cn) <= 1e-15 + 1e-12 * cn
        _, sa, sb = gpu.read_dsf_alpha_beta(tmp_path / f"first_SalphaSbeta_q{iq:05d}me00000zz", s_maxitr, s_maxitr)
        sa_ref, sb_ref = oracle.dsf_lanczos(oop, vx / np.sqrt(cn * cl.n), cn, s_maxitr)
        assert np.allclose(sa[:20], sa_ref[:20], rtol=1e-10)
        assert np.allclose(sb[:20], sb_ref[:20], rtol=1e-10)
        w2, d2, c2 = oracle.renbun(sa_ref, sb_ref, e0, delt, i_range, split, s_maxitr)
        d2 = oracle.normalize_dsf(d2, split, cn)
        rows = np.array([[float(t) for t in ln.split()] for ln in blocks[k].splitlines()[1:] if ln.strip()])
        assert rows.shape == (i_range, 5)
        assert np.allclose(rows[:, 2], w2, atol=1e-10)
        assert np.allclose(rows[:, 3], d2, rtol=1e-7, atol=2e-16 * np.abs(d2).max() + 1e-15)
        # calc_extended_DSF continued those files to 2 s_maxitr steps: the library's own
        # continuation (pned_dsflnc1z_span from indx + 1, bitwise equal to one long span and
        # held to the oracle in tests/test_gpu_resume.py) from the same files, bit for bit
        s2 = 2 * s_maxitr
        _, ea, eb = gpu.read_dsf_alpha_beta(tmp_path / f"SalphaSbeta_q{iq:05d}me00000zz", s2, s2)
        w1, w0 = pctx.vector(), pctx.vector()
        indx = pctx.read_vectors(tmp_path / f"first_vec_q{iq:05d}me00000zz", [w1, w0])
        ca, cb = np.zeros(s2), np.zeros(s2 + 1)
        ca[: indx + 1], cb[: indx + 1] = sa[: indx + 1], sb[: indx + 1]
        pctx.dsflnc1z_span(w1, w0, cn, ca, cb, first=indx + 1)
        assert np.array_equal(ea, ca) and np.array_equal(eb, cb[:s2])
        w3, d3, c3 = pctx.renbun(ca, cb, e0, delt, i_range, split, s2, sq=cn)
        rows = np.array([[float(t) for t in ln.split()] for ln in ext_blocks[k].splitlines()[1:] if ln.strip()])
        assert np.allclose(rows[:, 3], d3, rtol=1e-14, atol=1e-15 * np.abs(d3).max())
        assert np.array_equal(rows[:, 4].astype(int), c3)
    # parallelized_TPS: one SalphaSbeta_me00000seed<i>zz per sample, start vector seed + i
    oopt = oracle.OracleOp(opt)
    for i in range(first, last + 1):
        a, b, c = gpu.read_alpha_beta_coef(tmp_path / f"SalphaSbeta_me00000seed{i:05d}zz")
        a_ref, b_ref, c_ref = oracle.tps(oopt, cl.site_vec, sq_q, opt.shift_L, max_itr, oracle.gaussian(d, seed + i))
        assert np.allclose(a, a_ref, rtol=1e-11)
        assert np.allclose(b, b_ref, rtol=1e-11)
        assert np.allclose(c[:, :c_ref.shape[1]], c_ref, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("dtype,tol", [(gpu.C128, 1e-10), (gpu.C64, 2e-4)])
def test_tps_fused_matches_oracle(dtype, tol):
    _say("tps fused", dtype)
    cl = lattice.reference_cluster(12)
    op = model.build_operator(cl, 20, 90.0, tps=True)
    ctx = gpu.Context(op)
    ctx.tps_set_observables(cl.site_vec, cl.sq_momenta())
    nvec = 8
    v1, va, vb = (ctx.vector(nvec, dtype) for _ in range(3))
    ctx.set_initial_vec(v1, 21)
    X = v1.download()
    steps = 25
    al, be, co = ctx.tps_fused(v1, va, vb, op.shift_L, steps)
    for c in (0, 5):
        x = np.ascontiguousarray(X[:, c]).astype(np.complex128)
        a_ref, b_ref, c_ref = oracle.tps(oracle.OracleOp(op), cl.site_vec, cl.sq_momenta(), op.shift_L, steps, x)
        assert np.allclose(al[:, c], a_ref, rtol=tol), c
        assert np.allclose(be[:, c], b_ref, rtol=tol), c
        assert np.allclose(co[:, c, :], c_ref, rtol=max(tol, 1e-9), atol=1e-9 if dtype == gpu.C128 else 1e-4), c


@pytest.mark.parametrize("dtype", [gpu.C128, gpu.C64])
@pytest.mark.parametrize("steps,graph_steps", [(20, 6), (4, 0), (32, 3)])
def test_tps_fused_graph_matches_launches(dtype, steps, graph_steps):
    """The fused mTPQ run replayed from a captured HIP graph gives the one-launch-at-a-time
    run's alpha, beta and coef rows bit for bit: step 1 un-captured, then whole replays
    (each read back and checked) and the remaining steps un-captured; (4, 0) never replays."""
    _say("tps fused graph", dtype, steps, graph_steps)
    cl = lattice.reference_cluster(12)
    op = model.build_operator(cl, 20, 90.0, tps=True)
    ctx = gpu.Context(op)
    ctx.tps_set_observables(cl.site_vec, cl.sq_momenta())
    nvec = 8
    v1, va, vb = (ctx.vector(nvec, dtype) for _ in range(3))
    runs = []
    for g in (None, graph_steps, None):   # launches, graph, launches again (context still usable)
        ctx.set_initial_vec(v1, 21)
        runs.append(ctx.tps_fused(v1, va, vb, op.shift_L, steps, graph_steps=g))
    for k in range(3):
        assert np.array_equal(runs[1][k], runs[0][k]), k
        assert np.array_equal(runs[2][k], runs[0][k]), k
    assert np.all(np.isfinite(runs[1][2]))


def test_tps_fused_c64_n24_r64():
    """The batched form at its configured size: 3x4 torus (N=24), R=64 samples, mTPQ steps of
    (l - H/N) with the bond-norm shift.  The fused complex64 run against the fused complex128
    run on the same 64 start vectors (every column's E/N at every step within 1e-6), and the
    first steps of two columns of each against the oracle's mTPQ restatement of the same start
    vector (the oracle takes about 5 s per step at this size).  Both precisions run the H.V in
    column groups (one 256-byte run of columns per tile), which takes 3 passes here where all
    64 columns in the tile take 4."""
    _say("tps fused c64 N=24 R=64")
    cl = lattice.torus(3, 4)
    lb = model.bond_norm_shift(cl, 20, 90.0)
    op = model.build_operator(cl, 20, 90.0, tps=True, tps_lb=lb)
    ctx = gpu.Context(op)
    ctx.tps_set_observables(cl.site_vec, cl.sq_momenta())
    R, k = 64, 20
    seeds = [20261202 + 7919 * c for c in range(R)]
    runs = {}
    for dtype in (gpu.C64, gpu.C128):
        v1, va, vb = (ctx.vector(R, dtype) for _ in range(3))
        for c in range(R):
            ctx.set_initial_column(v1, c, seeds[c])
        runs[dtype] = ctx.tps_fused(v1, va, vb, op.shift_L, k)
        assert ctx.plan_info()[0] == 3, dtype
        for v in (v1, va, vb):
            v.free()
    a64, a128 = runs[gpu.C64][0], runs[gpu.C128][0]
    e_err = np.max(np.abs(a64 - a128)) / cl.n
    print(f"[result] N=24 R=64 fused complex64 vs complex128: max |dE/N| {e_err:.2e} over {k} steps", flush=True)
    assert e_err <= 1e-6
    oop = oracle.OracleOp(op)
    ko = 3
    for c in (0, 37):
        a_ref, b_ref, c_ref = oracle.tps(oop, cl.site_vec, cl.sq_momenta(), op.shift_L, ko,
                                         oracle.gaussian(op.dim, seeds[c]))
        for dtype, tol in ((gpu.C128, 1e-10), (gpu.C64, 2e-4)):
            al, be, co = runs[dtype]
            assert np.allclose(al[:ko, c], a_ref, rtol=tol), (dtype, c)
            assert np.allclose(be[:ko, c], b_ref, rtol=tol), (dtype, c)
            assert np.allclose(co[:ko, c, :], c_ref, rtol=max(tol, 1e-9),
                               atol=1e-9 if dtype == gpu.C128 else 1e-4), (dtype, c)


def test_spin1_tiled_vs_gather_n18():
    """cfg4 size (3^18 states): the tiled spin-S kernel against the per-state gather kernel
    (both pinned to the oracle at N=8, 12 by test_hv_parity)."""
    _say("spin1 tiled n18")
    op = model.build_operator(lattice.reference_cluster(18), 20, 108.0, q=3)
    ctx = gpu.Context(op)
    v1, a, b = ctx.vector(), ctx.vector(), ctx.vector()
    ctx.set_initial_vec(v1, 7)
    ctx.spin_kernel(True)
    p1 = ctx.hv_fresh(v1, a, 0.02)
    ctx.spin_kernel(False)
    p2 = ctx.hv_fresh(v1, b, 0.02)
    ctx.axpby(1.0, a, -1.0, b)           # b = a - b (axpby: y = a x + b y)
    diff = np.sqrt(ctx.calc_vec_inprdct(b, b).real)
    ref = np.sqrt(ctx.calc_vec_inprdct(a, a).real)
    assert diff <= 1e-12 * ref
    assert abs(p1 - p2) <= 1e-12 * abs(p2)


def test_set_matrices_refits_plan():
    """Replacing the bond matrices on a planned context keeps its tile passes when the
    new operator fits them (coefficients refitted) and still matches the oracle; an
    operator without field terms fits with those coefficients zero."""
    _say("set_matrices refit")
    cl = lattice.reference_cluster(16)
    opa = model.build_operator(cl, 20, 126.0, field=0.1)
    ctx = gpu.Context(opa)
    x = _rand(ctx.dim, 21)
    v1, v0 = ctx.vector(), ctx.vector()
    v1.upload(x)
    ctx.hv_fresh(v1, v0)
    plan_a = ctx.plan_info()
    for phi, h in ((37.0, 0.25), (200.0, 0.0), (300.0, 0.4)):
        opb = model.build_operator(cl, 20, phi, field=h)
        ctx.set_matrices(opb)
        pr = ctx.hv_fresh(v1, v0, 0.03)
        want = np.zeros_like(x)
        pr_ref = oracle.hv(oracle.OracleOp(opb), x, want, eperbd=0.03)
        got = v0.download()
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want), (phi, h)
        assert abs(pr - pr_ref) <= 1e-13 * np.linalg.norm(x) * np.linalg.norm(want)
        assert ctx.plan_info() == plan_a, "plan rebuilt although the operator fits it"


@pytest.mark.parametrize("spec,k", [(12, 6), (16, 8), (16, 5), ("2x3", 5), (18, 9)])
def test_sector_hv_parity(spec, k):
    """S^z sector basis (parity_type = 0): H.v and <v|H|v> against the oracle's
    list-table H_dot_v on the same sector vectors."""
    _say("sector hv", spec, k)
    cl = lattice.get_cluster(spec)
    op = model.build_operator(cl, 20, 0.0)
    ctx = gpu.Context(op)
    ctx.sz_sector(k)
    list1, _, _ = oracle.sector_tables(cl.n, k)
    assert ctx.dim == len(list1)
    x = _rand(ctx.dim, 31)
    y = _rand(ctx.dim, 32)
    v1, v0 = ctx.vector(), ctx.vector()
    v1.upload(x)
    v0.upload(y)
    pr = ctx.p_mltply_v3(v1, v0, 0.05)
    want = y.copy()
    pr_ref = oracle.hv_sector(oracle.OracleOp(op), k, x, want, eperbd=0.05)
    assert np.linalg.norm(v0.download() - want) <= 1e-12 * np.linalg.norm(want)
    assert abs(pr - pr_ref) <= 1e-13 * np.linalg.norm(x) * np.linalg.norm(want - y)


@pytest.mark.parametrize("spec,k", [(12, 5), (16, 7), ("2x3", 6)])
def test_intalv_xy_and_sector_dsf(spec, k):
    """S^+(q) (intalv_xy) from a k-set-bit sector context to a (k+1) one against the
    oracle, then the DSF Lanczos in the target sector against the oracle's full-space
    Lanczos of the embedded vector (H conserves S^z, so the two runs are the same
    Lanczos)."""
    _say("intalv_xy + sector dsf", spec, k)
    cl = lattice.get_cluster(spec)
    op = model.build_operator(cl, 20, 0.0)
    src, dst = gpu.Context(op), gpu.Context(op)
    src.sz_sector(k)
    dst.sz_sector(k + 1)
    x = _rand(src.dim, 41)
    xv = src.vector()
    xv.upload(x)
    w, v0 = dst.vector(), dst.vector()
    q = cl.q_list[min(2, len(cl.q_list) - 1)]
    cn = dst.intalv_xy(xv, w, q, cl.site_vec)
    v_ref, cn_ref = oracle.intalv_xy(cl.n, cl.site_vec, q, k, x)
    assert abs(cn - cn_ref) <= 1e-13 * cn_ref
    assert np.linalg.norm(w.download() - v_ref) <= 1e-13
    with pytest.raises(gpu.PnedError):   # x must come from the sector with one bit fewer
        src.intalv_xy(xv, src.vector(), q, cl.site_vec)
    sa, sb = dst.dsflnc1z(w, v0, cn, 30)
    lst, _, _ = oracle.sector_tables(cl.n, k + 1)
    vf = np.zeros(1 << cl.n, complex)
    vf[lst.astype(np.int64)] = v_ref
    sa_ref, sb_ref = oracle.dsf_lanczos(oracle.OracleOp(op), vf, cn_ref, 30)
    assert np.allclose(sa[:20], sa_ref[:20], rtol=1e-11)
    assert np.allclose(sb[:20], sb_ref[:20], rtol=1e-11)


@pytest.mark.parametrize("spec,q,ipm", [(12, 2, 1), (16, 2, -1), (8, 3, 1), (12, 3, -1), (8, 5, 1)])
def test_intalv_pm_full_space(spec, q, ipm):
    """Full-space S^{+/-}(q) (parity_type != 0) against the oracle, for the S=1/2 tree and
    spin S > 1/2 with ladder factors; for spin 1 also the DSF Lanczos that follows."""
    _say("intalv_pm full space", spec, q, ipm)
    cl = lattice.get_cluster(spec)
    op = model.build_operator(cl, 20, 108.0, q=q) if q > 2 else model.build_operator(cl, 20, 54.0, field=0.1)
    ctx = gpu.Context(op)
    x = _rand(ctx.dim, 51)
    xv, w, v0 = ctx.vector(), ctx.vector(), ctx.vector()
    xv.upload(x)
    qq = cl.q_list[min(1, len(cl.q_list) - 1)]
    cn = ctx.intalv_xy(xv, w, qq, cl.site_vec, ipm)
    v_ref, cn_ref = oracle.intalv_pm(cl.n, q, cl.site_vec, qq, ipm, x)
    assert abs(cn - cn_ref) <= 1e-13 * cn_ref
    assert np.linalg.norm(w.download() - v_ref) <= 1e-13
    if q == 3:
        sa, sb = ctx.dsflnc1z(w, v0, cn, 30)
        sa_ref, sb_ref = oracle.dsf_lanczos(oracle.OracleOp(op), v_ref, cn_ref, 30)
        assert np.allclose(sa[:20], sa_ref[:20], rtol=1e-11)
        assert np.allclose(sb[:20], sb_ref[:20], rtol=1e-11)


def test_intalv_xy_sector_lowering_matches_full_space():
    """S=1/2 sectors with ipm = -1 (k -> k-1) equal the full-space bit-clearing sum of the
    embedded vector, restricted to the target sector."""
    _say("intalv_xy sector ipm=-1")
    cl = lattice.get_cluster(16)
    op = model.build_operator(cl, 20, 0.0)
    src, dst = gpu.Context(op), gpu.Context(op)
    src.sz_sector(8)
    dst.sz_sector(7)
    x = _rand(src.dim, 61)
    xv, w = src.vector(), dst.vector()
    xv.upload(x)
    qq = cl.q_list[2]
    cn = dst.intalv_xy(xv, w, qq, cl.site_vec, -1)
    s_src, _, _ = oracle.sector_tables(cl.n, 8)
    s_dst, _, _ = oracle.sector_tables(cl.n, 7)
    xf = np.zeros(1 << cl.n, complex)
    xf[s_src.astype(np.int64)] = x
    vf, cnf = oracle.intalv_pm(cl.n, 2, cl.site_vec, qq, -1, xf)
    assert abs(cn - cnf) <= 1e-13 * cnf
    assert np.linalg.norm(w.download() - vf[s_dst.astype(np.int64)]) <= 1e-13


def test_sector_lanczos_ground_state():
    """The Heisenberg ground state of the 16-site cluster lies in the S^z = 0 sector:
    Lanczos in that sector reaches the oracle's full-space ground energy."""
    _say("sector lanczos")
    cl = lattice.reference_cluster(16)
    op = model.build_operator(cl, 20, 0.0)
    ctx = gpu.Context(op)
    ctx.sz_sector(8)
    v1, v0 = ctx.vector(), ctx.vector()
    ctx.set_initial_vec(v1, 7)
    itr, alpha, beta, ene = ctx.lnc1(v1, v0, max_itr=1024, nvec=2)
    x = oracle.gaussian(2 ** cl.n, 7)
    _, _, _, e_full = oracle.lanczos(oracle.OracleOp(op), x, max_itr=1024, nvec=2)
    assert abs(ene[0] - e_full[0]) <= 1e-10 * abs(e_full[0])


def test_sector_rejects_non_conserving_operator():
    _say("sector reject")
    op = model.build_operator(lattice.reference_cluster(12), 20, 54.0)
    ctx = gpu.Context(op)
    with pytest.raises(gpu.PnedError):
        ctx.sz_sector(6)
    # back to the full space: vectors of the other basis are refused
    op0 = model.build_operator(lattice.reference_cluster(12), 20, 0.0)
    ctx0 = gpu.Context(op0)
    a, b = ctx0.vector(), ctx0.vector()
    ctx0.sz_sector(6)
    with pytest.raises(gpu.PnedError):
        ctx0.p_mltply_v3(a, b)


def test_ctpq_from_gpu_mtpq():
    """cTPQ post-processing of a batched mTPQ run on the GPU (4 samples, fused complex128)
    against the oracle's mTPQ + cTPQ restatements of the same samples."""
    _say("ctpq from gpu mtpq")
    cl = lattice.reference_cluster(12)
    op = model.build_operator(cl, 20, 90.0, tps=True)
    ctx = gpu.Context(op)
    ctx.tps_set_observables(cl.site_vec, cl.sq_momenta())
    nvec, steps = 4, 150
    v1, va, vb = (ctx.vector(nvec, gpu.C128) for _ in range(3))
    ctx.set_initial_vec(v1, 41)
    X = v1.download()
    al, be, co = ctx.tps_fused(v1, va, vb, op.shift_L, steps)
    A, B, Cf = gpu.tps_arrays(al, be, co, nvec, co.shape[-1])
    got = gpu.ctpq(A, B, Cf, 0.1, 10.0, 30)
    RA, RB, RC = [], [], []
    for c in range(nvec):
        a_ref, b_ref, c_ref = oracle.tps(oracle.OracleOp(op), cl.site_vec, cl.sq_momenta(), op.shift_L, steps,
                                         np.ascontiguousarray(X[:, c]))
        RA.append(a_ref)
        RB.append(b_ref)
        RC.append(c_ref.T)
    want = oracle.ctpq(np.array(RA), np.array(RB), np.array(RC), 0.1, 10.0, 30)
    assert np.allclose(got[:, :2], want[:, :2], rtol=1e-7, atol=1e-9)


def test_vectors_outlive_their_context():
    """pned_ctx_destroy detaches live vectors: freeing them later is safe and another
    context refuses them."""
    _say("vectors after their context is destroyed")
    op = model.build_operator(lattice.get_cluster(8), 20, 54.0)
    ctx = gpu.Context(op)
    v, w = ctx.vector(), ctx.vector()
    ctx.set_initial_vec(v, 1)
    ctx.close()
    other = gpu.Context(op)
    with pytest.raises(gpu.PnedError):
        other.calc_vec_inprdct(v, v)
    v.free()
    w.free()
    u = other.vector()
    other.set_initial_vec(u, 1)
    assert abs(other.calc_vec_inprdct(u, u) - 1.0) < 1e-14
    other.close()
    u.free()


def test_argument_errors_leave_the_context_usable():
    """Bad calls come back as PnedError with the library's code, before any kernel runs,
    and the context keeps working afterwards."""
    _say("argument errors")
    op = model.build_operator(lattice.get_cluster(12), 20, 126.0, field=0.1)
    ctx = gpu.Context(op)
    v1, v0 = ctx.vector(), ctx.vector()
    wide = ctx.vector(2)
    three = ctx.vector(3)
    small = ctx.vector(1, gpu.C64)
    ctx.set_initial_vec(v1, 3)
    cl = lattice.get_cluster(12)
    ctx.tps_set_observables(cl.site_vec, cl.sq_momenta())
    other = gpu.Context(model.build_operator(lattice.get_cluster(8), 20, 54.0))
    foreign = other.vector()                                                # 2^8 states, another context
    one = np.ones(1, np.int32)
    cases = [
        (lambda: ctx.p_mltply_v3(v1, v1), gpu.E_ARG),                       # v1 and v0 the same vector
        (lambda: ctx.lnc1(v1, v1, fixed_steps=5), gpu.E_ARG),
        (lambda: ctx.lnc1(foreign, v0, fixed_steps=5), gpu.E_ARG),
        (lambda: ctx.lanczos_batch_steps(v1, v1, one, np.zeros(1), np.zeros(1), 5), gpu.E_ARG),
        (lambda: ctx.lanczos_batch_steps(foreign, v0, one, np.zeros(1), np.zeros(1), 5), gpu.E_ARG),
        (lambda: ctx.lanczos_batch_step(v1, foreign, one, np.zeros(1), np.zeros(1)), gpu.E_ARG),
        (lambda: ctx.tps_fused(v1, v0, v0, 1.0, 2), gpu.E_ARG),
        (lambda: ctx.tps_fused(v1, foreign, v0, 1.0, 2), gpu.E_ARG),
        (lambda: ctx.tps_fused(v1, v0, v0, 1.0, 2, graph_steps=6), gpu.E_ARG),
        (lambda: ctx.tps_fused(v1, wide, v0, 1.0, 2, graph_steps=6), gpu.E_ARG),
        (lambda: ctx.tps_fused(v1, small, v0, 1.0, 8, graph_steps=4), gpu.E_ARG),   # not a multiple of 3
        (lambda: ctx.autotune(v1, v1), gpu.E_ARG),
        (lambda: ctx.autotune(foreign, v0), gpu.E_ARG),
        (lambda: ctx.p_mltply_v3(v1, wide), gpu.E_ARG),                     # column counts differ
        (lambda: ctx.axpby(1.0, v1, 1.0, small), gpu.E_ARG),                # storage types differ
        (lambda: ctx.p_mltply_v3(three, ctx.vector(3)), gpu.E_UNSUPPORTED),  # 3 columns
        (lambda: ctx.intalv_xy(v1, v0, [0.0, 0.0], np.zeros((12, 2)), ipm=2), gpu.E_ARG),
        (lambda: ctx.lanczos_batch_steps(v1, v0, np.ones(1, np.int32), np.zeros(1), np.zeros(1), 0), gpu.E_ARG),
        (lambda: ctx.cocg(-3.0, 0.1, v1, v0, *(ctx.vector() for _ in range(3))), gpu.E_UNSUPPORTED),  # complex H
    ]
    for k, (call, code) in enumerate(cases):
        with pytest.raises(gpu.PnedError) as err:
            call()
        assert err.value.code == code, (k, err.value)
    v0.zero()
    pr = ctx.p_mltply_v3(v1, v0)
    want = np.zeros(ctx.dim, complex)
    p_ref = oracle.hv(oracle.OracleOp(op), v1.download(), want)
    assert np.linalg.norm(v0.download() - want) <= 1e-12 * np.linalg.norm(want)
    assert abs(pr - p_ref) <= 1e-12 * abs(p_ref)
