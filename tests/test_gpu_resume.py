"""Checkpoint / restart and the DSF extension on the GPU (the reference's interval_dump /
interval_dump_read restart branches, S_one_half/Hybrid_TPS_Lanczos.f90:87-93,
Hybrid_TPS_TPS.f90:352-358, Hybrid_TPS_DSF.f90:213-225, and calc_extended_DSF,
Hybrid_TPS_DSF.f90:515-553).  A run stopped at the checkpoint point, written to the
reference's files, read back into fresh vectors and continued gives alpha/beta equal bit
for bit to the same entry point run in one span; the one-span results match the oracle."""
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


@pytest.mark.parametrize("spec,q,stop", [(16, 2, 57), ("3x4", 2, 41), (8, 3, 38)])
def test_lnc1_resume_is_bitwise(tmp_path, spec, q, stop):
    _say("lnc1 resume", spec, q, stop)
    cl = lattice.get_cluster(spec)
    op = model.build_operator(cl, 20, 126.0, q=q, field=0.1)
    ctx = gpu.Context(op)
    v1, v0 = ctx.vector(), ctx.vector()
    max_itr, nvec = 600, 4
    a1, b1 = np.zeros(max_itr), np.zeros(max_itr)
    ctx.set_initial_vec(v1, 9)
    x = v1.download()
    itr1, ene1 = ctx.lnc1_span(v1, v0, a1, b1, max_itr=max_itr, nvec=nvec)
    assert itr1 > stop
    if ctx.dim <= 1 << 16:   # oracle: the reference's loop on the same start vector
        itr_ref, a_ref, b_ref, e_ref = oracle.lanczos(oracle.OracleOp(op), x, max_itr=max_itr, nvec=nvec)
        # the stop step is where the 1e-13 relative change test between checks (every 5
        # steps) first passes; rounding at 1e-15 in alpha/beta (another summation order) can
        # move it by a check or two, so it is bounded, not matched
        assert abs(itr1 - itr_ref) <= 10, (itr1, itr_ref)
        assert abs(ene1[0] - e_ref[0]) < 1e-10 * abs(e_ref[0])
        assert np.allclose(a1[:20], a_ref[:20], rtol=1e-12)
    # the same run stopped at the top of step `stop`, checkpointed, resumed in new vectors
    a2, b2 = np.zeros(max_itr), np.zeros(max_itr)
    ctx.set_initial_vec(v1, 9)
    itr, _ = ctx.lnc1_span(v1, v0, a2, b2, stop=stop, max_itr=max_itr, nvec=nvec)
    assert itr == stop
    f = tmp_path / "vecdump_me00000"
    coef = np.zeros((max_itr, nvec + 1))
    ctx.interval_dump(f, ctx.dim, max_itr, reals=[a2, b2, coef, np.zeros(nvec)], vecs=[v0, v1])
    w1, w0 = ctx.vector(), ctx.vector()
    ds, vs, (ra, rb, _, _) = ctx.interval_dump_read(f, [max_itr, max_itr, max_itr * (nvec + 1), nvec], [w0, w1])
    assert (ds, vs) == (ctx.dim, max_itr)
    itr2, ene2 = ctx.lnc1_span(w1, w0, ra, rb, first=stop, max_itr=max_itr, nvec=nvec)
    assert itr2 == itr1
    assert np.array_equal(ra[:itr1], a1[:itr1]) and np.array_equal(rb[:itr1], b1[:itr1])
    assert np.array_equal(ene2, ene1)


def test_tps_resume_is_bitwise(tmp_path):
    _say("tps resume")
    cl = lattice.reference_cluster(12)
    op = model.build_operator(cl, 20, 90.0, tps=True)
    ctx = gpu.Context(op)
    ctx.tps_set_observables(cl.site_vec, cl.sq_momenta())
    nc, steps, stop = 1, 24, 11   # the reference runs (and checkpoints) one sample at a time
    ncoef = 7 + len(cl.sq_momenta())
    v1, v0, vg, vt = (ctx.vector(nc) for _ in range(4))
    A, B, Cf = np.zeros(steps * nc), np.zeros(steps * nc), np.zeros(steps * nc * ncoef)
    ctx.set_initial_vec(v1, 13)
    X = v1.download()
    ctx.tps_span(v1, v0, vg, vt, op.shift_L, A, B, Cf)
    X = X.reshape(-1, nc)
    for c in range(nc):
        a_ref, b_ref, c_ref = oracle.tps(oracle.OracleOp(op), cl.site_vec, cl.sq_momenta(), op.shift_L, steps,
                                         np.ascontiguousarray(X[:, c]))
        assert np.allclose(A.reshape(steps, nc)[:, c], a_ref, rtol=1e-11)
        assert np.allclose(Cf.reshape(steps, nc, ncoef)[:, c, :], c_ref, rtol=1e-9, atol=1e-12)
    A2, B2, C2 = np.zeros_like(A), np.zeros_like(B), np.zeros_like(Cf)
    ctx.set_initial_vec(v1, 13)
    ctx.tps_span(v1, v0, vg, vt, op.shift_L, A2, B2, C2, stop=stop)
    f = tmp_path / "vecdump_me00000"
    ctx.interval_dump(f, ctx.dim, steps, reals=[A2, B2, C2], vecs=[v0, v1, vg])
    w1, w0, wg, wt = (ctx.vector(nc) for _ in range(4))
    _, _, (ra, rb, rc) = ctx.interval_dump_read(f, [A.size, B.size, Cf.size], [w0, w1, wg])
    ctx.tps_span(w1, w0, wg, wt, op.shift_L, ra, rb, rc, first=stop)
    assert np.array_equal(ra, A) and np.array_equal(rb, B) and np.array_equal(rc, Cf)


@pytest.mark.parametrize("q", [2, 3])
def test_dsf_resume_and_extension_are_bitwise(tmp_path, q):
    _say("dsf resume and extension", q)
    cl = lattice.reference_cluster(12 if q == 2 else 8)
    op = model.build_operator(cl, 20, 54.0 if q == 2 else 108.0, q=q)
    ctx = gpu.Context(op)
    x, v1, v0 = ctx.vector(), ctx.vector(), ctx.vector()
    ctx.set_initial_vec(x, 3)
    qv = cl.q_list[1]
    cn = ctx.intalv_zz(x, v1, qv, cl.site_vec)
    exc = v1.download()
    s_full, s_short, stop = 60, 35, 23
    sa, sb = np.zeros(s_full), np.zeros(s_full + 1)
    icf = ctx.dsflnc1z_span(v1, v0, cn, sa, sb)
    assert icf == s_full - 1
    sa_ref, sb_ref = oracle.dsf_lanczos(oracle.OracleOp(op), exc, cn, s_full)
    assert np.allclose(sa[:20], sa_ref[:20], rtol=1e-11) and np.allclose(sb[:20], sb_ref[:20], rtol=1e-11)
    # checkpoint at the top of step `stop` (interval_dump: vec1 = v1, vec2 = v0), then resume
    sa2, sb2 = np.zeros(s_full), np.zeros(s_full + 1)
    v1.upload(exc)
    ctx.dsflnc1z_span(v1, v0, cn, sa2, sb2, stop=stop)
    f = tmp_path / "vecdump_me00000"
    ctx.interval_dump(f, ctx.dim, s_full, reals=[sa2, sb2], vecs=[v1, v0])
    w1, w0 = ctx.vector(), ctx.vector()
    _, _, (ra, rb) = ctx.interval_dump_read(f, [s_full, s_full + 1], [w1, w0])
    ctx.dsflnc1z_span(w1, w0, cn, ra, rb, first=stop)
    assert np.array_equal(ra, sa) and np.array_equal(rb, sb)
    # COMP=zr extension: a shorter run's SalphaSbeta_q / vec_q files (dump_alpha_beta),
    # read back (read_alpha_beta2) and continued from indx + 1 to the longer s_maxitr
    sa3, sb3 = np.zeros(s_short), np.zeros(s_short + 1)
    v1.upload(exc)
    icf = ctx.dsflnc1z_span(v1, v0, cn, sa3, sb3)
    fa, fv = tmp_path / "SalphaSbeta_q00002me00000zz", tmp_path / "vec_q00002me00000zz"
    gpu.dump_dsf_alpha_beta(fa, cn, sa3, sb3[:s_short])
    ctx.dump_vectors(fv, icf, [v1, v0])
    w1, w0 = ctx.vector(), ctx.vector()
    indx = ctx.read_vectors(fv, [w1, w0])
    assert indx == s_short - 1
    sq, ta, tb = gpu.read_dsf_alpha_beta(fa, indx + 1, indx + 1)
    ea, eb = np.zeros(s_full), np.zeros(s_full + 1)
    ea[: indx + 1], eb[: indx + 1] = ta, tb
    ctx.dsflnc1z_span(w1, w0, sq.real, ea, eb, first=indx + 1)
    assert np.array_equal(ea, sa) and np.array_equal(eb[:s_full], sb[:s_full])
    # and the spectrum of the extended coefficients is the one of the long run
    ww, d1, c1 = ctx.renbun(sa, sb, -5.0, 0.05, 400, 1e-2, s_full, sq=cn)
    ww, d2, c2 = ctx.renbun(ea, eb, -5.0, 0.05, 400, 1e-2, s_full, sq=cn)
    assert np.array_equal(d1, d2) and np.array_equal(c1, c2)
