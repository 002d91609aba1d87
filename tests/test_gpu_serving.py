"""Batch contexts and the continuous-batching server (GPU, through the C ABI).

* per-column operators: each column of a batched H.v equals the oracle H.v of its own
  operator (1e-12 normwise);
* batch invariance: a request's alpha/beta sequence is bit-identical whether its pool's
  other slots are idle or busy with other requests;
* serving: every result's E0 agrees with the oracle's Lanczos ground energy of the same
  operator to 1e-10 relative, and its residual bound is below 1e-10.
"""
import math
import sys

import numpy as np
import pytest

import oracle
from pned_amd import gpu, lattice, model, serving

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if gpu.device_count() < 1:
        pytest.skip("no GPU")


def test_batch_hv_per_column_operators():
    _say("batch hv")
    cl = lattice.get_cluster("2x4")
    rng = np.random.default_rng(3)
    ops = [model.build_operator(cl, 20, float(rng.uniform(0, 360)), field=float(rng.uniform(0, 0.5)))
           for _ in range(4)]
    ctx = gpu.Context(ops[0], batch_ops=ops)
    d = ctx.dim
    X = rng.standard_normal((d, 4)) + 1j * rng.standard_normal((d, 4))
    Y = rng.standard_normal((d, 4)) + 1j * rng.standard_normal((d, 4))
    v1, v0 = ctx.vector(4), ctx.vector(4)
    v1.upload(X)
    v0.upload(Y)
    pr = ctx.p_mltply_v3(v1, v0, 0.03)
    got = v0.download()
    for c in range(4):
        want = Y[:, c].copy()
        p_ref = oracle.hv(oracle.OracleOp(ops[c]), np.ascontiguousarray(X[:, c]), want, eperbd=0.03)
        assert np.linalg.norm(got[:, c] - want) <= 1e-12 * np.linalg.norm(want), c
        assert abs(pr[c] - p_ref) <= 1e-13 * np.linalg.norm(X[:, c]) * np.linalg.norm(want - Y[:, c])
    # replace one column's operator
    new = model.build_operator(cl, 20, 201.0, field=0.37)
    ctx.set_column_matrices(2, new)
    ctx.hv_fresh(v1, v0)
    want = np.zeros(d, complex)
    oracle.hv(oracle.OracleOp(new), np.ascontiguousarray(X[:, 2]), want)
    assert np.linalg.norm(v0.download()[:, 2] - want) <= 1e-12 * np.linalg.norm(want)


@pytest.mark.parametrize("slots", [1, 4])
def test_batch_steps_equal_single_steps(slots):
    """pned_lanczos_batch_steps gives bit for bit the alpha/beta of the same number of
    single steps (batch context with an idle column, and one column of a plain context)."""
    _say(f"batch steps vs single steps, {slots} slots")
    reqs = serving.make_requests(slots, ["2x4"], seed=9)
    runs = []
    for chunked in (False, True):
        pool = serving.Pool("2x4", slots, 1e-10, 300)
        for c, r in enumerate(reqs):
            if slots == 1 or c != 2:
                pool.admit(c, r, 0.0)
        mode, ap, bp = pool.mode.copy(), pool.aprev.copy(), pool.bprev.copy()
        if chunked:
            A, B = pool.ctx.lanczos_batch_steps(pool.v1, pool.v0, mode, ap, bp, 7)
        else:
            rows_a, rows_b = [], []
            for _ in range(7):
                a, b = pool.ctx.lanczos_batch_step(pool.v1, pool.v0, mode, ap, bp)
                rows_a.append(a)
                rows_b.append(b)
                mode = np.where(mode != 0, 2, 0).astype(np.int32)
                ap, bp = a, b
            A, B = np.array(rows_a), np.array(rows_b)
        runs.append((A, B))
        pool.close()
    (A1, B1), (A2, B2) = runs
    assert A1.shape == A2.shape == (7, slots)
    assert np.array_equal(A1, A2) and np.array_equal(B1, B2)
    if slots > 1:
        assert not A2[:, 2].any() and not B2[:, 2].any()
    assert (B2[:, 0] > 0).all()


def test_alone_equals_in_the_mix():
    _say("batch invariance")
    reqs = serving.make_requests(8, ["3x3"], seed=5)
    target = reqs[3]

    def run(batch):
        pool = serving.Pool("3x3", 8, 1e-10, 300)
        for c, r in enumerate(batch):
            if r is not None:
                pool.admit(c, r, 0.0)
        got = None
        while pool.active():
            for res in pool.step(lambda: 0.0):
                if res.rid == target.rid:
                    got = res
        pool.close()
        return got

    alone = run([None, None, None, target, None, None, None, None])
    mixed = run(reqs)
    assert alone is not None and mixed is not None
    assert alone.steps == mixed.steps
    assert np.array_equal(alone.alpha, mixed.alpha)
    assert np.array_equal(alone.beta, mixed.beta)
    assert alone.e0 == mixed.e0


def test_server_results_match_oracle():
    _say("server")
    reqs = serving.make_requests(10, ["2x4", "3x3", "2x5"], seed=11)
    srv = serving.Server(hbm_budget_bytes=50e9, tol=1e-10)
    res = srv.run(reqs)
    assert len(res) == len(reqs)
    for r, q in zip(res, reqs):
        assert r.rid == q.rid
        assert r.resid < 1e-10
        op = serving.operator_for(q)
        x = oracle.gaussian(op.dim, q.seed)
        itr, a, b, ene = oracle.lanczos(oracle.OracleOp(op), x, max_itr=1500, nvec=1)
        assert abs(r.e0 - ene[0]) <= 1e-10 * abs(ene[0]), (q, r.e0, ene[0])


def test_sequential_server_matches_oracle():
    """The one-slot (plain context, matrices replaced per request) serving path."""
    _say("sequential server")
    reqs = serving.make_requests(6, ["2x4", "3x3"], seed=13)
    srv = serving.Server(hbm_budget_bytes=50e9, tol=1e-10, sequential=True)
    res = srv.run(reqs)
    assert len(res) == len(reqs)
    for r, q in zip(sorted(res, key=lambda z: z.rid), sorted(reqs, key=lambda z: z.rid)):
        assert r.rid == q.rid and r.resid < 1e-10
        op = serving.operator_for(q)
        x = oracle.gaussian(op.dim, q.seed)
        itr, a, b, ene = oracle.lanczos(oracle.OracleOp(op), x, max_itr=1500, nvec=1)
        assert abs(r.e0 - ene[0]) <= 1e-10 * abs(ene[0]), (q, r.e0, ene[0])


def test_server_large_sizes_match_plain_lanczos():
    """The sizes that carry the serving work (3x4 N=24 and 2x7 N=28, one slot per pool) in a
    mix with small ones: every served ground state has residual < 1e-10, and its E0 equals
    the plain Lanczos (lnc1, pinned to the oracle at N=16 and N=24) from the same start
    vector to 1e-10."""
    _say("server with N=24 and N=28 requests")
    reqs = serving.make_requests(3, ["3x4"], seed=17) + serving.make_requests(1, ["2x7"], seed=19) + \
        serving.make_requests(4, ["2x4", "3x3"], seed=23)
    for i, q in enumerate(reqs):
        q.rid = i
    srv = serving.Server(hbm_budget_bytes=120e9, tol=1e-10)
    res = srv.run(reqs)
    assert len(res) == len(reqs)
    got = {r.rid: r for r in res}
    for q in reqs:
        r = got[q.rid]
        assert r.resid < 1e-10, (q, r.resid)
        op = serving.operator_for(q)
        ctx = gpu.Context(op)
        v1, v0 = ctx.vector(), ctx.vector()
        ctx.set_initial_vec(v1, q.seed)
        itr, a, b, ene = ctx.lnc1(v1, v0, max_itr=1500, nvec=1)
        _say(f"{q.spec}: served E0 {r.e0:.13f}, plain lnc1 {ene[0]:.13f}")
        assert abs(r.e0 - ene[0]) <= 1e-10 * abs(ene[0]), (q, r.e0, ene[0])
        for v in (v1, v0):
            v.free()
        ctx.close()
