"""The precision form on the GPU: complex64 vector storage with FP64 alpha/beta, dots and
norms, and Gram-Schmidt reorthogonalisation, for the ground state and the DSF continued
fraction.  Errors are stated against the FP64 oracle (the reference's arithmetic on the same
start vectors): complex64 rounds every stored component and every H.v product to single
precision, which bounds the energy to ~1e-7 relative and the spectrum to ~1e-6."""
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


def _basis(ctx, x, m, dtype):
    vecs = [ctx.vector(1, dtype) for _ in range(m)]
    vecs[0].upload(x)
    return vecs, ctx.vector(1, dtype)


@pytest.mark.parametrize("dtype,e_tol,orth_tol", [(gpu.C128, 1e-12, 1e-12), (gpu.C64, 1e-6, 1e-5)])
def test_reorthogonalised_lanczos_ground_state(dtype, e_tol, orth_tol):
    _say("reorth lanczos", dtype)
    cl = lattice.reference_cluster(16)
    op = model.build_operator(cl, 20, 126.0, field=0.1)
    ctx = gpu.Context(op)
    x = oracle.gaussian(ctx.dim, 20261202)
    _, _, _, e_ref = oracle.lanczos(oracle.OracleOp(op), x, max_itr=1024, nvec=1)
    basis, w = _basis(ctx, x, 160, dtype)
    itr, a, b, ene = ctx.lnc1_reorth(basis, w, nvec=1)
    err = abs(ene[0] - e_ref[0]) / abs(e_ref[0])
    Q = np.array([v.download().astype(np.complex128) for v in basis[: itr + 1]])
    G = Q.conj() @ Q.T
    orth = np.max(np.abs(G - np.eye(len(Q))))
    print(f"[result] dtype {dtype}: E0 {ene[0]:.15f} vs FP64 oracle {e_ref[0]:.15f}, rel err {err:.2e}, "
          f"{itr} steps, max |<q_i|q_j> - delta_ij| {orth:.2e}", flush=True)
    assert err <= e_tol
    assert orth <= orth_tol


def _reorth_lanczos_fp64(oop, q1, steps):
    """The reorthogonalised recursion in FP64 on the CPU (the oracle's H.v, numpy Gram-Schmidt,
    two classical passes as the library does): alpha_i, beta_i."""
    d = len(q1)
    Q = [q1 / np.linalg.norm(q1)]
    a, b = [], []
    for i in range(steps):
        w = np.zeros(d, dtype=complex)
        oracle.hv(oop, Q[-1], w)
        a.append(np.vdot(Q[-1], w).real)
        w = w - a[-1] * Q[-1] - (b[-1] * Q[-2] if i > 0 else 0)
        for _ in range(2):
            c = [np.vdot(q, w) for q in Q]
            w = w - sum(cj * q for cj, q in zip(c, Q))
        b.append(np.linalg.norm(w))
        Q.append(w / b[-1])
    return np.array(a), np.array(b), np.array(Q)


@pytest.mark.parametrize("dtype,a_tol,d_tol", [(gpu.C128, 1e-12, 1e-9), (gpu.C64, 2e-6, 1e-4)])
def test_reorthogonalised_dsf(dtype, a_tol, d_tol):
    """S^z(q) continued fraction of the ground state (N=12) with the reorthogonalised
    recursion, against the same recursion in FP64 on the CPU.  The individual coefficients
    of this excitation are ill-conditioned functions of the start vector (a 1e-15 change
    of it moves alpha_29 by ~1e-6 in FP64), so the comparison is made on what they define:
    the spectrum S(q, w) of a converged fraction (100 steps; at 30 steps a complex64
    rounding of the start alone changes it by 1e-5, at 60 and more by 4e-9), and on the
    first 5 coefficients.  Stated beside it: the difference to the reference's
    unreorthogonalised FP64 recursion with the same number of steps."""
    _say("reorth dsf", dtype)
    cl = lattice.reference_cluster(12)
    op = model.build_operator(cl, 20, 54.0)
    oop = oracle.OracleOp(op)
    d = op.dim
    g = oracle.gaussian(d, 7)
    _, _, _, e_ref = oracle.lanczos(oop, g, max_itr=1024, nvec=1)
    _, x0 = oracle.inverse_iteration(oop, e_ref[0], g)
    q = cl.q_list[3]
    vx = oracle.sz_q(cl.n, 2, cl.site_vec, q, x0)
    cn = np.vdot(vx, vx).real / cl.n
    exc = vx / np.linalg.norm(vx)
    s_maxitr, i_range, split, eta = 100, 2000, 2.5e-3, 0.05
    ra, rb, Q = _reorth_lanczos_fp64(oop, exc, s_maxitr - 1)
    assert np.max(np.abs(Q.conj() @ Q.T - np.eye(len(Q)))) < 1e-12
    sa_r = np.zeros(s_maxitr)
    sb_r = np.zeros(s_maxitr)
    sa_r[: s_maxitr - 1], sb_r[0], sb_r[1:] = ra, cn, rb
    _, d_r, _ = oracle.renbun(sa_r, sb_r, e_ref[0], eta, i_range, split, s_maxitr)
    d_r = oracle.normalize_dsf(d_r, split, cn)
    sa_ref, sb_ref = oracle.dsf_lanczos(oop, exc, cn, s_maxitr)
    _, d_ref, _ = oracle.renbun(sa_ref, sb_ref, e_ref[0], eta, i_range, split, s_maxitr)
    d_ref = oracle.normalize_dsf(d_ref, split, cn)
    ctx = gpu.Context(op)
    basis, w = _basis(ctx, exc, s_maxitr, dtype)
    sa, sb = ctx.dsflnc1z_reorth(basis, w, cn, s_maxitr)
    _, dsf, _ = ctx.renbun(sa, sb, e_ref[0], eta, i_range, split, s_maxitr, sq=cn)
    ea = np.max(np.abs(sa[:5] - ra[:5]))
    # floor for the first coefficients: the same FP64 recursion started from the excitation
    # rounded to the storage type (each later Lanczos vector is rounded the same way)
    ex_s = exc.astype(np.complex64 if dtype == gpu.C64 else np.complex128).astype(np.complex128)
    ra_s, _, _ = _reorth_lanczos_fp64(oop, ex_s / np.linalg.norm(ex_s), 5)
    floor = np.max(np.abs(ra_s[:5] - ra[:5]))
    ed = np.max(np.abs(dsf - d_r)) / np.max(np.abs(d_r))
    eref = np.max(np.abs(dsf - d_ref)) / np.max(np.abs(d_ref))
    print(f"[result] dtype {dtype}, reorthogonalised DSF ({s_maxitr} steps): vs the FP64 reorthogonalised "
          f"recursion max |d salpha_1..5| {ea:.2e} (start rounded to the storage type alone: {floor:.2e}), max |d S(q,w)|/max S {ed:.2e}; vs the reference's FP64 "
          f"recursion without reorthogonalisation max |d S(q,w)|/max S {eref:.2e}", flush=True)
    assert ea <= max(a_tol, 20 * floor)
    assert ed <= d_tol
