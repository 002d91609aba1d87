"""Pins of the CPU oracle (oracle/pned_oracle.c) before the GPU path is compared with it.

What pins each oracle function (the reference ships no tests or golden vectors,
so every pin is a fixture extracted from its text or a mathematical fact):

* lattice tables      -- tests/golden/ref_bond_tables.json (the reference's literal
                         tables, extracted as text) + honeycomb invariants;
* orc_hv (H_dot_v)    -- dense Kronecker-product H built from textbook spin operators
                         (tests/dense.py), S=1/2 with/without [111] field and spin-1;
                         Hermiticity; the eperbd shift identity;
* orc_philox          -- Random123 published known-answer vectors;
* gaussian start      -- moments of 2^20 samples; u1 in (0,1] never gives inf;
* orc_bisec/vec12     -- scipy.linalg.eigh_tridiagonal;
* orc_lanczos         -- dense eigvalsh ground energy; alpha/beta reproduce the
                         tridiagonal of the Krylov basis built by numpy;
* orc_inverse_iteration -- residual ||Hx - E0 x|| and overlap with the dense GS;
* orc_tps             -- (L - H)^k psi by dense matrix powers; energies per step;
                         every eval_phys_val column (internal energy, z-bond
                         correlation with the reference's bit offset, static structure
                         factors) from dense operators of the textbook S^z;
* orc_sz_q            -- sum_j e^{iq.r_j} S^z_j of the textbook S^z (S=1/2 and spin 1);
* DSF CF              -- backward evaluation of the same fraction; m=1 Lorentzian
                         closed form; normalisation integral equals S(q).
"""
import json
import math
import os

import numpy as np
import scipy.sparse as sp
import pytest
import scipy.linalg as sla

import oracle
from pned_amd import lattice, model
from tests import dense

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_reference_tables_match_fixture():
    d = json.load(open(os.path.join(GOLD, "ref_bond_tables.json")))
    for n, v in d.items():
        c = lattice.reference_cluster(int(n))
        assert list(c.ipair) == v["ipair"]
        assert list(c.ibond_type) == v["ibond_type"]


@pytest.mark.parametrize("spec", [8, 12, 16, 18, 24, 32, "2x4", "3x3", "2x5", "3x4", "2x7"])
def test_honeycomb_invariants(spec):
    c = lattice.get_cluster(spec)
    assert c.ibond == 3 * c.n // 2
    per_site = {k: [] for k in range(1, c.n + 1)}
    seen = set()
    for a, b, t in c.bonds:
        assert a != b
        key = (min(a, b), max(a, b))
        assert key not in seen
        seen.add(key)
        per_site[a].append(t)
        per_site[b].append(t)
    for k, ts in per_site.items():
        assert sorted(ts) == [1, 2, 3], (k, ts)   # one x, one y, one z bond per site


def test_torus44_is_reference_32():
    c = lattice.torus(4, 4)
    assert c.n == 32 and c.name == "ref32"
    # x bonds join 2c-1,2c; the z bond of cell (i, j=0) wraps to j=3 (distance 25)
    assert c.bonds[0] == (1, 2, 1) and c.bonds[2] == (1, 2 * (3 * 4 + 0) + 2, 3)


def _oop(cluster, phi, q=2, h=0.0, tps=False):
    op = model.build_operator(cluster, 20, phi, q=q, field=h, tps=tps)
    return op, oracle.OracleOp(op)


@pytest.mark.parametrize("spec,phi,h", [(8, 54.0, 0.0), (12, 126.0, 0.0), (12, 126.0, 0.1),
                                         ("2x3", 77.0, 0.3)])
def test_hv_vs_dense_s12(spec, phi, h):
    c = lattice.get_cluster(spec)
    op, oop = _oop(c, phi, h=h)
    J, K = dense.model20_couplings(phi)
    H = dense.kh_dense(c, J, K, K, K, q=2, h=h)
    rng = np.random.default_rng(1)
    v = rng.standard_normal(2 ** c.n) + 1j * rng.standard_normal(2 ** c.n)
    v0 = rng.standard_normal(2 ** c.n) + 1j * rng.standard_normal(2 ** c.n)
    out = v0.copy()
    pr = oracle.hv(oop, v, out)
    ref = v0 + H @ v
    assert np.linalg.norm(out - ref) <= 1e-13 * np.linalg.norm(ref)
    assert abs(pr - np.vdot(v, H @ v)) <= 1e-12 * abs(np.vdot(v, H @ v))


def test_hv_vs_dense_spin1():
    c = lattice.reference_cluster(8)
    op, oop = _oop(c, 108.0, q=3)
    J, K = dense.model20_couplings(108.0)
    H = dense.kh_dense(c, J, K, K, K, q=3)
    rng = np.random.default_rng(2)
    v = rng.standard_normal(3 ** 8) + 1j * rng.standard_normal(3 ** 8)
    out = np.zeros_like(v)
    oracle.hv(oop, v, out)
    assert np.linalg.norm(out - H @ v) <= 1e-13 * np.linalg.norm(H @ v)


def test_hv_spin1_field_and_shift():
    c = lattice.reference_cluster(8)
    op, oop = _oop(c, 30.0, q=3, h=0.2)
    J, K = dense.model20_couplings(30.0)
    H = dense.kh_dense(c, J, K, K, K, q=3, h=0.2)
    rng = np.random.default_rng(3)
    v = rng.standard_normal(3 ** 8) + 1j * rng.standard_normal(3 ** 8)
    out = np.zeros_like(v)
    ep = 0.37
    oracle.hv(oop, v, out, eperbd=ep)
    # shift identity: eperbd is subtracted on every bond diagonal -> -ibond*eperbd
    ref = H @ v - c.ibond * ep * v
    assert np.linalg.norm(out - ref) <= 1e-13 * np.linalg.norm(ref)


def test_hermiticity_and_linearity_n16():
    c = lattice.reference_cluster(16)
    op, oop = _oop(c, 200.0, h=0.15)
    rng = np.random.default_rng(4)
    d = 2 ** 16
    x = rng.standard_normal(d) + 1j * rng.standard_normal(d)
    y = rng.standard_normal(d) + 1j * rng.standard_normal(d)
    hx, hy = np.zeros(d, complex), np.zeros(d, complex)
    oracle.hv(oop, x, hx)
    oracle.hv(oop, y, hy)
    a = np.vdot(y, hx)
    b = np.conj(np.vdot(x, hy))
    assert abs(a - b) <= 1e-12 * abs(a)
    hz = np.zeros(d, complex)
    oracle.hv(oop, 2.0 * x - 3j * y, hz)
    assert np.linalg.norm(hz - (2.0 * hx - 3j * hy)) <= 1e-12 * np.linalg.norm(hz)


def test_kitaev_flux_commutes_n16():
    """Pure Kitaev (MODEL 20, phi=90 -> J=0): plaquette W_p = prod sigma^{gamma_out} commutes."""
    c = lattice.reference_cluster(16)
    op, oop = _oop(c, 90.0)
    # find one hexagon: a 6-cycle in the bond graph
    adj = {k: [] for k in range(c.n)}
    for a, b, t in c.bonds:
        adj[a - 1].append((b - 1, t))
        adj[b - 1].append((a - 1, t))
    hexagon = None

    def dfs(path):
        nonlocal hexagon
        if hexagon:
            return
        if len(path) == 6:
            if any(nb == path[0] for nb, _ in adj[path[-1]]):
                hexagon = list(path)
            return
        for nb, _ in adj[path[-1]]:
            if nb not in path:
                dfs(path + [nb])
    dfs([0])
    assert hexagon
    # outgoing bond type of each hexagon site = the one not used inside the loop
    d = 2 ** c.n
    idx = np.arange(d)
    rng = np.random.default_rng(5)
    v = rng.standard_normal(d) + 1j * rng.standard_normal(d)

    def W(x):
        y = x.copy()
        for k, s in enumerate(hexagon):
            inner = {t for nb, t in adj[s] if nb in (hexagon[k - 1], hexagon[(k + 1) % 6])}
            g = ({1, 2, 3} - inner).pop()
            bit = (idx >> s) & 1
            if g == 3:     # sigma^z: +1 on up (bit 0)
                y = y * (1 - 2 * bit)
            elif g == 1:   # sigma^x: flip
                y = y[idx ^ (1 << s)]
            else:          # sigma^y |b> : flip with i(1-2b') phase
                y = y[idx ^ (1 << s)] * (1j * (1 - 2 * bit)) * (-1)
        return y
    hv1 = np.zeros(d, complex)
    oracle.hv(oop, W(v), hv1)
    hv2 = np.zeros(d, complex)
    oracle.hv(oop, v, hv2)
    assert np.linalg.norm(hv1 - W(hv2)) <= 1e-12 * np.linalg.norm(hv1)


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10
    assert list(oracle.philox([0, 0, 0, 0], [0, 0])) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert list(oracle.philox([0xffffffff] * 4, [0xffffffff] * 2)) == \
        [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert list(oracle.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                              [0xa4093822, 0x299f31d0])) == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_gaussian_moments():
    v = oracle.gaussian(1 << 20, 20261202, normalize=False)
    assert np.all(np.isfinite(v))
    for part in (v.real, v.imag):
        assert abs(part.mean()) < 5e-3
        assert abs(part.var() - 1.0) < 5e-3
    assert abs(np.mean(v.real * v.imag)) < 5e-3
    w = oracle.gaussian(1 << 12, 7)
    assert abs(np.linalg.norm(w) - 1.0) < 1e-14
    assert not np.allclose(w, oracle.gaussian(1 << 12, 8))


def test_bisec_vec12_vs_scipy():
    rng = np.random.default_rng(6)
    m = 60
    a = rng.standard_normal(m) * 3
    b = rng.random(m) + 0.1
    E = oracle.bisec(a, b, 8)
    ev, evec = sla.eigh_tridiagonal(a, b[:-1])
    assert np.allclose(E, ev[:8], atol=1e-12)
    V = oracle.vec12(a, b, E, 3)
    for k in range(3):
        assert abs(abs(V[k] @ evec[:, k]) - 1.0) < 1e-10


def test_lanczos_ground_energy_n12():
    c = lattice.reference_cluster(12)
    op, oop = _oop(c, 126.0, h=0.1)
    J, K = dense.model20_couplings(126.0)
    H = dense.kh_dense(c, J, K, K, K, h=0.1)
    e_exact = np.linalg.eigvalsh(H)
    v = oracle.gaussian(2 ** 12, 20261202)
    itr, alpha, beta, ene = oracle.lanczos(oop, v, max_itr=1024, nvec=8)
    assert itr > 0
    assert abs(ene[0] - e_exact[0]) < 1e-10 * abs(e_exact[0])
    # first steps: alpha/beta are the tridiagonal of the Krylov basis (numpy Gram-Schmidt)
    q0 = v.copy()
    w = H @ q0
    a0 = np.vdot(q0, w).real
    r = w - a0 * q0
    assert abs(alpha[0] - a0) < 1e-12 * abs(a0)
    assert abs(beta[0] - np.linalg.norm(r)) < 1e-12 * np.linalg.norm(r)
    q1 = r / np.linalg.norm(r)
    a1 = np.vdot(q1, H @ q1).real
    assert abs(alpha[1] - a1) < 1e-12 * abs(a1)


def test_inverse_iteration_n12():
    c = lattice.reference_cluster(12)
    op, oop = _oop(c, 54.0)
    J, K = dense.model20_couplings(54.0)
    H = dense.kh_dense(c, J, K, K, K)
    e, vecs = np.linalg.eigh(H)
    b = oracle.gaussian(2 ** 12, 11)
    it, x = oracle.inverse_iteration(oop, e[0], b, max_inv=50)
    assert it > 0
    assert np.linalg.norm(H @ x - e[0] * x) < 1e-8
    assert abs(abs(np.vdot(vecs[:, 0], x)) - 1.0) < 1e-10


def test_tps_vs_dense_powers_n8():
    c = lattice.reference_cluster(8)
    op_tps = model.build_operator(c, 20, 90.0, tps=True)
    oop = oracle.OracleOp(op_tps)
    J, K = dense.model20_couplings(90.0)
    H = dense.kh_dense(c, J, K, K, K)
    L = op_tps.shift_L
    A = L * np.eye(2 ** 8) - H
    psi = oracle.gaussian(2 ** 8, 3)
    steps = 12
    sv = c.site_vec
    sq = c.sq_momenta()
    alpha, beta, coef = oracle.tps(oop, sv, sq, L, steps, psi)
    v = psi.copy()
    for k in range(steps):
        e = np.vdot(v, H @ v).real
        assert abs(alpha[k] - e) < 1e-11 * max(1.0, abs(e)), k
        w = A @ v
        assert abs(beta[k] - np.linalg.norm(w)) < 1e-11 * np.linalg.norm(w)
        v = w / np.linalg.norm(w)
    assert coef.shape == (steps, 7 + len(sq))
    # coef(:,4) = l_max - alpha' = <H>, coef(:,2) = l_max
    assert np.allclose(coef[:, 3], alpha, atol=1e-11)
    assert np.allclose(coef[:, 1], L)


def _sz_diag(n, q):
    """Diagonals of the textbook S^z_j (tests/dense.py basis order) for j = 0..n-1."""
    _, _, sz = dense.site_ops(q)
    return [dense.embed(sz, j, n, q).diagonal().real for j in range(n)]


@pytest.mark.parametrize("spec,phi,h", [(8, 54.0, 0.1), (12, 126.0, 0.0)])
def test_tps_observables_vs_dense_operators(spec, phi, h):
    """coef(:, 5..7+sq_num) of eval_phys_val (H/TPS.f90:83-111) from dense operators built
    with the textbook S^z (basis: bit 0 = up, so a bit value b gives b - 1/2 = -S^z):
      col 5   Re<vg|H|v1>  (= Re(l<vg|v1> - <vg|(l-H)v1>), calc_internal_energy :114-125);
      col 6/7 <v1|Z|v1>, Re<vg|Z|v1> with Z = sum over type-3 bonds of f(p1) f(p2), where
              p = the 1-based ipair entry used as a bit position (the reference's
              off-by-one, :290-316) and f(p) = -S^z_p for p < n, -1/2 (bit of rank id 0)
              for p = n;
      sq      per momentum q_k: Re<v1|A|v1>, Re<vg|A|v1>, Re<v1|B|v1>, Re<vg|B|v1> with
              A = sum_j e^{iq.r_j} S^z_j S^z_0 (calc_sq_zz) and
              B = sum_j e^{iq.r_j} (1/8 - S^z_j S^z_0 / 2) (= [s_j != s_0]/4, calc_sq_pp);
              of these 4*sq_num values only the first sq_num are stored (:106-109).
    vg and v1 are the normalised (l - H)^{k-2} psi and (l - H)^{k-1} psi (vg = 0 at k = 1)."""
    c = lattice.get_cluster(spec)
    n = c.n
    J, K = dense.model20_couplings(phi)
    H = dense.kh_dense(c, J, K, K, K, h=h)
    op_tps = model.build_operator(c, 20, phi, field=h, tps=True)
    L = op_tps.shift_L
    psi = oracle.gaussian(2 ** n, 17)
    steps = 6
    sq = c.sq_momenta()
    nsq = len(sq)
    _, _, coef = oracle.tps(oracle.OracleOp(op_tps), c.site_vec, sq, L, steps, psi)
    szd = _sz_diag(n, 2)
    f = [-szd[p] if p < n else -0.5 * np.ones(2 ** n) for p in range(n + 1)]
    zdiag = sum(f[a] * f[b] for (a, b, t) in c.bonds if t == 3)
    A = [sum(np.exp(1j * (qk @ c.site_vec[j])) * szd[j] * szd[0] for j in range(n)) for qk in sq]
    B = [sum(np.exp(1j * (qk @ c.site_vec[j])) * (0.125 - 0.5 * szd[j] * szd[0]) for j in range(n)) for qk in sq]
    vg = np.zeros(2 ** n, dtype=complex)
    v1 = psi.copy()
    for k in range(steps):
        row = coef[k]
        assert abs(row[4] - np.vdot(vg, H @ v1).real) < 1e-11, k
        assert abs(row[5] - np.vdot(v1, zdiag * v1).real) < 1e-12, k
        assert abs(row[6] - np.vdot(vg, zdiag * v1).real) < 1e-12, k
        four = []
        for a, b in zip(A, B):
            four += [np.vdot(v1, a * v1).real, np.vdot(vg, a * v1).real,
                     np.vdot(v1, b * v1).real, np.vdot(vg, b * v1).real]
        assert np.allclose(row[7:7 + nsq], four[:nsq], rtol=0, atol=1e-12), k
        w = L * v1 - H @ v1
        vg, v1 = v1, w / np.linalg.norm(w)
    # the off-by-one matters on these clusters: the textbook z-bond sum differs
    zz_true = sum(szd[a - 1] * szd[b - 1] for (a, b, t) in c.bonds if t == 3)
    assert np.abs(zz_true - zdiag).max() > 0.1


def test_tps_spin1_vs_dense_operators():
    """The spin-S mTPQ (S_arbitrary/Hybrid_TPS9_TPS.f90:45-409) at spin 1, N=8, against
    dense operators built from the textbook spin-1 matrices:
      * the shifted operator of change_to_TPS_hamiltonian's spin-S form
        (S_arbitrary/Hybrid_TPS9_params.f90:1660-1686), which adds l_b only to the local
        indices 0..3 of each bond (local index q*d(max site) + d(min site)):
        A = sum_b l_b P_b - H, P_b the projector on those four local states;
        alpha_k = L - <v|A|v>, beta_k = ||A v||;
      * coef columns 5..7+sq_num with the tree's weights on the BITS of the base-q state
        index (not its digits): zz bonds (b_p1 - 1/2)(b_p2 - 1/2) at the 1-based ipair
        entries (0 past bit n-1), structure factors (b_j - S)(b_0 - S) and
        my_xor(b_j, b_0), where my_xor = sx + sy of the tree's transposed ladder matrices =
        (S^x - S^y)[b_j, b_0] of the textbook ones."""
    c = lattice.reference_cluster(8)
    n, q = c.n, 3
    D = q ** n
    phi, h = 108.0, 0.1
    J, K = dense.model20_couplings(phi)
    H = dense.kh_dense(c, J, K, K, K, q=q, h=h)
    op_tps = model.build_operator(c, 20, phi, q=q, field=h, tps=True)
    op = model.build_operator(c, 20, phi, q=q, field=h)
    lb = op_tps.shift_L / c.ibond
    idx = np.arange(D)
    digits = [(idx // q ** k) % q for k in range(n)]
    shift = np.zeros(D)
    for (a, b, _) in c.bonds:
        lo, hi = min(a, b) - 1, max(a, b) - 1
        shift += lb * ((q * digits[hi] + digits[lo]) <= 3)
    A = sp.diags(shift) - H
    psi = oracle.gaussian(D, 23)
    steps = 6
    sq = c.sq_momenta()
    nsq = len(sq)
    alpha, beta, coef = oracle.tps(oracle.OracleOp(op_tps), c.site_vec, sq, op_tps.shift_L, steps, psi)
    bit = [((idx >> p) & 1).astype(float) for p in range(n)]
    f = [bit[p] - 0.5 if p < n else -0.5 * np.ones(D) for p in range(n + 1)]
    zdiag = sum(f[a] * f[b] for (a, b, t) in c.bonds if t == 3)
    Sx, Sy, _ = dense.site_ops(q)
    mx = Sx - Sy
    ph = [np.exp(1j * (qk @ c.site_vec[j])) for qk in sq for j in range(n)]
    Aq = [sum(np.exp(1j * (qk @ c.site_vec[j])) * (bit[j] - 1.0) * (bit[0] - 1.0) for j in range(n)) for qk in sq]
    Bq = [sum(np.exp(1j * (qk @ c.site_vec[j])) * mx[bit[j].astype(int), bit[0].astype(int)] for j in range(n))
          for qk in sq]
    del ph
    vg = np.zeros(D, dtype=complex)
    v1 = psi.copy()
    L = op_tps.shift_L
    for k in range(steps):
        Av = A @ v1
        assert abs(alpha[k] - (L - np.vdot(v1, Av).real)) < 1e-11 * L, k
        assert abs(beta[k] - np.linalg.norm(Av)) < 1e-11 * np.linalg.norm(Av), k
        row = coef[k]
        assert abs(row[4] - (L * np.vdot(vg, v1) - np.vdot(vg, Av)).real) < 1e-11, k
        assert abs(row[5] - np.vdot(v1, zdiag * v1).real) < 1e-12, k
        assert abs(row[6] - np.vdot(vg, zdiag * v1).real) < 1e-12, k
        four = []
        for a, b in zip(Aq, Bq):
            four += [np.vdot(v1, a * v1).real, np.vdot(vg, a * v1).real,
                     np.vdot(v1, b * v1).real, np.vdot(vg, b * v1).real]
        assert np.allclose(row[7:7 + nsq], four[:nsq], rtol=0, atol=1e-12), k
        vg, v1 = v1, Av / np.linalg.norm(Av)
    # the unshifted H.v of the same tables is the textbook H (the shift is the only change)
    x = oracle.gaussian(D, 5)
    y = np.zeros(D, dtype=complex)
    oracle.hv(oracle.OracleOp(op), x, y)
    assert np.linalg.norm(y - H @ x) < 1e-12 * np.linalg.norm(H @ x)

@pytest.mark.parametrize("spec,q", [(8, 2), (12, 2), (8, 3)])
def test_sz_q_is_the_dense_sz_sum(spec, q):
    """orc_sz_q (intalv_zz before normalisation, H/DSF.f90:18-87, S/DSF.f90:18-89) against
    sum_j e^{iq.r_j} S^z_j from the textbook S^z.  Basis conventions (both trees): the
    S=1/2 tree's bit 0 is up, so its weight (b - 1/2) is -S^z_j; the spin-S tree's digit d
    has m = d - S, so its weight (d - S) is +S^z_j."""
    c = lattice.get_cluster(spec)
    n = c.n
    sign = -1.0 if q == 2 else 1.0
    szd = _sz_diag(n, q)
    x = np.random.default_rng(5).standard_normal(q ** n) + 1j * np.random.default_rng(6).standard_normal(q ** n)
    for qv in c.q_list:
        v = oracle.sz_q(n, q, c.site_vec, qv, x)
        ref = sign * sum(np.exp(1j * (qv @ c.site_vec[j])) * szd[j] for j in range(n)) * x
        assert np.linalg.norm(v - ref) <= 1e-13 * np.linalg.norm(ref)


def test_cf_closed_forms():
    # continued fraction: the modified Lentz recursion must equal the backward
    # evaluation of f = t + a1/(b1 + a2/(b2 + ...)), a_m = -sbeta_m^2, b_m = w+E-alpha_m+i eta
    rng = np.random.default_rng(9)
    depth = 6
    salpha = rng.standard_normal(depth) + 1.0
    sbeta = rng.random(depth + 1) + 0.2
    eta, E = 0.05, -1.3
    w, dsf, conv = oracle.renbun(salpha, sbeta, E, eta, 2000, 2.5e-3, depth + 1)
    tiny = 1e-10 + 1e-10j
    tail = np.zeros_like(w, dtype=complex)
    for m in range(depth - 1, -1, -1):
        b = w + E - salpha[m] + 1j * eta
        tail = -sbeta[m] ** 2 / (b + tail)
    ref = (tiny + tail).imag
    # early exit at |C D - 1|^2 < 5e-29 truncates the tail only below that level
    assert np.allclose(dsf, ref, rtol=1e-9, atol=1e-13)
    nd = oracle.normalize_dsf(dsf, 2.5e-3, 0.3)
    integral = np.sum(0.5 * 2.5e-3 * (nd[1:] + nd[:-1]))
    assert abs(integral - 0.3) < 1e-12


@pytest.mark.parametrize("n,k", [(8, 4), (8, 3), (9, 4), (12, 6), (12, 0), (12, 12), (13, 5)])
def test_sector_tables_enumerate_and_index(n, k):
    """mk_list12: list1 is every state with k set bits in increasing order (left half
    major, right half minor is numeric order), and la/lb index each of them back."""
    list1, la, lb = oracle.sector_tables(n, k)
    want = np.array([s for s in range(1 << n) if bin(s).count("1") == k], dtype=np.uint32)
    assert np.array_equal(list1, want)
    ihf = (n + 1) // 2
    idx = la[list1 & ((1 << ihf) - 1)].astype(np.int64) + lb[list1 >> ihf].astype(np.int64) - 1
    assert np.array_equal(idx, np.arange(len(list1)))


@pytest.mark.parametrize("spec,k", [(8, 4), (8, 3), (12, 6), ("2x3", 5)])
def test_sector_hv_is_the_projected_dense_operator(spec, k):
    """The sector H.v equals the dense Heisenberg H restricted to the k-down sector
    (phi = 0: J S.S), with the eperbd shift on every bond diagonal."""
    cl = lattice.get_cluster(spec)
    op = model.build_operator(cl, 20, 0.0)
    J, K = dense.model20_couplings(0.0)
    H = dense.kh_dense(cl, J, K, K, K)
    H = H.toarray() if hasattr(H, "toarray") else np.asarray(H)
    list1, _, _ = oracle.sector_tables(cl.n, k)
    Hs = H[np.ix_(list1.astype(np.int64), list1.astype(np.int64))]
    rng = np.random.default_rng(k)
    x = rng.standard_normal(len(list1)) + 1j * rng.standard_normal(len(list1))
    y0 = rng.standard_normal(len(list1)) + 0j
    y = y0.copy()
    ep = 0.04
    pr = oracle.hv_sector(oracle.OracleOp(op), k, x, y, eperbd=ep)
    want = y0 + Hs @ x - cl.ibond * ep * x
    assert np.linalg.norm(y - want) <= 1e-12 * np.linalg.norm(want)
    assert abs(pr - np.vdot(x, Hs @ x - cl.ibond * ep * x)) <= 1e-12 * np.linalg.norm(x) ** 2 * np.linalg.norm(Hs, 2)
    # the sector of the complement count is the spin-flipped image with the same spectrum
    w_sec = np.linalg.eigvalsh(Hs)
    w_full = np.linalg.eigvalsh(H)
    assert np.all(np.min(np.abs(w_full[None, :] - w_sec[:, None]), axis=1) < 1e-9)


@pytest.mark.parametrize("spec,k", [(8, 3), (8, 4), (8, 0), (12, 5), ("2x3", 6)])
def test_intalv_xy_is_the_dense_bit_setting_sum(spec, k):
    """intalv_xy equals sum_j e^{i q.r_j} |1><0|_j (the reference's ipm = +1: the bit of
    site j goes from 0 to 1) built from dense single-site operators on the full space,
    restricted to the sector with k + 1 set bits and normalised; cnorm = ||.||^2 / N."""
    cl = lattice.get_cluster(spec)
    n = cl.n
    src, _, _ = oracle.sector_tables(n, k)
    dst, _, _ = oracle.sector_tables(n, k + 1)
    rng = np.random.default_rng(7 + k)
    x = rng.standard_normal(len(src)) + 1j * rng.standard_normal(len(src))
    xf = np.zeros(1 << n, complex)
    xf[src.astype(np.int64)] = x
    q = (0.7, -1.3)
    raise_bit = np.array([[0, 0], [1, 0]], dtype=complex)   # |1><0| in the bit basis
    vf = np.zeros(1 << n, complex)
    for j in range(n):
        ph = np.exp(1j * (q[0] * cl.site_vec[j][0] + q[1] * cl.site_vec[j][1]))
        vf += ph * (dense.embed(raise_bit, j, n, 2) @ xf)
    outside = np.ones(1 << n, bool)
    outside[dst.astype(np.int64)] = False
    assert np.all(vf[outside] == 0)
    want = vf[dst.astype(np.int64)]
    v, cn = oracle.intalv_xy(n, cl.site_vec, q, k, x)
    assert abs(cn - np.vdot(want, want).real / n) <= 1e-13 * cn
    assert np.linalg.norm(v - want / np.linalg.norm(want)) <= 1e-13


@pytest.mark.parametrize("spec,q,ipm", [(8, 3, 1), (8, 3, -1), (8, 4, 1), (8, 5, -1), (12, 2, 1), (12, 2, -1)])
def test_intalv_pm_full_space_is_the_dense_ladder_sum(spec, q, ipm):
    """Full-space intalv_xy equals sum_j e^{i q.r_j} S^{+/-}_j from the textbook spin-S
    ladder matrices (digit basis m = d - S); for q = 2 the S=1/2 tree's bit set / clear."""
    cl = lattice.get_cluster(spec) if isinstance(spec, str) else lattice.reference_cluster(spec)
    n = cl.n
    rng = np.random.default_rng(3 + q)
    x = rng.standard_normal(q ** n) + 1j * rng.standard_normal(q ** n)
    if q == 2:
        step = np.array([[0, 0], [1, 0]], dtype=complex) if ipm == 1 else np.array([[0, 1], [0, 0]], dtype=complex)
    else:
        sx, sy, _ = dense.site_ops(q)
        splus = sx + 1j * sy            # raises the digit: sp[d+1, d] = sqrt((S-m)(S+m+1))
        step = splus if ipm == 1 else splus.conj().T
    qv = (0.4, 1.1)
    want = np.zeros_like(x)
    for j in range(n):
        ph = np.exp(1j * (qv[0] * cl.site_vec[j][0] + qv[1] * cl.site_vec[j][1]))
        want += ph * (dense.embed(step, j, n, q) @ x)
    v, cn = oracle.intalv_pm(n, q, cl.site_vec, qv, ipm, x)
    assert abs(cn - np.vdot(want, want).real / n) <= 1e-12 * cn
    assert np.linalg.norm(v - want / np.linalg.norm(want)) <= 1e-12


def test_sector_hv_rejects_non_conserving_operator():
    cl = lattice.reference_cluster(8)
    op = model.build_operator(cl, 20, 54.0)   # Kitaev terms flip two aligned spins
    list1, _, _ = oracle.sector_tables(cl.n, 4)
    x = np.ones(len(list1), complex)
    with pytest.raises(ValueError):
        oracle.hv_sector(oracle.OracleOp(op), 4, x, np.zeros_like(x))


def test_ctpq_three_step_expansion():
    """calc_fac/eval_q written out by hand for three mTPQ steps: the even weights are
    1 and (bt^2/8) beta_2^2, the odd ones bt beta_2 and (bt^2/8) beta_2^2 bt beta_3 / 3;
    beta_1 and the last step's columns do not enter."""
    rng = np.random.default_rng(2)
    beta = rng.random(3) + 0.5
    c4 = rng.standard_normal(3)
    c5 = rng.standard_normal(3)
    for bt in (0.3, 1.7, 6.0):
        w2 = bt * bt / 8.0 * beta[1] ** 2
        u1 = bt * beta[1]
        u2 = w2 * bt * beta[2] / 3.0
        want = (c4[0] + w2 * c4[1] + u1 * c5[0] + u2 * c5[1]) / (1.0 + w2 + u1 + u2)
        assert abs(oracle.ctpq_sample(beta, c4, c5, bt) - want) <= 1e-14 * max(1.0, abs(want))


def test_ctpq_constant_norm_series():
    """With every norm equal to b the even weights sum to cosh x and the odd ones to
    2 sinh x (x = bt b / 2); constant columns come back unchanged."""
    nd1, b = 400, 3.0
    beta = np.full(nd1, b)
    for bt in (0.5, 2.0, 5.0):
        x = bt * b / 2.0
        c4 = np.full(nd1, 1.7)
        assert abs(oracle.ctpq_sample(beta, c4, c4, bt) - 1.7) < 1e-13
        got = oracle.ctpq_sample(beta, c4, np.zeros(nd1), bt)
        want = 1.7 * np.cosh(x) / (np.cosh(x) + 2.0 * np.sinh(x))
        assert abs(got - want) < 1e-12


class _HostVec:
    def __init__(self, d):
        self.a = np.zeros(d, dtype=complex)

    def upload(self, x, offset=0):
        self.a[offset:offset + x.size] = x


@pytest.mark.parametrize("spec,q,phi,h,low", [(12, 2, 126.0, 0.1, 7), ("2x3", 2, 77.0, 0.3, 6),
                                               (8, 3, 108.0, 0.0, 5), (8, 3, 30.0, 0.2, 3)])
def test_product_state_closed_forms(spec, q, phi, h, low):
    """The full-size GPU checks' closed forms (<v|H|v>, <Hv|Hv> of a product state from the
    textbook operator terms) and their chunked product-state writer, against the oracle's
    H.v on small clusters; a field of the wrong sign must move <v|H|v>."""
    cl = lattice.get_cluster(spec)
    op, oop = _oop(cl, phi, q=q, h=h)
    phis = dense.random_product_sites(cl.n, q, 5)
    e1, e2 = dense.product_closed_forms(dense.kh_terms(cl, q, phi, h), phis)
    v = _HostVec(q ** cl.n)
    dense.upload_product(v, phis, q, low)
    assert abs(np.vdot(v.a, v.a) - 1.0) < 1e-14
    out = np.zeros_like(v.a)
    pr = oracle.hv(oop, v.a, out)
    assert abs(pr - e1) <= 1e-13 * np.sqrt(abs(e2))
    assert abs(np.vdot(out, out) - e2) <= 1e-13 * abs(e2)
    if h:
        wrong, _ = dense.product_closed_forms(dense.kh_terms(cl, q, phi, -h), phis)
        assert abs(pr - wrong) > 1e-6
