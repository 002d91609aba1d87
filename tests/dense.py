"""Independent dense Hamiltonians (numpy Kronecker products) for small clusters.

Built from textbook spin operators and the couplings, NOT from the bond matrices the
oracle and the product consume, so a mistake in either (dropped term, wrong sign,
transposed operand, wrong site order) shows up as a mismatch.
"""
import math

import numpy as np
import scipy.sparse as sp_


def site_ops(q):
    """Standard (Sx, Sy, Sz) in the basis order the reference uses for that tree.

    q=2 (S=1/2 tree): index 0 = up (Sz=+1/2).  q>2 (spin-S tree): digit i has m = i - S.
    """
    if q == 2:
        sp = np.array([[0, 1], [0, 0]], dtype=complex)   # |0>=up <- |1>=down
        sz = np.diag([0.5, -0.5]).astype(complex)
    else:
        S = 0.5 * (q - 1)
        sp = np.zeros((q, q), dtype=complex)
        for i in range(q - 1):
            m = i - S
            sp[i + 1, i] = math.sqrt((S - m) * (S + m + 1))
        sz = np.diag([i - S for i in range(q)]).astype(complex)
    sm = sp.conj().T
    return 0.5 * (sp + sm), -0.5j * (sp - sm), sz


def embed(op, site, n, q):
    """Operator on ``site`` (0-based; basis index = sum d_k q^k) of an n-site system."""
    out = sp_.csr_matrix(np.array([[1.0 + 0j]]))
    for k in range(n - 1, -1, -1):
        out = sp_.kron(out, sp_.csr_matrix(op) if k == site else sp_.identity(q, format="csr"),
                       format="csr")
    return out


def kh_dense(cluster, J, Kx, Ky, Kz, q=2, h=0.0, hdir=(1, 1, 1), Jzz=None):
    """H = sum_b [J S.S + K_gamma S^g S^g] + h sum_k S_k.n  on ``cluster`` (1-based bonds).

    The reference's MODEL 20 writes the bond as diag(Jxx, Jyy, Jzz) with the Kitaev
    axis coupling K1 = J + K: pass Kx/Ky/Kz as the extra coupling added on that axis.
    """
    n = cluster.n
    sx, sy, sz = site_ops(q)
    S = [sx, sy, sz]
    ops = [[embed(a, k, n, q) for a in S] for k in range(n)]
    D = q ** n
    H = sp_.csr_matrix((D, D), dtype=complex)
    for (a, b, t) in cluster.bonds:
        i, j = a - 1, b - 1
        K = (Kx, Ky, Kz)[t - 1]
        for g in range(3):
            c = J + (K if g == t - 1 else 0.0)
            H += c * ops[i][g] @ ops[j][g]
    if h:
        nvec = np.asarray(hdir, dtype=float)
        for k in range(n):
            H += h * sum(nvec[g] * ops[k][g] for g in range(3))
    return H.toarray() if D <= 4096 else H.tocsr()


def model20_couplings(phi_deg):
    a = math.pi / 180.0 * phi_deg
    J = math.cos(a)
    K1 = math.cos(a) + 2 * math.sin(a)
    return J, K1 - J


# -- product states: closed forms of <v|H|v> and <v|H H|v> --------------------------------

def kh_terms(cluster, q, phi, h):
    """H = sum_T c_T prod_{k in T} A_T,k as (c, {site: matrix}) from textbook operators."""
    J, K = model20_couplings(phi)
    S = site_ops(q)
    out = []
    for (a, b, t) in cluster.bonds:
        for g in range(3):
            c = J + (K if g == t - 1 else 0.0)
            if c != 0.0:
                out.append((c, {a - 1: S[g], b - 1: S[g]}))
    if h:
        ns = S[0] + S[1] + S[2]
        for k in range(cluster.n):
            out.append((h, {k: ns}))
    return out


def product_closed_forms(terms, phis):
    """<v|H|v> and <v|H H|v> for the normalised product state of the single-site phis."""
    def ev(k, m):
        return np.vdot(phis[k], m @ phis[k])

    e1 = sum(c * np.prod([ev(k, m) for k, m in A.items()]) for c, A in terms)
    e2 = 0.0
    for c, A in terms:
        for c2, B in terms:
            val = 1.0 + 0j
            for k in set(A) | set(B):
                m = A.get(k)
                m2 = B.get(k)
                if m is None:
                    val *= ev(k, m2)
                elif m2 is None:
                    val *= ev(k, m)
                else:
                    val *= ev(k, m @ m2)
            e2 += c * c2 * val
    return e1, e2


def upload_product(vec, phis, q, low):
    """Writes prod_k phi_k[d_k] at index sum_k d_k q^k in chunks of q^low states."""
    def kron_sites(ks):
        out = np.ones(1, dtype=complex)
        for k in ks:
            out = np.outer(phis[k], out).ravel()   # site k is the slower index
        return out

    n = len(phis)
    B = kron_sites(range(low))
    A = kron_sites(range(low, n))
    buf = np.empty_like(B)
    for hi in range(A.size):
        np.multiply(B, A[hi], out=buf)
        vec.upload(buf, offset=hi * B.size)


def random_product_sites(n, q, seed):
    r = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        z = r.standard_normal(q) + 1j * r.standard_normal(q)
        out.append(z / np.linalg.norm(z))
    return out
