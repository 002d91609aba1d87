"""Couplings and q^2 x q^2 bond matrices (the reference's ``xlocal_Ham``).

Restates ``set_lattice`` (S_one_half/Hybrid_TPS_params.f90:144-185; spin-S
S_arbitrary/Hybrid_TPS9_params.f90:140-182), ``set_xlocal_Ham``
(Hybrid_TPS_params.f90:1388-1597; Hybrid_TPS9_params.f90:1412-1657) and
``change_to_TPS_hamiltonian`` (Hybrid_TPS_params.f90:1600-1623;
Hybrid_TPS9_params.f90:1660-1686).

Array convention (the C ABI consumes it verbatim): ``xlocal[t, out, in]`` in C order
equals the reference's Fortran ``xlocal_Ham(in, out, t+1)``; the local index of a bond
is ``q*digit(max site) + digit(min site)``.  The H.v kernel applies
``v0(out) += xlocal_Ham(in, out) * v1(in)`` exactly as ``H_dot_v`` does
(Hybrid_TPS_parallel.f90:434-436; Hybrid_TPS9_parallel.f90:349-353).

Which operator that formula applies depends on how the matrices were built:

* ``tree="s12"`` (S_one_half): standard spin matrices with basis bit 0 = up
  (Hybrid_TPS_params.f90:1434-1439), so the kernel applies H^T (= H for every
  reference model, whose bond matrices are real symmetric);
* ``tree="spin"`` (S_arbitrary): transposed ladder matrices with digit 0 = m=-S
  (Hybrid_TPS9_params.f90:1480-1490), so the kernel applies H.

The [111] field the reference lacks (h (Sx+Sy+Sz) on every site) is folded into the
bond matrices as (h/c) (S_i + S_j).n per bond, c being the coordination number.  It is
built in the ``spin`` orientation (matrices transposed) for both trees, so the kernel
applies exactly H + field.  For ``h = 0`` this changes nothing.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

EPS_MASK = 0.5e-28   # Hybrid_TPS_params.f90:1391


@dataclass
class Couplings:
    J1: float = 0.0
    K1: float = -1.0
    J1p: float = 0.0
    K1p: float = -1.0
    J1pp: float = 0.0
    K1pp: float = -1.0
    I11: float = 0.0
    I12: float = 0.0
    I11p: float = 0.0
    I12p: float = 0.0
    I11pp: float = 0.0
    I12pp: float = 0.0
    J2: float = 0.0
    K2: float = 0.0
    I21: float = 0.0
    I22: float = 0.0
    J3: float = 0.0
    K3: float = 0.0
    ratio_l: float = 1.0
    c_count: int = 3


def set_lattice(model: int, param: float, tree: str = "s12") -> Couplings:
    """MODEL/PARAM -> couplings (Hybrid_TPS_params.f90:144-185).

    MODEL 0 : J = 1-alpha, K = 1-3 alpha with alpha = PARAM.
    MODEL 20: J = cos(phi), K1 = cos(phi) + 2 sin(phi), phi = PARAM in degrees.
    other   : module defaults -- pure FM Kitaev K=-1 (S=1/2 tree) or J=K=1
              Heisenberg (spin-S tree, Hybrid_TPS9_params.f90:12-15).
    """
    c = Couplings()
    if tree == "spin":
        c.J1 = c.J1p = c.J1pp = 1.0
        c.K1 = c.K1p = c.K1pp = 1.0
    if model == 0:
        a = float(param)
        c.J1 = c.J1p = c.J1pp = 1.0 - a
        c.K1 = c.K1p = c.K1pp = 1.0 - 3.0 * a
        c.I11 = c.I12 = c.I11p = c.I12p = c.I11pp = c.I12pp = 0.0
        c.J2 = c.K2 = c.I21 = c.I22 = c.J3 = c.K3 = 0.0
        c.ratio_l = 1.0
    elif model == 20:
        a = 4.0 * math.atan(1.0) / 180.0 * float(param)
        c.J1 = math.cos(a)
        c.K1 = math.cos(a) + 2.0 * math.sin(a)
        c.J1p = c.J1pp = c.J1
        c.K1p = c.K1pp = c.K1
        c.I11 = c.I12 = c.I11p = c.I12p = c.I11pp = c.I12pp = 0.0
        c.J2 = c.K2 = c.I21 = c.I22 = c.J3 = c.K3 = 0.0
        c.ratio_l = 1.0
    return c


def spin_matrices(q: int, tree: str) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Single-site (sx, sy, sz) as the chosen tree builds them."""
    if tree == "s12":
        assert q == 2
        sx = np.array([[0, 0.5], [0.5, 0]], dtype=complex)
        sy = np.array([[0, -0.5j], [0.5j, 0]], dtype=complex)
        sz = np.array([[0.5, 0], [0, -0.5]], dtype=complex)
        return sx, sy, sz
    ss = 0.5 * (q - 1)
    su = np.zeros((q, q), dtype=complex)
    sd = np.zeros((q, q), dtype=complex)
    for i in range(q - 1):
        sm = i - ss
        su[i, i + 1] = math.sqrt((ss - sm) * (ss + sm + 1.0))
    for i in range(1, q):
        sm = i - ss
        sd[i, i - 1] = math.sqrt((ss + sm) * (ss - sm + 1.0))
    sz = np.diag([i - ss for i in range(q)]).astype(complex)
    sx = 0.5 * (su + sd)
    sy = -0.5j * (su - sd)
    return sx, sy, sz


def _pair(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """M[q*j1+i1, q*j2+i2] = a[i1,i2] * b[j1,j2]: site1 (min) is the low digit."""
    q = a.shape[0]
    m = np.zeros((q * q, q * q), dtype=complex)
    for j2 in range(q):
        for i2 in range(q):
            for j1 in range(q):
                for i1 in range(q):
                    m[q * j1 + i1, q * j2 + i2] = a[i1, i2] * b[j1, j2]
    return m


def bond_matrices(c: Couplings, q: int = 2, tree: str = "s12",
                  field: float = 0.0, field_dir=(1.0, 1.0, 1.0)) -> np.ndarray:
    """Return ``xlocal`` with shape (9, q*q, q*q) indexed [type, out?...].

    The returned array ``H[t]`` is the reference's ``xlocal_Ham(:,:,t+1)`` with the
    Fortran index order kept: ``H[t][row, col] = xlocal_Ham(row, col, t+1)``.
    The kernel reads it as ``xlocal_Ham(in, out)``.
    """
    sx, sy, sz = spin_matrices(q, tree)
    S = [sx, sy, sz]
    names = "xyz"
    prod = {}
    for i, a in enumerate(S):
        for j, b in enumerate(S):
            prod[names[i] + names[j]] = _pair(a, b)
    rl = c.ratio_l
    if tree == "s12":
        J1z = np.array([[c.J1 * rl, c.I11 * rl, c.I12 * rl],
                        [c.I11 * rl, c.J1 * rl, c.I12 * rl],
                        [c.I12 * rl, c.I12 * rl, c.K1]])
        J1x = np.array([[c.K1p, c.I12pp * rl, c.I12p * rl],
                        [c.I12pp * rl, c.J1pp * rl, c.I11p * rl],
                        [c.I12p * rl, c.I11p * rl, c.J1p * rl]])
        J1y = np.array([[c.J1pp * rl, c.I12pp * rl, c.I11p * rl],
                        [c.I12pp * rl, c.K1p, c.I12p * rl],
                        [c.I11p * rl, c.I12p * rl, c.J1p * rl]])
    else:
        J1z = np.array([[c.J1, c.I11 * rl, c.I12 * rl],
                        [c.I11 * rl, c.J1, c.I12 * rl],
                        [c.I12 * rl, c.I12 * rl, c.K1]])
        J1x = np.array([[c.K1p, c.I12pp * rl, c.I12p * rl],
                        [c.I12pp * rl, c.J1pp, c.I11p * rl],
                        [c.I12p * rl, c.I11p * rl, c.J1p]])
        J1y = np.array([[c.J1pp, c.I12pp * rl, c.I11p * rl],
                        [c.I12pp * rl, c.K1p, c.I12p * rl],
                        [c.I11p * rl, c.I12p * rl, c.J1p]])
    J2x = np.array([[c.K2 * rl, c.I21 * rl, c.I22 * rl],
                    [c.I21 * rl, c.J2 * rl, c.I22 * rl],
                    [c.I22 * rl, c.I22 * rl, c.J2 * rl]])
    J2y = np.array([[c.J2 * rl, c.I21 * rl, c.I22 * rl],
                    [c.I21 * rl, c.K2 * rl, c.I22 * rl],
                    [c.I22 * rl, c.I22 * rl, c.J2 * rl]])
    J2z = np.array([[c.J2 * rl, c.I21 * rl, c.I22 * rl],
                    [c.I21 * rl, c.J2 * rl, c.I22 * rl],
                    [c.I22 * rl, c.I22 * rl, c.K2 * rl]])
    J3z = np.diag([c.J3 * rl] * 3)

    def contract(J):
        m = np.zeros_like(prod["xx"])
        for i in range(3):
            for j in range(3):
                m += J[i, j] * prod[names[i] + names[j]]
        return m

    H = np.zeros((9, q * q, q * q), dtype=complex)
    x1x, x1y, x1z = contract(J1x), contract(J1y), contract(J1z)
    x2x, x2y, x2z = contract(J2x), contract(J2y), contract(J2z)
    x3z = contract(J3z)
    if tree == "spin":
        # second-neighbour x/y types are zeroed unless simple=3/4, which no model sets
        x2x = np.zeros_like(x2x)
        x2y = np.zeros_like(x2y)
    H[0], H[1], H[2] = x1x, x1y, x1z
    H[3], H[4], H[5] = x2x, x2y, x2z
    H[6], H[7], H[8] = x3z, x3z, x3z
    if field != 0.0:
        H[:3] += field_bond_term(q, field, field_dir, c.c_count, tree)
    return H


def field_bond_term(q: int, h: float, direction=(1.0, 1.0, 1.0), c_count: int = 3,
                    tree: str = "s12") -> np.ndarray:
    """(h/c_count) (S_1 + S_2).n in the kernel's (in, out) orientation, i.e. transposed.

    With v0(out) += M(in, out) v1(in) this applies exactly h S.n on every site of a
    cluster whose sites all have ``c_count`` bonds.  Basis: spin-S tree (digit 0 = m=-S).
    For ``tree="s12"`` the S=1/2 tree's basis (bit 0 = up) is used instead.
    """
    if tree == "s12":
        sx, sy, sz = spin_matrices(2, "s12")
    else:
        # standard orientation for the spin-S basis: digit i has m = i - S
        ss = 0.5 * (q - 1)
        sp = np.zeros((q, q), dtype=complex)
        for i in range(q - 1):
            m = i - ss
            sp[i + 1, i] = math.sqrt((ss - m) * (ss + m + 1.0))
        sm = sp.conj().T
        sx = 0.5 * (sp + sm)
        sy = -0.5j * (sp - sm)
        sz = np.diag([i - ss for i in range(q)]).astype(complex)
    n = np.asarray(direction, dtype=float)
    s_n = n[0] * sx + n[1] * sy + n[2] * sz
    eye = np.eye(q, dtype=complex)
    m = _pair(s_n, eye) + _pair(eye, s_n)       # standard (row=out, col=in)
    return (h / c_count) * m.T                   # stored as (in, out)


def local_mask(H: np.ndarray, q: int, tree: str = "s12") -> Tuple[np.ndarray, int]:
    """``llocal_Ham`` (True = skipped element) and ``parity_type`` as the reference sets them.

    S=1/2: Hybrid_TPS_params.f90:1567-1594.  spin-S: Hybrid_TPS9_params.f90:1604-1643.
    Returned mask has the same [t, row, col] layout as ``H``.
    """
    qq = q * q
    mask = np.ones_like(H, dtype=bool)
    if tree == "s12":
        parity = 0
        for t in range(9):
            if abs(np.conj(H[t, 1, 0]) * H[t, 1, 0]) > EPS_MASK:
                parity |= 2
            if abs(np.conj(H[t, 2, 0]) * H[t, 2, 0]) > EPS_MASK:
                parity |= 2
            if abs(np.conj(H[t, 3, 0]) * H[t, 3, 0]) > EPS_MASK:
                parity |= 1
        if parity == 0:
            for t in range(9):
                for (k, i) in ((0, 0), (1, 1), (1, 2), (2, 1), (2, 2), (3, 3)):
                    mask[t, k, i] = False
        else:
            for t in range(9):
                for i in range(4):
                    for k in range(4):
                        if abs(np.conj(H[t, k, i]) * H[t, k, i]) > EPS_MASK or k == i:
                            mask[t, k, i] = False
        return mask, parity
    parity = 0
    for t in range(9):
        for j in range(qq):
            for i in range(j, qq):
                s1, s2 = j % q, j // q
                s3, s4 = i % q, i // q
                if s1 + s2 != s3 + s4 and abs(np.conj(H[t, i, j]) * H[t, i, j]) > EPS_MASK:
                    parity |= 1
    for t in range(9):
        for j in range(qq):
            for i in range(qq):
                if parity == 0:
                    s1, s2 = j % q, j // q
                    s3, s4 = i % q, i // q
                    if s1 + s2 == s3 + s4 or i == j:
                        mask[t, i, j] = False
                else:
                    if abs(np.conj(H[t, i, j]) * H[t, i, j]) > EPS_MASK or i == j:
                        mask[t, i, j] = False
    return mask, parity


def tps_shift(H: np.ndarray, q: int, ibond: int, c_count: int = 3,
              tree: str = "s12", lb: Optional[float] = None) -> Tuple[np.ndarray, float, float]:
    """``change_to_TPS_hamiltonian``: h_b <- l_b I - h_b.  Returns (H', l_b, L = ibond*l_b).

    S=1/2 (Hybrid_TPS_params.f90:1600-1623): l_b = sum_t sum_k |h_kk|^2 / c_count.
    spin-S (Hybrid_TPS9_params.f90:1660-1686): l_b = 2 sum_t sum_{k<4} S^2 |h_kk| / c_count / q^2,
    and only local indices 0..3 of each diagonal are shifted (reference quirk, kept).
    ``lb`` given: that per-bond shift instead of the reference's formula.
    """
    Hs = -H.copy()
    if lb is not None:
        pass
    elif tree == "s12":
        lb = 0.0
        for t in range(9):
            for i in range(4):
                lb += abs((np.conj(H[t, i, i]) * H[t, i, i]).real)
        lb /= c_count
    else:
        ss = 0.25 * (q - 1) * (q - 1)
        lb = 0.0
        for t in range(9):
            for i in range(4):
                lb += ss * math.sqrt(abs((np.conj(H[t, i, i]) * H[t, i, i]).real))
        lb = 2.0 * lb / c_count / (q * q)
    for t in range(9):
        for i in range(4):
            Hs[t, i, i] = lb + Hs[t, i, i]
    return Hs, lb, lb * ibond


@dataclass
class Operator:
    """Everything the C ABI needs to build a context (the reference's data_mem tables)."""

    n: int
    q: int
    ipair: np.ndarray        # int32 (2*ibond,), 1-based
    ibond_type: np.ndarray   # int32 (ibond,), 1..9
    xlocal: np.ndarray       # complex128 (9, q^2, q^2): xlocal[t, row, col] = xlocal_Ham(row,col,t+1)
    llocal: np.ndarray       # bool, same shape
    parity_type: int
    tree: str
    shift_L: float = 0.0     # TPS: L = ibond * l_b (0 for plain H)

    @property
    def ibond(self) -> int:
        return len(self.ibond_type)

    @property
    def dim(self) -> int:
        return self.q ** self.n

    def fortran_xlocal(self) -> np.ndarray:
        """complex128 flat array in the Fortran memory order of xlocal_Ham(0:qq-1,0:qq-1,9)."""
        return np.ascontiguousarray(np.transpose(self.xlocal, (0, 2, 1))).reshape(-1)

    def fortran_llocal(self) -> np.ndarray:
        """int32 flat array in the Fortran memory order of llocal_Ham (1 = skipped)."""
        return np.ascontiguousarray(np.transpose(self.llocal, (0, 2, 1))).reshape(-1).astype(np.int32)


def bond_norm_shift(cluster, model: int = 20, param: float = 0.0, q: int = 2, field: float = 0.0,
                    factor: float = 1.2) -> float:
    """A per-bond mTPQ shift l_b = factor * max_t ||h_t||_2 over the bond matrices in use:
    then L = ibond * l_b >= factor * sum_b ||h_b|| bounds the spectrum of H from above, and
    l = L / N = factor * max|bond sum| / N (no reference counterpart: the reference's l_b is
    sum_t sum_k |h_kk|^2 / c_count)."""
    tree = "s12" if q == 2 else "spin"
    H = bond_matrices(set_lattice(model, param, tree), q, tree, field=field)
    used = sorted({int(t) - 1 for t in cluster.ibond_type})
    return factor * max(float(np.linalg.norm(H[t], 2)) for t in used)


def build_operator(cluster, model: int = 20, param: float = 0.0, q: int = 2,
                   tree: str = None, field: float = 0.0, tps: bool = False,
                   tps_lb: Optional[float] = None) -> Operator:
    """Reference data_mem tables for a cluster/model; ``param`` in the reference's units.
    ``tps``: the shifted operator l_b - h_b of change_to_TPS_hamiltonian, with the
    reference's l_b or, when given, ``tps_lb``."""
    if tree is None:
        tree = "s12" if q == 2 else "spin"
    c = set_lattice(model, param, tree)
    H = bond_matrices(c, q, tree, field=field)
    mask, parity = local_mask(H, q, tree)
    L = 0.0
    if tps:
        H, lb, L = tps_shift(H, q, cluster.ibond, c.c_count, tree, lb=tps_lb)
    return Operator(cluster.n, q, cluster.ipair, cluster.ibond_type, H, mask, parity, tree, L)
