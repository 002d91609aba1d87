"""Honeycomb cluster tables: bond list, bond types, site positions, momenta.

Restates the geometry of ``set_bond_info`` (S_one_half/Hybrid_TPS_params.f90:344-1369,
identical in S_arbitrary/Hybrid_TPS9_params.f90:345-1407) for the nearest-neighbour
(``simple == 0``) clusters the reference defines: N = 8, 12, 16, 18, 24, 32.  Every
reference model sets ``simple = 0`` (Hybrid_TPS_params.f90:150,165), so only the
three nearest-neighbour bond types x=1, y=2, z=3 occur.

Conventions kept from the reference:

* ``ipair`` is 1-based, two entries per bond, in the reference's bond order;
* site ``k`` (1-based) is basis bit ``k-1``;
* ``site_vec`` and ``q_list`` use the reference's primitive vectors
  a1=(1/2, 1/(2*sqrt3)), a2=(1/2, -1/(2*sqrt3)), a3=(0, 1/sqrt3) and
  k1=(2*pi, 0), k2=(0, 2*pi/sqrt3) (Hybrid_TPS_params.f90:367-374).

The hand-built clusters N = 8, 12, 18, 24 have no generating rule in the reference:
their bond tables are the reference's cluster definitions, kept as data so the
Hamiltonian is the same operator.  N = 16 and N = 32 follow the reference's loops.

New clusters that the reference does not define (rectangular ``L1 x L2`` PBC
tori: 2x4, 3x3, 2x5, 3x4, 2x7, 4x4 ...) are built by :func:`torus`, which
generalises the reference's N = 32 construction (Hybrid_TPS_params.f90:376-403)
to any L1, L2.  ``torus(4, 4)`` reproduces the reference N = 32 cluster.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

_S3 = math.sqrt(3.0)
A1 = np.array([0.5, 0.5 / _S3])
A2 = np.array([0.5, -0.5 / _S3])
A3 = np.array([0.0, 1.0 / _S3])
K1 = np.array([2.0 * math.pi, 0.0])
K2 = np.array([0.0, 2.0 * math.pi / _S3])


@dataclass
class Cluster:
    """One finite cluster.  ``bonds`` holds (site1, site2, type) with 1-based sites."""

    n: int
    bonds: List[Tuple[int, int, int]]
    site_vec: np.ndarray                 # (n, 2)
    q_list: np.ndarray                   # (q_num, 2)
    q_list_symmetric: Optional[List[int]] = None   # 1-based indices into q_list
    name: str = ""
    c_count: int = 3

    @property
    def ibond(self) -> int:
        return len(self.bonds)

    @property
    def ipair(self) -> np.ndarray:
        out = np.empty(2 * self.ibond, dtype=np.int32)
        for b, (s, t, _) in enumerate(self.bonds):
            out[2 * b] = s
            out[2 * b + 1] = t
        return out

    @property
    def ibond_type(self) -> np.ndarray:
        return np.array([t for _, _, t in self.bonds], dtype=np.int32)

    def sq_momenta(self) -> np.ndarray:
        """Momenta used by the mTPQ static structure factor (q_list_symmetric)."""
        if not self.q_list_symmetric:
            return np.zeros((0, 2))
        return np.array([self.q_list[k - 1] for k in self.q_list_symmetric])


def _pairs(flat: List[int], types: List[int]) -> List[Tuple[int, int, int]]:
    assert len(flat) == 2 * len(types)
    return [(flat[2 * i], flat[2 * i + 1], types[i]) for i in range(len(types))]


# --- reference N=8 (Hybrid_TPS_params.f90:791-806, 900-937) ------------------------
def _cluster8() -> Cluster:
    flat = [1, 2, 3, 7, 5, 4, 6, 8,
            7, 1, 2, 3, 6, 5, 4, 8,
            1, 6, 8, 2, 3, 4, 5, 7]
    bonds = _pairs(flat, [1] * 4 + [2] * 4 + [3] * 4)
    sv = np.zeros((8, 2))
    sv[0] = 1.5 * A3 + A1
    sv[1] = sv[0] + A1
    sv[2] = sv[1] + A2
    sv[3] = sv[2] - A3
    sv[4] = sv[3] - A1
    sv[5] = sv[4] - A2
    sv[6] = sv[4] - A3
    sv[7] = sv[1] + A3
    q = [K1 / 2 * i + K2 / 2 * j for j in (0, 2) for i in (0, 2)]
    q.append(K1 / 2 + K2 / 2)
    ql = np.zeros((9, 2))
    ql[:5] = np.array(q)
    return Cluster(8, bonds, sv, ql, [1, 2, 3, 4, 5], name="ref8")


# --- reference N=12 (Hybrid_TPS_params.f90:939-953, 1033-1083) ---------------------
def _cluster12() -> Cluster:
    flat = [1, 2, 5, 6, 9, 10, 3, 4, 7, 8, 11, 12,
            2, 5, 6, 9, 10, 1, 4, 7, 8, 11, 12, 3,
            1, 4, 5, 8, 9, 12, 3, 2, 7, 6, 11, 10]
    bonds = _pairs(flat, [1] * 6 + [2] * 6 + [3] * 6)
    sv = np.zeros((12, 2))
    k = 0
    for i in range(3):
        base = 0.5 * A3 + i * (A1 + A2)
        sv[k] = base
        sv[k + 1] = sv[k] + A1
        sv[k + 2] = sv[k + 1] + A3
        sv[k + 3] = sv[k + 2] + A1
        k += 4
    ql, sym = [], []
    for j in (0, 2):
        for i in range(4):
            ql.append(K1 / 3 * i + K2 / 2 * j)
            if i in (0, 3) and j in (0, 2):
                sym.append(len(ql))
    return Cluster(12, bonds, sv, np.array(ql), sym, name="ref12")


# --- reference N=16 (Hybrid_TPS_params.f90:1186-1240, 1340-1352) -------------------
def _cluster16() -> Cluster:
    L1 = L2 = 2
    bonds = []
    sv = np.zeros((16, 2))
    for j in range(L2):
        for i in range(L1):
            c = 4 * L1 * j + 4 * i + 1
            ip = 4 * L1 * j + 4 * ((i + 1) % L1) + 1
            jp = 4 * L1 * ((j + 1) % L2) + 4 * i + 1
            bonds += [(c + 0, c + 1, 1), (c + 2, ip + 3, 1),
                      (c + 1, c + 2, 3), (c + 3, jp + 0, 3),
                      (c + 1, ip + 0, 2), (c + 3, c + 2, 2)]
            base = 0.5 * A3 + (A1 + A2) * i + A3 * (3 * j)
            k = c - 1
            sv[k] = base
            sv[k + 1] = base + A1
            sv[k + 2] = base + A1 + A3
            sv[k + 3] = base + A1 + A3 - A2
    ql, sym = [], []
    for j in range(L2 + 1):
        for i in range(L1 + 1):
            ql.append(K1 / L1 * i + K2 / L2 * j)
            if i in (0, L1) and j in (0, L2):
                sym.append(len(ql))
    return Cluster(16, bonds, sv, np.array(ql), sym, name="ref16")


# --- reference N=18 (Hybrid_TPS_params.f90:1084-1185) ------------------------------
def _cluster18() -> Cluster:
    flat = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18,
            6, 1, 2, 3, 4, 5, 12, 7, 8, 9, 10, 11, 18, 13, 14, 15, 16, 17,
            1, 14, 3, 16, 5, 18, 7, 2, 9, 4, 11, 6, 13, 8, 15, 10, 17, 12]
    bonds = _pairs(flat, [1] * 9 + [2] * 9 + [3] * 9)
    sv = np.zeros((18, 2))
    k = 0
    for base, cnt in ((0.5 * A3, 6), (1.5 * A3 + A1, 6), (2.5 * A3 + 2 * A1, 3)):
        for i in range(1, cnt + 1):
            if i % 2 == 1:
                sv[k] = base + 0.5 * (i - 1) * (A1 + A2)
            else:
                sv[k] = sv[k - 1] + A1
            k += 1
    # the reference fills 15 of 18 positions (three chains of 6, 6, 3 sites); the
    # remaining rows stay zero, as an unset allocatable would read on most systems
    ql = [K1 / 3 * i + K2 / 2 * j for j in (0, 2) for i in range(4)]
    # q_list_symmetric is never allocated for N=18 in the reference (D10 quirk)
    return Cluster(18, bonds, sv, np.array(ql), None, name="ref18")


# --- reference N=24 (Hybrid_TPS_params.f90:516-538, 746-789) -----------------------
def _cluster24() -> Cluster:
    flat = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12,
            13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24,
            1, 18, 2, 13, 3, 20, 4, 15, 5, 22, 6, 17,
            7, 24, 8, 19, 9, 14, 10, 21, 11, 16, 12, 23,
            1, 12, 3, 2, 5, 4, 7, 6, 9, 8, 11, 10,
            13, 24, 15, 14, 17, 16, 19, 18, 21, 20, 23, 22]
    bonds = _pairs(flat, [1] * 12 + [3] * 12 + [2] * 12)
    sv = np.zeros((24, 2))
    base = 0.5 * A3
    for i in range(1, 13):
        sv[i - 1] = base + (A1 + A2) * 0.5 * (i - 1) if i % 2 == 1 else sv[i - 2] + A1
    base = A1 + A3 + sv[0]
    for i in range(13, 25):
        sv[i - 1] = base + (A1 + A2) * 0.5 * (i - 13) if i % 2 == 1 else sv[i - 2] + A1
    ql, sym = [], []
    for j in (0, 2):
        for i in range(4):
            ql.append(K1 / 3 * i + K2 / 2 * j)
            if i in (0, 3) and j in (0, 2):
                sym.append(len(ql))
    for i in (1, 3, 5):
        ql.append(K1 / 6 * i + K2 / 2 * 1)
        if i == 3:
            sym.append(len(ql))
    ql.append(np.zeros(2))   # q_num = 12 but only 11 momenta are set by the reference
    return Cluster(24, bonds, sv, np.array(ql), sym, name="ref24")


def torus(L1: int, L2: int) -> Cluster:
    """Rectangular L1 x L2 honeycomb PBC cluster, N = 2*L1*L2.

    Generalises the reference's N = 32 loop (Hybrid_TPS_params.f90:376-403, site
    positions 466-478, momenta 480-510): cell (i, j) holds sites s = 2*(j*L1+i)+1
    (sublattice A) and s+1 (sublattice B); bond x joins A(i,j)-B(i,j), y joins
    A(i,j)-B(i-1,j) and z joins A(i,j)-B(i,j-1), both with periodic wrap.
    """
    n = 2 * L1 * L2
    bonds = []
    for j in range(L2):
        for i in range(L1):
            s = 2 * (j * L1 + i) + 1
            bonds.append((s, 2 * (j * L1 + i) + 2, 1))
            bonds.append((s, 2 * (j * L1 + (i - 1) % L1) + 2, 2))
            bonds.append((s, 2 * (((j - 1) % L2) * L1 + i) + 2, 3))
    sv = np.zeros((n, 2))
    base = 0.5 * A3
    k = 0
    for j in range(L2):
        for i in range(L1):
            sv[k] = base + (A1 + A2) * i + (A1 + A3) * j
            sv[k + 1] = sv[k] + A1
            k += 2
    if (L1, L2) == (4, 4):
        ql, sym = [], []
        for j in range(3):
            for i in range(3):
                ql.append(K1 / 2 * i + K2 / 2 * j)
                if i % 2 == 0 and j % 2 == 0:
                    sym.append(len(ql))
        for j in range(2):
            for i in range(2):
                ql.append(K1 / 2 * i + K2 / 2 * j + (K1 + K2) / 4)
                if i == 1 and j == 1:
                    sym.append(len(ql))
        return Cluster(n, bonds, sv, np.array(ql), sym, name="ref32")
    ql, sym = [], []
    for m2 in range(L2):
        for m1 in range(L1):
            ql.append(torus_momentum(L1, L2, m1, m2))
    sym = [1]
    return Cluster(n, bonds, sv, np.array(ql), sym, name=f"torus{L1}x{L2}")


def torus_momentum(L1: int, L2: int, m1: int, m2: int) -> np.ndarray:
    """Allowed momentum (m1/L1) b1 + (m2/L2) b2 of the L1 x L2 torus.

    The torus translations are T1 = a1 + a2 (cell i -> i+1) and T2 = a1 + a3
    (cell j -> j+1); b1, b2 are the dual vectors with b_a . T_b = 2 pi delta_ab.
    """
    T = np.array([A1 + A2, A1 + A3])          # rows are T1, T2
    B = 2.0 * math.pi * np.linalg.inv(T).T    # rows are b1, b2
    return (m1 / L1) * B[0] + (m2 / L2) * B[1]


_REF = {8: _cluster8, 12: _cluster12, 16: _cluster16, 18: _cluster18,
        24: _cluster24, 32: lambda: torus(4, 4)}

_NAMED = {"2x4": (2, 4), "3x3": (3, 3), "2x5": (2, 5), "3x4": (3, 4),
          "2x7": (2, 7), "4x4": (4, 4), "2x2": (2, 2), "2x3": (2, 3), "1x4": (1, 4)}


def reference_cluster(n: int) -> Cluster:
    """The reference's cluster of N sites (Hybrid_TPS_params.f90:1364-1366 rejects others)."""
    if n not in _REF:
        raise ValueError(f"Lattice is not supported: N={n} (reference clusters: {sorted(_REF)})")
    return _REF[n]()


def get_cluster(spec) -> Cluster:
    """``spec`` is an int (reference cluster) or a string 'L1xL2' (new torus) or 'refN'."""
    if isinstance(spec, (int, np.integer)):
        return reference_cluster(int(spec))
    s = str(spec)
    if s.startswith("ref"):
        return reference_cluster(int(s[3:]))
    if s in _NAMED:
        return torus(*_NAMED[s])
    if "x" in s:
        a, b = s.split("x")
        return torus(int(a), int(b))
    return reference_cluster(int(s))
