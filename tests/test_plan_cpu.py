"""CPU checks of the S=1/2 H.v planner (no GPU): the term decomposition and pass
assignment exported by libpned_gpu.so are replayed with numpy and compared with the
oracle's p_mltply_v3 restatement.  A dropped term, a wrong selector or a pass whose
tile does not contain a term's flip bits fails here."""
import numpy as np
import pytest

import oracle
from pned_amd import gpu, lattice, model


def emulate(plan, n, v1, v0_in, eperbd, ibond, colbits=0):
    """Apply the planned passes exactly as the kernel does (memory index = state)."""
    D = 1 << n
    m = np.arange(D, dtype=np.int64)
    s = m
    out = v0_in.copy()
    # diagonal
    d = np.full(D, plan["c0diag"], dtype=complex)
    for kind, shift, mask, wr, wi in plan["diag"]:
        mask = int(mask)
        f = (s & mask) if kind == 0 else ((s ^ (s >> int(shift))) & mask)
        pc = np.array([bin(x).count("1") for x in range(1 << 16)], dtype=np.int64)
        cnt = pc[f & 0xFFFF] + pc[(f >> 16) & 0xFFFF]
        d += cnt * (wr + 1j * wi)
    d += -ibond * eperbd
    out += d * v1
    seen = np.zeros(len(plan["gmask"]), dtype=int)
    for p in plan["passes"]:
        tb = p["tb"]
        for t in range(p["first"], p["first"] + p["n"]):
            seen[t] += 1
            tm = int(plan["tmask"][t])
            mm = 0
            for k in range(p["K"]):
                if tm >> k & 1:
                    mm |= 1 << tb[k]
            assert (mm >> colbits) == plan["gmask"][t]
            sel = np.array([bin(x).count("1") & 1 for x in range(1 << 16)])
            sm = int(plan["selmask"][t])
            sv = (sel[(s & sm) & 0xFFFF] ^ sel[((s & sm) >> 16) & 0xFFFF]).astype(bool)
            c = np.where(sv, plan["c1"][t], plan["c0"][t])
            out += c * v1[m ^ mm]
    assert np.all(seen == 1)
    assert plan["passes"][0]["diag"] and not any(p["diag"] for p in plan["passes"][1:])
    return out


@pytest.mark.parametrize("spec,phi,h,K", [(12, 54.0, 0.0, 8), (12, 126.0, 0.1, 9), (16, 200.0, 0.15, 10),
                                          ("2x4", 252.0, 0.3, 12), (8, 54.0, 0.2, 12)])
def test_plan_matches_oracle(spec, phi, h, K):
    c = lattice.get_cluster(spec)
    op = model.build_operator(c, 20, phi, field=h)
    plan = gpu.s12_plan(op, 0, K, 4)
    rng = np.random.default_rng(7)
    D = 2 ** c.n
    v1 = rng.standard_normal(D) + 1j * rng.standard_normal(D)
    v0 = rng.standard_normal(D) + 1j * rng.standard_normal(D)
    ep = 0.11
    ref = v0.copy()
    oracle.hv(oracle.OracleOp(op), v1, ref, eperbd=ep)
    got = emulate(plan, c.n, v1, v0, ep, c.ibond)
    assert np.linalg.norm(got - ref) <= 1e-13 * np.linalg.norm(ref)


def test_plan_tile_sizes_n32():
    op = model.build_operator(lattice.reference_cluster(32), 20, 126.0, field=0.1)
    plan = gpu.s12_plan(op, 0, 12, 4)
    assert all(p["K"] == 12 for p in plan["passes"])
    assert all(p["tb"][:4] == [0, 1, 2, 3] for p in plan["passes"])
    # 48 double flips + 32 single flips
    assert len(plan["gmask"]) == 80
    assert len(plan["passes"]) <= 8


def _mem_index(site_bit, n):
    """Memory index of every basis state under a layout (site k at memory bit site_bit[k])."""
    s = np.arange(1 << n, dtype=np.int64)
    m = np.zeros_like(s)
    for k in range(n):
        m |= ((s >> k) & 1) << int(site_bit[k])
    return m


def _swap_table(n, pairs):
    sb = np.arange(n, dtype=np.int32)
    for a, b in pairs:
        sb[a], sb[b] = b, a
    return sb


@pytest.mark.parametrize("spec,phi,h,K,pairs", [
    (12, 54.0, 0.0, 8, [(0, 7), (2, 11)]),
    (12, 126.0, 0.1, 9, [(1, 5)]),
    ("2x4", 252.0, 0.3, 11, [(0, 15), (1, 9), (2, 4), (3, 12)]),
    (18, 33.0, 0.2, 12, None),            # the layout the planner chooses for this cluster
])
def test_plan_with_layout_matches_oracle(spec, phi, h, K, pairs):
    """A layout moves the terms' bits; applied to the vectors laid out the same way, the
    planned passes must reproduce the oracle's H.v in basis order."""
    c = lattice.get_cluster(spec)
    op = model.build_operator(c, 20, phi, field=h)
    sb = gpu.s12_layout(op) if pairs is None else _swap_table(c.n, pairs)
    assert not np.array_equal(sb, np.arange(c.n))
    assert np.array_equal(sb[sb], np.arange(c.n))       # an involution
    plan = gpu.s12_plan(op, 0, K, 3, site_bit=sb)
    mem = _mem_index(sb, c.n)
    rng = np.random.default_rng(11)
    D = 2 ** c.n
    v1 = rng.standard_normal(D) + 1j * rng.standard_normal(D)
    v0 = rng.standard_normal(D) + 1j * rng.standard_normal(D)
    v1m, v0m = np.empty_like(v1), np.empty_like(v0)
    v1m[mem] = v1
    v0m[mem] = v0
    ep = 0.05
    got = emulate(plan, c.n, v1m, v0m, ep, c.ibond)[mem]
    ref = v0.copy()
    oracle.hv(oracle.OracleOp(op), v1, ref, eperbd=ep)
    assert np.linalg.norm(got - ref) <= 1e-13 * np.linalg.norm(ref)


def test_layout_saves_a_pass_at_n32():
    """The N=32 torus with the [111] field: 2^13-state tiles need 5 passes with sites 1-3 in
    the forced low bits, 4 with the layout the planner chooses."""
    op = model.build_operator(lattice.reference_cluster(32), 20, 126.0, field=0.1)
    sb = gpu.s12_layout(op)
    assert np.array_equal(sb[sb], np.arange(32))
    ident = gpu.s12_plan(op, 0, 13, 3)
    lay = gpu.s12_plan(op, 0, 13, 3, site_bit=sb)
    assert len(ident["passes"]) == 5
    assert len(lay["passes"]) == 4
    for plan in (ident, lay):
        assert all(p["tb"][:3] == [0, 1, 2] for p in plan["passes"])
        assert len(plan["gmask"]) == 80
