"""CPU checks of the spin-S tile planner (no GPU): the passes and partner-shift classes that
libpned_gpu.so exports are replayed with numpy and compared with the oracle's H.v
restatement (a dropped element, a wrong shift or coefficient, or a bond applied in a pass
whose tile does not hold both of its sites fails here), and the cover the tiled kernel
runs at spin-1 N=18 is checked for its pass count and forced low sites."""
import numpy as np
import pytest

import oracle
from pned_amd import gpu, lattice, model


def emulate(plan, op, v1, v0_in):
    """v0 + H v1 from the plan: per pass, each class of each bond adds
    coef[out(s)] * v1[s + da q^sa + db q^sb]; each bond adds its matrix diagonal (eperbd 0)."""
    q, n = op.q, op.n
    qq = q * q
    D = q ** n
    s = np.arange(D, dtype=np.int64)
    digit = [(s // q ** k) % q for k in range(n)]
    xl = op.fortran_xlocal().reshape(9, qq, qq)   # [t][out][in]
    ll = op.fortran_llocal().reshape(9, qq, qq)
    ip = np.asarray(op.ipair).reshape(-1, 2) - 1
    out = v0_in.copy()
    applied = np.zeros(op.ibond, dtype=int)
    for p in plan["passes"]:
        tile = set(p["sites"])
        for b in p["bonds"]:
            sa, sb = int(min(ip[b])), int(max(ip[b]))
            assert sa in tile and sb in tile
            applied[b] += 1
            t = int(op.ibond_type[b]) - 1
            o = digit[sa] + q * digit[sb]
            dg = np.array([0 if ll[t, k, k] else xl[t, k, k] for k in range(qq)])
            out += dg[o] * v1
        for c in p["classes"]:
            b = c["bond"]
            assert b in p["bonds"]
            sa, sb = int(min(ip[b])), int(max(ip[b]))
            o = digit[sa] + q * digit[sb]
            ok = ((digit[sa] + c["da"] >= 0) & (digit[sa] + c["da"] < q) & (digit[sb] + c["db"] >= 0)
                  & (digit[sb] + c["db"] < q))
            partner = np.where(ok, s + c["da"] * q ** sa + c["db"] * q ** sb, s)
            coef = c["coef"][o]
            assert np.all(coef[~ok] == 0)
            out += coef * v1[partner]
    assert np.all(applied == 1)
    return out


@pytest.mark.parametrize("spec,q,phi,h,tile_sites,low_sites", [
    (8, 3, 108.0, 0.0, 0, 0),
    (8, 3, 108.0, 0.2, 5, 0),       # field: complex classes; smaller tiles, more passes
    (8, 3, 54.0, 0.1, 4, 2),
    (8, 4, 108.0, 0.1, 4, 0),       # spin-3/2
    (8, 5, 72.0, 0.0, 3, 0),        # spin-2
])
def test_spin_plan_replays_oracle_hv(spec, q, phi, h, tile_sites, low_sites):
    cl = lattice.get_cluster(spec)
    op = model.build_operator(cl, 20, phi, q=q, field=h)
    plan = gpu.spin_plan(op, tile_sites=tile_sites, low_sites=low_sites)
    rng = np.random.default_rng(3)
    v1 = rng.standard_normal(op.dim) + 1j * rng.standard_normal(op.dim)
    v0 = rng.standard_normal(op.dim) + 1j * rng.standard_normal(op.dim)
    want = v0.copy()
    oracle.hv(oracle.OracleOp(op), v1, want, eperbd=0.0)
    got = emulate(plan, op, v1, v0)
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    L = plan["low_sites"]
    for p in plan["passes"]:
        assert p["sites"][:L] == list(range(L))
        assert len(p["classes"]) <= 64


def test_spin1_n18_cover():
    """The spin-1 N=18 workload: 8-site tiles (3^8 states), 5 passes with the lowest 3 sites
    forced (runs of 27 states, 432 bytes) -- the greedy cover alone needs 7 at L=3 -- and no
    pass count gained by forcing only 2."""
    cl = lattice.reference_cluster(18)
    op = model.build_operator(cl, 20, 108.0, q=3)
    plan = gpu.spin_plan(op)
    assert plan["low_sites"] == 3
    assert len(plan["passes"]) == 5
    assert all(len(p["sites"]) == 8 for p in plan["passes"])
    assert len(gpu.spin_plan(op, low_sites=2)["passes"]) >= 5
    assert sorted(b for p in plan["passes"] for b in p["bonds"]) == list(range(op.ibond))
