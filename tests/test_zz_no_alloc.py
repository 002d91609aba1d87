"""The allocation-free H.v (pned_p_mltply_v3_noalloc) against pned_p_mltply_v3.

Both run the same pass kernels; the allocation-free entry keeps its per-block partial sums
in a static device array and writes <v1|(H - shift)|v1> to device memory.  Two calls in a
row on one GPU, the second on a smaller context (the scratch keeps the first call's
partials), must each equal the ordinary entry point's v0 and prdct exactly.
"""
import sys

import numpy as np
import pytest

from pned_amd import gpu, lattice, model

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if gpu.device_count() < 1:
        pytest.skip("no GPU")


def _rand(d, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(d) + 1j * r.standard_normal(d)


def _case(spec, phi, h, eperbd, seed):
    cl = lattice.get_cluster(spec)
    op = model.build_operator(cl, 20, phi, field=h)
    ctx = gpu.Context(op)
    d = ctx.dim
    x, y = _rand(d, seed), _rand(d, seed + 1)
    v1, v0, out = ctx.vector(), ctx.vector(), ctx.vector()
    v1.upload(x)
    v0.upload(y)
    pr_ref = ctx.p_mltply_v3(v1, v0, eperbd)     # also builds the pass tables
    want = v0.download()
    v0.upload(y)
    out.upload(np.full(d, np.nan + 1j * np.nan))
    ctx.p_mltply_v3_noalloc(v1, v0, eperbd, out.data_ptr())
    ctx.sync()
    got = v0.download()
    pr = out.download(0, 1)[0]
    return ctx, got, want, pr, pr_ref


def test_noalloc_matches_p_mltply_v3_twice():
    _say("noalloc N=16 field, then N=12")
    ctx_a, got, want, pr, pr_ref = _case(16, 126.0, 0.1, 0.05, 3)
    assert np.array_equal(got, want)
    assert pr == pr_ref
    ctx_b, got, want, pr, pr_ref = _case(12, 54.0, 0.0, 0.0, 7)
    assert np.array_equal(got, want)
    assert pr == pr_ref


def test_noalloc_needs_a_plan():
    _say("noalloc before any H.v")
    op = model.build_operator(lattice.get_cluster(8), 20, 54.0)
    ctx = gpu.Context(op)
    v1, v0, out = ctx.vector(), ctx.vector(), ctx.vector()
    with pytest.raises(gpu.PnedError) as e:
        ctx.p_mltply_v3_noalloc(v1, v0, 0.0, out.data_ptr())
    assert e.value.code == gpu.E_ARG
    ctx.p_mltply_v3(v1, v0, 0.0)
    ctx.p_mltply_v3_noalloc(v1, v0, 0.0, out.data_ptr())
    ctx.sync()
