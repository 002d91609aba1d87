"""The Lanczos tridiagonal's eigenpairs on the GPU (pned_ctx_bisec / pned_ctx_vec12) against
the host entry points pned_bisec / pned_vec12 on the same inputs: bit for bit, since both
run one source with the same rounding.  The host solvers themselves are pinned against
LAPACK in test_abi_cpu.py."""
import sys
import time

import numpy as np
import pytest

from pned_amd import gpu, lattice, model

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    if gpu.device_count() < 1:
        pytest.skip("no GPU")
    c = gpu.Context(model.build_operator(lattice.reference_cluster(12), 20, 54.0))
    yield c
    c.close()


def _tridiag(n, seed, kind):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(n)
    b = rng.random(n) + 0.05
    if kind == "graded":      # off-diagonals over 12 decades
        b *= 10.0 ** rng.uniform(-12, 0, n)
    elif kind == "pairs":     # two decoupled identical blocks: every eigenvalue twice
        h, c = a[: n // 2], b[: n // 2]
        a = np.concatenate([h, h])
        b = np.concatenate([c[:-1], [0.0], c[:-1], [0.0]])
    elif kind == "zero":      # eigenvalues near zero: the stop width ends the bisection
        a = np.round(a, 1)
        b *= 1e-9
    return a, b


@pytest.mark.parametrize("n,seed,kind", [(1, 1, "plain"), (2, 2, "plain"), (7, 3, "plain"), (80, 4, "graded"),
                                         (300, 5, "plain"), (40, 6, "pairs"), (400, 7, "pairs"), (33, 8, "zero"),
                                         (2000, 9, "plain"), (8000, 10, "graded")])
def test_device_tridiagonal_equals_host(ctx, n, seed, kind):
    _say("tridiagonal", n, seed, kind)
    a, b = _tridiag(n, seed, kind)
    n = len(a)
    ne = min(8, n)
    t0 = time.perf_counter()
    e_host = gpu.bisec(a, b, ne)
    t1 = time.perf_counter()
    e_dev = ctx.bisec(a, b, ne)
    t2 = time.perf_counter()
    assert np.array_equal(e_dev, e_host)
    nv = min(4, n)
    v_host = gpu.vec12(a, b, e_host, nv)
    v_dev = ctx.vec12(a, b, e_host, nv)
    assert np.array_equal(v_dev, v_host)
    print(f"[result] n={n}: bisec host {1e3 * (t1 - t0):.2f} ms, device {1e3 * (t2 - t1):.2f} ms", flush=True)


def test_lanczos_checks_use_device_bisection(ctx):
    """lnc1's energies are the device bisection of its own alpha/beta, and equal the host
    bisection of them."""
    _say("lnc1 tridiagonal")
    v1, v0 = ctx.vector(), ctx.vector()
    ctx.set_initial_vec(v1, 5)
    itr, alpha, beta, ene = ctx.lnc1(v1, v0, nvec=8)
    assert np.array_equal(ene, gpu.bisec(alpha, beta, 8))
    assert np.array_equal(ene, ctx.bisec(alpha, beta, 8))
    coef_h = gpu.vec12(alpha, beta, ene[:2], 2)
    coef_d = ctx.vec12(alpha, beta, ene[:2], 2)
    assert np.array_equal(coef_h, coef_d)
    v1.free()
    v0.free()


def test_device_tridiagonal_argument_errors(ctx):
    _say("tridiagonal errors")
    a, b = np.ones(5), np.ones(5)
    big = np.ones(8001)
    for call in (lambda: ctx.bisec(a, b, 6), lambda: ctx.bisec(big, big, 1), lambda: ctx.bisec(big, big, 0),
                 lambda: gpu.bisec(big, big, 0), lambda: ctx.vec12(a, b, np.zeros(0), 0)):
        with pytest.raises(gpu.PnedError) as err:
            call()
        assert err.value.code == gpu.E_ARG
    assert np.array_equal(ctx.bisec(a, b, 2), gpu.bisec(a, b, 2))
    # inverse iteration takes any size, as the host's does
    rng = np.random.default_rng(11)
    a2, b2 = rng.standard_normal(9000), rng.random(9000) + 0.05
    e2 = gpu.bisec(a2[:8000], b2[:8000], 2)
    assert np.array_equal(ctx.vec12(a2, b2, e2, 2), gpu.vec12(a2, b2, e2, 2))
