"""CPU checks of the C ABI: the library loads without a GPU, exports every symbol that
include/pned_gpu.h declares, and reports the missing device as an error."""
import ctypes
import os
import re

import numpy as np

from pned_amd import gpu, lattice, model

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    out = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            out |= set(re.findall(r"\b(pned_[a-z0-9_]+)\s*\(", text))
    return out


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(gpu.LIB_PATH)
    missing = [s for s in sorted(declared_symbols()) if not hasattr(L, s)]
    assert not missing, missing
    assert len(declared_symbols()) > 30


def test_no_gpu_is_an_error_not_a_fallback():
    if gpu.device_count() > 0:
        return
    op = model.build_operator(lattice.reference_cluster(8), 20, 54.0)
    try:
        gpu.Context(op)
    except gpu.PnedError as e:
        assert e.code == gpu.E_HIP
    else:
        raise AssertionError("context created without a GPU")


def test_host_tridiagonal_solvers_match_oracle():
    import oracle
    rng = np.random.default_rng(1)
    a = rng.standard_normal(80)
    b = rng.random(80) + 0.05
    E = gpu.bisec(a, b, 8)
    assert np.array_equal(E, oracle.bisec(a, b, 8))
    V = gpu.vec12(a, b, E, 3)
    assert np.allclose(V, oracle.vec12(a, b, E, 3), atol=1e-14)
