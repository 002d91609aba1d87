"""Vector files written from and read into device memory (vec_q*, basis*): round trips,
subrecord splits, and a Fortran program reading the GPU-written file."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from pned_amd import gpu, lattice, model

pytestmark = pytest.mark.gpu
FC = shutil.which("amdflang") or "/opt/rocm/llvm/bin/amdflang"


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if gpu.device_count() < 1:
        pytest.skip("no GPU")


@pytest.mark.parametrize("limit", [0, 1000, 4096 + 8])
def test_vector_file_round_trip(tmp_path, limit):
    _say("vector file", limit)
    gpu.io_subrecord_limit(limit)
    try:
        op = model.build_operator(lattice.reference_cluster(12), 20, 54.0)
        ctx = gpu.Context(op)
        a, b = ctx.vector(), ctx.vector()
        ctx.set_initial_vec(a, 1)
        ctx.set_initial_vec(b, 2)
        xa, xb = a.download(), b.download()
        f = tmp_path / "vec_q00003me00000zz"
        ctx.dump_vectors(f, 17, [a, b])
        assert os.path.getsize(f) >= 2 * 16 * ctx.dim + 12
        c, d = ctx.vector(), ctx.vector()
        assert ctx.read_vectors(f, [c, d]) == 17
        assert np.array_equal(c.download(), xa) and np.array_equal(d.download(), xb)
        g = tmp_path / "basisme00000"
        ctx.basis_write(g, -7.125, a)
        e = ctx.vector()
        assert ctx.basis_read(g, e) == -7.125
        assert np.array_equal(e.download(), xa)
    finally:
        gpu.io_subrecord_limit(0)


READER = """
program rv
  implicit none
  integer :: indx
  complex(kind(0d0)) :: v1(4096), v2(4096)
  open(11, file='vec_q00003me00000zz', form='unformatted', access='sequential', status='old')
  read(11) indx
  read(11) v1
  read(11) v2
  close(11)
  write(*,'(I6)') indx
  write(*,'(4ES25.17)') v1(1), v2(4096)
end program rv
"""


@pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not available")
def test_fortran_reads_gpu_vector_file(tmp_path):
    _say("fortran reads vector file")
    op = model.build_operator(lattice.reference_cluster(12), 20, 54.0)
    ctx = gpu.Context(op)
    a, b = ctx.vector(), ctx.vector()
    ctx.set_initial_vec(a, 5)
    ctx.set_initial_vec(b, 6)
    ctx.dump_vectors(tmp_path / "vec_q00003me00000zz", 3, [a, b])
    (tmp_path / "rv.f90").write_text(READER)
    subprocess.run([FC, "rv.f90", "-o", "rv"], cwd=tmp_path, check=True, capture_output=True)
    out = subprocess.run(["./rv"], cwd=tmp_path, check=True, capture_output=True, text=True).stdout.split()
    assert out[0] == "3"
    v = [float(x) for x in out[1:]]
    xa, xb = a.download(), b.download()
    assert v[0] == xa[0].real and v[1] == xa[0].imag and v[2] == xb[-1].real and v[3] == xb[-1].imag


def test_interval_dump_round_trip(tmp_path):
    """Checkpoint as the TPS restart writes it (sa, sb, sc, se, vec1, vec2, vec3) and read
    back into fresh device vectors; the real arrays keep Fortran order."""
    _say("interval_dump round trip")
    op = model.build_operator(lattice.reference_cluster(12), 20, 54.0)
    ctx = gpu.Context(op)
    vs = [ctx.vector() for _ in range(3)]
    for k, v in enumerate(vs):
        ctx.set_initial_vec(v, 10 + k)
    want = [v.download() for v in vs]
    r = np.random.default_rng(4)
    sa, sb, sc, se = r.standard_normal(20), r.standard_normal(20), r.standard_normal((20, 9)), r.standard_normal(8)
    f = tmp_path / "vecdump_me00000"
    ctx.interval_dump(f, ctx.dim, 20, reals=[sa, sb, sc, se], vecs=vs)
    back = [ctx.vector() for _ in range(3)]
    ds, vsz, arrs = ctx.interval_dump_read(f, real_counts=[20, 20, 180, 8], vecs=back)
    assert (ds, vsz) == (ctx.dim, 20)
    assert np.array_equal(arrs[0], sa) and np.array_equal(arrs[1], sb) and np.array_equal(arrs[3], se)
    assert np.array_equal(arrs[2].reshape(9, 20).T, sc)
    for v, w in zip(back, want):
        assert np.array_equal(v.download(), w)
