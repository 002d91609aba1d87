"""The reference's result files (Fortran unformatted sequential) through the library's
writers and readers: byte layout of records and subrecords, round trips, and files read
and written by a Fortran program using the reference's own read/write statements."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from pned_amd import gpu

FC = shutil.which("amdflang") or "/opt/rocm/llvm/bin/amdflang"


def _records(path):
    """(leading marker, trailing marker, payload) of every subrecord in the file."""
    out = []
    data = open(path, "rb").read()
    pos = 0
    while pos < len(data):
        lead = struct.unpack_from("<i", data, pos)[0]
        n = abs(lead)
        payload = data[pos + 4: pos + 4 + n]
        trail = struct.unpack_from("<i", data, pos + 4 + n)[0]
        out.append((lead, trail, payload))
        pos += 8 + n
    return out


@pytest.fixture(autouse=True)
def _default_limit():
    gpu.io_subrecord_limit(0)
    yield
    gpu.io_subrecord_limit(0)


def _tps_arrays(nd1=11, nd2=7, seed=0):
    r = np.random.default_rng(seed)
    return r.standard_normal(nd1), r.random(nd1), r.standard_normal((nd1, nd2))


def test_alpha_beta_coef_layout_and_round_trip(tmp_path):
    a, b, c = _tps_arrays()
    f = tmp_path / "SalphaSbeta_me00000seed00001tp"
    gpu.dump_alpha_beta_coef(f, a, b, c)
    recs = _records(f)
    assert [r[0] for r in recs] == [4, 4, 88, 88, 616] and all(r[0] == r[1] for r in recs)
    assert struct.unpack("<i", recs[0][2])[0] == 11 and struct.unpack("<i", recs[1][2])[0] == 7
    assert np.array_equal(np.frombuffer(recs[4][2], dtype="<f8"), c.T.reshape(-1))   # Fortran order
    a2, b2, c2 = gpu.read_alpha_beta_coef(f)
    assert np.array_equal(a2, a) and np.array_equal(b2, b) and np.array_equal(c2, c)


def test_subrecords_follow_the_documented_split(tmp_path):
    """616-byte coef record with a 200-byte limit: subrecords of 200, 200, 200, 16 bytes;
    leading marker negated when another subrecord follows, trailing one negated when
    another precedes."""
    gpu.io_subrecord_limit(200)
    a, b, c = _tps_arrays()
    f = tmp_path / "ab"
    gpu.dump_alpha_beta_coef(f, a, b, c)
    recs = _records(f)
    assert [(r[0], r[1]) for r in recs[4:]] == [(-200, 200), (-200, -200), (-200, -200), (16, -16)]
    assert b"".join(r[2] for r in recs[4:]) == np.ascontiguousarray(c.T).tobytes()
    a2, b2, c2 = gpu.read_alpha_beta_coef(f)
    assert np.array_equal(c2, c)


def test_dsf_alpha_beta_round_trip(tmp_path):
    r = np.random.default_rng(3)
    sa, sb = r.standard_normal(30), r.random(31)
    f = tmp_path / "SalphaSbeta_q00001me00000zz"
    gpu.dump_dsf_alpha_beta(f, 0.375 - 0.125j, sa, sb)
    # sq is complex(kind(0d0)) in the reference (Hybrid_TPS_DSF.f90:358): a 16-byte record
    recs = _records(f)
    assert [x[0] for x in recs] == [16, 240, 248]
    assert recs[0][2] == np.array([0.375, -0.125]).tobytes()
    sq, sa2, sb2 = gpu.read_dsf_alpha_beta(f, 30, 31)
    assert sq == 0.375 - 0.125j and np.array_equal(sa2, sa) and np.array_equal(sb2, sb)


FORTRAN_READ = """
program rd
  implicit none
  integer :: num_data1, num_data2
  double precision, allocatable :: ta(:), tb(:), tc(:,:)
  complex(kind(0d0)) :: sq
  double precision :: sa(30), sb(31)
  open(11, file='SalphaSbeta_me00000seed00001tp', form='unformatted', access='sequential', status='old')
  read(11) num_data1
  read(11) num_data2
  allocate(ta(num_data1), tb(num_data1), tc(num_data1, num_data2))
  read(11) ta
  read(11) tb
  read(11) tc
  close(11)
  write(*,'(2I6)') num_data1, num_data2
  write(*,'(ES25.17)') ta(3), tb(5), tc(4,6), sum(tc)
  open(12, file='SalphaSbeta_q00001me00000zz', form='unformatted', access='sequential', status='old')
  read(12) sq
  read(12) sa
  read(12) sb
  close(12)
  write(*,'(ES25.17)') dble(sq), dimag(sq), sa(7), sb(31)
end program rd
"""

FORTRAN_WRITE = """
program wr
  implicit none
  integer :: i, k
  double precision :: alpha(9), beta(9), coef(9,8)
  complex(kind(0d0)) :: sq
  do i = 1, 9
    alpha(i) = 0.5d0 * i
    beta(i) = 1d0 / i
    do k = 1, 8
      coef(i,k) = i + 100d0 * k
    end do
  end do
  open(11, file='SalphaSbeta_me00000seed00002tp', form='unformatted', access='sequential', status='unknown')
  write(11) 9
  write(11) 8
  write(11) alpha
  write(11) beta
  write(11) coef
  close(11)
  sq = dcmplx(1.5d0, -0.25d0)
  open(12, file='SalphaSbeta_q00002me00000zz', form='unformatted', access='sequential', status='unknown')
  write(12) sq
  write(12) alpha
  write(12) beta
  close(12)
end program wr
"""


@pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not available")
def test_fortran_program_reads_and_writes_the_same_files(tmp_path):
    a, b, c = _tps_arrays(12, 9, seed=4)
    gpu.dump_alpha_beta_coef(tmp_path / "SalphaSbeta_me00000seed00001tp", a, b, c)
    r = np.random.default_rng(5)
    sa, sb = r.standard_normal(30), r.random(31)
    gpu.dump_dsf_alpha_beta(tmp_path / "SalphaSbeta_q00001me00000zz", -2.25 + 0.5j, sa, sb)
    for name, src in (("rd", FORTRAN_READ), ("wr", FORTRAN_WRITE)):
        (tmp_path / f"{name}.f90").write_text(src)
        subprocess.run([FC, f"{name}.f90", "-o", name], cwd=tmp_path, check=True, capture_output=True)
    out = subprocess.run(["./rd"], cwd=tmp_path, check=True, capture_output=True, text=True).stdout.split()
    assert out[:2] == ["12", "9"]
    vals = [float(x) for x in out[2:]]
    assert vals[0] == a[2] and vals[1] == b[4] and vals[2] == c[3, 5]
    assert abs(vals[3] - c.sum()) <= 1e-12 * np.abs(c).sum()
    assert vals[4:6] == [-2.25, 0.5] and vals[6] == sa[6] and vals[7] == sb[30]
    subprocess.run(["./wr"], cwd=tmp_path, check=True, capture_output=True)
    a2, b2, c2 = gpu.read_alpha_beta_coef(tmp_path / "SalphaSbeta_me00000seed00002tp")
    i = np.arange(1, 10)
    assert np.array_equal(a2, 0.5 * i) and np.array_equal(b2, 1.0 / i)
    assert np.array_equal(c2, i[:, None] + 100.0 * np.arange(1, 9)[None, :])
    sq, sa3, sb3 = gpu.read_dsf_alpha_beta(tmp_path / "SalphaSbeta_q00002me00000zz", 9, 9)
    assert sq == 1.5 - 0.25j and np.array_equal(sa3, 0.5 * i) and np.array_equal(sb3, 1.0 / i)


FORTRAN_BINDINGS = """
program bind
  use iso_c_binding
  use pned_gpu
  use pned_parallelized_diag
  implicit none
  integer :: i, k
  double precision :: alpha(12), beta(12), coef(12,9), out(4*5)
  do i = 1, 12
    alpha(i) = -0.1d0 * i
    beta(i) = 1d0 + 0.01d0 * i
    do k = 1, 9
      coef(i,k) = -0.5d0 + 0.02d0 * i - 0.001d0 * k * i
    end do
  end do
  call dump_alpha_beta_coef('tp', 1, 0, alpha, beta, coef)
  call pned_check(pned_ctpq_eval(12_c_int, 9_c_int, 1_c_int, alpha, beta, coef, 0.5d0, 4d0, 4_c_int, out), 'ctpq')
  write(*,'(ES25.17)') out
end program bind
"""


@pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not available")
def test_fortran_module_bindings(tmp_path):
    """The ISO_C_BINDING layer's file and cTPQ entry points, called from Fortran: the file
    it writes is the one the library's own reader returns, and the cTPQ columns equal the
    library's through the Python binding."""
    here = os.path.dirname(gpu.LIB_PATH)
    mod = os.path.join(here, "fortran", "_build")
    if not os.path.exists(os.path.join(mod, "libpned_fortran.a")):
        pytest.skip("Fortran layer not built")
    (tmp_path / "bind.f90").write_text(FORTRAN_BINDINGS)
    subprocess.run([FC, "-I", mod, "bind.f90", os.path.join(mod, "libpned_fortran.a"), "-L", here,
                    "-lpned_gpu", f"-Wl,-rpath,{here}", "-o", "bind"], cwd=tmp_path, check=True,
                   capture_output=True)
    out = subprocess.run(["./bind"], cwd=tmp_path, check=True, capture_output=True, text=True).stdout
    got = np.array([float(x) for x in out.split()]).reshape(5, 4)
    i = np.arange(1, 13, dtype=float)
    k = np.arange(1, 10, dtype=float)
    alpha, beta = -0.1 * i, 1.0 + 0.01 * i
    coef = -0.5 + 0.02 * i[:, None] - 0.001 * k[None, :] * i[:, None]
    a2, b2, c2 = gpu.read_alpha_beta_coef(tmp_path / "SalphaSbeta_me00000seed00001tp")
    assert np.array_equal(a2, alpha) and np.array_equal(b2, beta) and np.array_equal(c2, coef)
    want = gpu.ctpq(alpha[None, :], beta[None, :], coef.T[None, :, :], 0.5, 4.0, 4)
    np.testing.assert_allclose(got, want, rtol=1e-15, atol=0)
