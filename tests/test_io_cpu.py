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
  double precision :: out(4*5)
  ! working_area arrays of the module, filled as the mTPQ loop fills them
  allocate(pned_alpha(12), pned_beta(12), pned_coef(12, 9))
  do i = 1, 12
    pned_alpha(i) = -0.1d0 * i
    pned_beta(i) = 1d0 + 0.01d0 * i
    do k = 1, 9
      pned_coef(i,k) = -0.5d0 + 0.02d0 * i - 0.001d0 * k * i
    end do
  end do
  ! the reference's argument list (comp, indx1, indx2, v1, v2, v3); the vectors are not written
  call dump_alpha_beta_coef('tp', 1, 13, c_null_ptr, c_null_ptr, c_null_ptr)
  call pned_check(pned_ctpq_eval(12_c_int, 9_c_int, 1_c_int, pned_alpha, pned_beta, pned_coef, 0.5d0, 4d0, &
                  4_c_int, out), 'ctpq')
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


FORTRAN_DSF_RES = """
program res
  implicit none
  double precision :: q(2), w(3), dsf(3), sq
  integer :: conv(3), k, i, j
  q = (/ 0.5d0, -1.25d0 /)
  w = (/ 0.0d0, 2.5d-3, 4.0d0 /)
  dsf = (/ 1.5d0, -0.0d0, 123456789.25d0 /)
  conv = (/ 1, 0, 1 /)
  sq = 0.1234567890123456789d0
  j = 100
  open(j, file='dsf_res_zz.datcl00000lme00000', status='unknown')
  do i = 1, 2
    write(j,'(A,F25.18)') '#        ', sq
    do k = 1, 3
      write(j,'(2F18.10,2F25.16,I5)') q(1), q(2), w(k), dsf(k), conv(k)
    end do
    write(j,*) ''
    write(j,*) ''
    write(j,*) ''
    write(j,*) ''
    q(1) = q(1) * 2d0
  end do
  close(j)
end program res
"""


@pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not available")
def test_dsf_res_text_matches_fortran(tmp_path):
    """write_dsf_res's text, byte for byte as a Fortran program with the reference's format
    statements writes it (two momenta appended)."""
    (tmp_path / "res.f90").write_text(FORTRAN_DSF_RES)
    subprocess.run([FC, "res.f90", "-o", "res"], cwd=tmp_path, check=True, capture_output=True)
    subprocess.run(["./res"], cwd=tmp_path, check=True, capture_output=True)
    want = (tmp_path / "dsf_res_zz.datcl00000lme00000").read_bytes()
    f = tmp_path / "mine"
    w, d, c = [0.0, 2.5e-3, 4.0], [1.5, -0.0, 123456789.25], [1, 0, 1]
    gpu.write_dsf_res(f, 0.1234567890123456789, (0.5, -1.25), w, d, c, append=False)
    gpu.write_dsf_res(f, 0.1234567890123456789, (1.0, -1.25), w, d, c, append=True)
    assert f.read_bytes() == want
    assert b"*" * 25 in want   # a value wider than F25.16 prints as asterisks in both


FORTRAN_INFO = """
program info
  implicit none
  character(len=128) :: dump_name
  integer :: t_np, t_max_itr, t_s_maxitr, t_n, t_ibond, t_nvec, t_total_sz, i, k
  character(len=2) :: t_dsf_comp
  logical :: position(7)
  integer :: itr(12)
  double precision :: sa(5), sb(5), sc(5,3), se(8)
  integer :: data_size, vec_size
  open(11, file='vecdump_me00000.info', form='unformatted', access='sequential', status='old')
  read(11) dump_name
  read(11) t_np
  read(11) position
  read(11) t_max_itr
  read(11) t_dsf_comp
  read(11) t_s_maxitr
  read(11) t_n
  read(11) t_ibond
  read(11) t_nvec
  read(11) t_total_sz
  do i = 1, 12
    read(11) k
    itr(i) = k
  end do
  close(11)
  write(*,'(A)') trim(dump_name)
  write(*,'(8I8)') t_np, t_max_itr, t_s_maxitr, t_n, t_ibond, t_nvec, t_total_sz, count(position)
  write(*,'(7L2)') position
  write(*,'(A)') t_dsf_comp
  write(*,'(12I6)') itr
  open(12, file='vecdump_me00000', form='unformatted', access='sequential', status='old')
  read(12) data_size
  read(12) vec_size
  read(12) sa
  read(12) sb
  read(12) sc
  read(12) se
  close(12)
  write(*,'(2I8)') data_size, vec_size
  write(*,'(4ES25.17)') sa(2), sb(5), sc(4,2), se(8)
  ! the same statements interval_dump uses, for the library to read back
  position = (/ .false., .true., .false., .true., .false., .false., .true. /)
  open(13, file='fortran_dump.info', form='unformatted', access='sequential', status='unknown')
  write(13) dump_name
  write(13) 2
  write(13) position
  write(13) 300
  write(13) 'xx'
  write(13) 150
  write(13) 24
  write(13) 36
  write(13) 8
  write(13) 12
  do i = 1, 12
    write(13) 10 * i
  end do
  close(13)
end program info
"""


@pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not available")
def test_interval_dump_files_against_fortran(tmp_path):
    """The checkpoint's .info and dump files: the library writes, a Fortran program reads
    them with check_interval_dump's and interval_dump_read's statements; and the library
    reads the .info that interval_dump's write statements produce."""
    pos = [1, 0, 0, 1, 1, 0, 0]
    itc = list(range(101, 113))
    gpu.dump_info_write(tmp_path / "vecdump_me00000.info", "vecdump_me00000", 1, pos, 1024, "zz", 300, 32, 48, 8,
                        16, itc, true_value=1)
    r = np.random.default_rng(9)
    sa, sb, sc, se = r.standard_normal(5), r.standard_normal(5), r.standard_normal((5, 3)), r.standard_normal(8)
    # no device vectors in this CPU test (their records are tested on the GPU)
    lib = gpu.lib()
    arrs = [np.ascontiguousarray(a.reshape(-1, order="F")) for a in (sa, sb, sc, se)]
    ptrs = (gpu.C.c_void_p * 4)(*[a.ctypes.data for a in arrs])
    cnts = (gpu.C.c_int64 * 4)(*[a.size for a in arrs])
    gpu.check(lib.pned_interval_dump(str(tmp_path / "vecdump_me00000").encode(), 4096, 300, 4, ptrs, cnts, 0,
                                     (gpu.C.c_void_p * 1)()))
    (tmp_path / "info.f90").write_text(FORTRAN_INFO)
    subprocess.run([FC, "info.f90", "-o", "info"], cwd=tmp_path, check=True, capture_output=True)
    out = subprocess.run(["./info"], cwd=tmp_path, check=True, capture_output=True, text=True).stdout.splitlines()
    assert out[0] == "vecdump_me00000"
    assert [int(x) for x in out[1].split()] == [1, 1024, 300, 32, 48, 8, 16, 3]
    assert out[2].split() == ["T", "F", "F", "T", "T", "F", "F"]
    assert out[3] == "zz"
    assert [int(x) for x in out[4].split()] == itc
    assert [int(x) for x in out[5].split()] == [4096, 300]
    got = [float(x) for x in out[6].split()]
    assert got == [sa[1], sb[4], sc[3, 1], se[7]]
    info = gpu.dump_info_read(tmp_path / "fortran_dump.info")
    assert info == {"dump_name": "vecdump_me00000", "np": 2,
                    "position": [False, True, False, True, False, False, True], "max_itr": 300,
                    "dsf_comp": "xx", "s_maxitr": 150, "n": 24, "ibond": 36, "nvec": 8, "total_sz": 12,
                    "itr_counter": [10 * i for i in range(1, 13)]}
    # Intel Fortran's default .true. (-1) reads back as .true. as well
    gpu.dump_info_write(tmp_path / "intel.info", "vecdump_me00003", 4, [1, 0, 1], 10, "zz", 5, 8, 12, 1, 4, itc,
                        true_value=-1)
    raw = (tmp_path / "intel.info").read_bytes()
    pos_at = (4 + 128 + 4) + (4 + 4 + 4) + 4   # name record, np record, position marker
    assert struct.unpack_from("<3i", raw, pos_at) == (-1, 0, -1)
    assert gpu.dump_info_read(tmp_path / "intel.info", npos=3)["position"] == [True, False, True]
