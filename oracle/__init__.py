"""ctypes front of the CPU oracle (``oracle/pned_oracle.c``).

TEST INFRASTRUCTURE ONLY: the reference restatement that the tests, the smoke entry and
the bench's ``cpu_baseline`` leg compare the HIP path against.  The product never
imports it.
Each wrapper names the reference routine the C function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libpned_oracle.so")
_lib = None

_zp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_bp = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


class _Z(C.Structure):
    """C ``double complex`` passed/returned by value (SysV: two SSE registers)."""
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


def build() -> str:
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        i32, i64, f64, u64, u32 = C.c_int, C.c_int64, C.c_double, C.c_uint64, C.c_uint32
        L.orc_hv_rows.argtypes = [i32, i32, i32, _ip, _ip, _zp, _bp, f64, _zp, _zp, _zp, i64, i64]
        L.orc_hv.argtypes = [i32, i32, i32, _ip, _ip, _zp, _bp, f64, _zp, _zp, _zp]
        L.orc_dot.argtypes = [i64, _zp, _zp]
        L.orc_dot.restype = _Z
        L.orc_beta2.argtypes = [i64, _Z, _zp, _zp]
        L.orc_beta2.restype = f64
        L.orc_philox4x32_10.argtypes = [_up, _up, _up]
        L.orc_gauss_fill.argtypes = [i64, u64, u32, _zp]
        L.orc_gauss_fill.restype = f64
        L.orc_set_initial_vec.argtypes = [i64, u64, u32, _zp]
        L.orc_bisec.argtypes = [_dp, _dp, i32, _dp, i32, f64]
        L.orc_bisec.restype = i32
        L.orc_vec12.argtypes = [_dp, _dp, i32, _dp, i32, _dp]
        L.orc_lanczos.argtypes = [i32, i32, i32, _ip, _ip, _zp, _bp, _zp, _zp, i32, i32, i32, _dp, _dp, _dp]
        L.orc_lanczos.restype = i32
        L.orc_inverse_iteration.argtypes = [i32, i32, i32, _ip, _ip, _zp, _bp, f64,
                                            _zp, _zp, _zp, _zp, _zp, i32, i32, i32]
        L.orc_inverse_iteration.restype = i32
        L.orc_sz_q.argtypes = [i32, i32, _dp, _dp, _zp, _zp]
        L.orc_intalv_xy.argtypes = [i32, _dp, _dp, i64, _up, _zp, i64, _up, _up, _zp]
        L.orc_intalv_pm.argtypes = [i32, i32, _dp, _dp, i32, _zp, _zp]
        L.orc_sq_apply.argtypes = [i32, i32, _dp, _dp, i32, _zp, _zp]
        L.orc_zz_bond_apply.argtypes = [i32, i32, i32, _ip, _ip, _zp, _zp]
        L.orc_tps.argtypes = [i32, i32, i32, _ip, _ip, _ip, _zp, _bp, _dp, i32, _dp, f64, i32,
                              _zp, _zp, _zp, _zp, _dp, _dp, _dp]
        L.orc_dsf_lanczos.argtypes = [i32, i32, i32, _ip, _ip, _zp, _bp, f64, i32, _zp, _zp, _dp, _dp]
        L.orc_dsf_lanczos.restype = i32
        L.orc_renbun.argtypes = [i32, i32, f64, f64, f64, _dp, _dp, _dp, _ip, _dp]
        L.orc_normalize_dsf.argtypes = [i32, f64, f64, _dp]
        L.orc_check1.argtypes = [i32, i32, i32, _ip, _ip, _zp, _bp, _zp, _zp, _zp, _zp, i32]
        L.orc_check1.restype = i32
        L.orc_fill_range.argtypes = [u64, i64, i64, C.c_void_p]
        L.orc_mk_list12.argtypes = [i32, i32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_mk_list12.restype = i64
        L.orc_hv_sector.argtypes = [i32, i32, _ip, _ip, _zp, _bp, f64, i32, i64, _up, _up, _up, _zp, _zp, _zp]
        L.orc_hv_sector.restype = i32
        L.orc_ctpq_sample.argtypes = [i32, f64, _dp, _dp, _dp]
        L.orc_ctpq_sample.restype = f64
        L.orc_ctpq.argtypes = [i32, i32, i32, _dp, _dp, _dp, f64, f64, i32, _dp]
        L.orc_num_threads.restype = i32
        _lib = L
    return _lib


class OracleOp:
    """The operator tables in the oracle's layout, built from a ``pned_amd.model.Operator``."""

    def __init__(self, op):
        self.n, self.q, self.ibond = op.n, op.q, op.ibond
        ip = np.asarray(op.ipair, dtype=np.int64).reshape(-1, 2) - 1
        self.site_pair = np.ascontiguousarray(np.sort(ip, axis=1).reshape(-1).astype(np.int32))
        self.btype = np.ascontiguousarray(np.asarray(op.ibond_type, dtype=np.int32))
        self.ipair = np.ascontiguousarray(np.asarray(op.ipair, dtype=np.int32))
        self.xl = np.ascontiguousarray(op.xlocal.astype(np.complex128).reshape(-1))
        self.mk = np.ascontiguousarray(op.llocal.astype(np.uint8).reshape(-1))
        self.dim = self.q ** self.n

    def args(self):
        return (self.n, self.q, self.ibond, self.site_pair, self.btype, self.xl, self.mk)


def hv(oop: OracleOp, v1: np.ndarray, v0: np.ndarray, eperbd: float = 0.0) -> complex:
    """p_mltply_v3 (Hybrid_TPS_parallel.f90:190-263): v0 += (H - ibond*eperbd) v1; returns prdct."""
    pr = np.zeros(1, dtype=np.complex128)
    lib().orc_hv(*oop.args(), float(eperbd), v1, v0, pr)
    return complex(pr[0])


def sector_tables(n: int, tot_sz: int):
    """mk_list12 (Hybrid_TPS_parallel.f90:452-578) at np=1: (list1, la, lb) for the states
    with tot_sz set bits; index(s) = la[s & right] + lb[s >> ihf] - 1 (0-based)."""
    ihf0, ihf = n // 2, (n + 1) // 2
    ldim = lib().orc_mk_list12(n, tot_sz, None, None, None)
    list1 = np.zeros(ldim, dtype=np.uint32)
    la = np.zeros(1 << ihf, dtype=np.uint32)
    lb = np.zeros(1 << ihf0, dtype=np.uint32)
    lib().orc_mk_list12(n, tot_sz, list1.ctypes.data, la.ctypes.data, lb.ctypes.data)
    return list1, la, lb


def hv_sector(oop: OracleOp, tot_sz: int, v1: np.ndarray, v0: np.ndarray, eperbd: float = 0.0) -> complex:
    """H_dot_v over the S^z sector basis (parity_type = 0): v0 += (H - ibond*eperbd) v1."""
    assert oop.q == 2
    list1, la, lb = sector_tables(oop.n, tot_sz)
    pr = np.zeros(1, dtype=np.complex128)
    rc = lib().orc_hv_sector(oop.n, oop.ibond, oop.site_pair, oop.btype, oop.xl, oop.mk, float(eperbd), int(tot_sz),
                             len(list1), list1, la, lb, v1, v0, pr)
    if rc != 0:
        raise ValueError("the operator does not conserve S^z")
    return complex(pr[0])


def ctpq_sample(beta, c4, c5, bt: float) -> float:
    """calc_fac + eval_q (Hybrid_TPS_eval_mpi.f90:236-303) for one mTPQ sample."""
    b = np.ascontiguousarray(beta, dtype=np.float64)
    return lib().orc_ctpq_sample(len(b), float(bt), b, np.ascontiguousarray(c4, dtype=np.float64),
                                 np.ascontiguousarray(c5, dtype=np.float64))


def ctpq(alpha, beta, coef, t_min: float, t_max: float, num_div: int) -> np.ndarray:
    """The cTPQ program (Hybrid_TPS_eval_mpi.f90:383-446).  alpha, beta: (seeds, nd1);
    coef: (seeds, nd2, nd1), i.e. the reference's Fortran arrays.  Returns
    (num_div+1, 4): T, C, variance, dC/dT."""
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    c = np.ascontiguousarray(coef, dtype=np.float64)
    seeds, nd1 = b.shape
    nd2 = c.shape[1]
    out = np.zeros(4 * (num_div + 1))
    lib().orc_ctpq(nd1, nd2, seeds, a.reshape(-1), b.reshape(-1), c.reshape(-1), float(t_min), float(t_max),
                   int(num_div), out)
    return out.reshape(num_div + 1, 4)


def hv_rows(oop: OracleOp, v1, v0, lo: int, hi: int, eperbd: float = 0.0) -> complex:
    pr = np.zeros(1, dtype=np.complex128)
    lib().orc_hv_rows(*oop.args(), float(eperbd), v1, v0, pr, int(lo), int(hi))
    return complex(pr[0])


def dot(x, y) -> complex:
    r = lib().orc_dot(len(x), x, y)
    return complex(r.re, r.im)


def philox(ctr, key):
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(np.asarray(ctr, dtype=np.uint32), np.asarray(key, dtype=np.uint32), out)
    return out


def gaussian(dim: int, seed: int, col: int = 0, normalize: bool = True) -> np.ndarray:
    v = np.zeros(dim, dtype=np.complex128)
    if normalize:
        lib().orc_set_initial_vec(dim, seed, col, v)
    else:
        lib().orc_gauss_fill(dim, seed, col, v)
    return v


def bisec(alpha, beta, ne: int, eps: float = 1e-25):
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    E = np.zeros(max(ne, 1))
    rc = lib().orc_bisec(a, b, len(a), E, ne, eps)
    if rc != 0:
        raise ValueError("bisec argument error")
    return E[:ne]


def vec12(alpha, beta, E, nvec: int):
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    e = np.ascontiguousarray(E, dtype=np.float64)
    v = np.zeros(len(a) * nvec)
    lib().orc_vec12(a, b, len(a), e, nvec, v)
    return v.reshape(nvec, len(a))


def lanczos(oop: OracleOp, v_start, max_itr=1024, nvec=8, fixed_steps=0):
    """parallelized_lnc1z (Hybrid_TPS_Lanczos.f90:48-163) from an injected start vector."""
    v1 = np.array(v_start, dtype=np.complex128, copy=True)
    v0 = np.zeros_like(v1)
    steps = fixed_steps if fixed_steps > 0 else max_itr
    alpha = np.zeros(steps + 1)
    beta = np.zeros(steps + 1)
    ene = np.zeros(max(nvec, 1))
    itr = lib().orc_lanczos(*oop.args(), v1, v0, max_itr, nvec, fixed_steps, alpha, beta, ene)
    return itr, alpha[:abs(itr)], beta[:abs(itr)], ene


def inverse_iteration(oop: OracleOp, eig: float, b_start, max_inv=5000, max_cg=500, saved_loop=True):
    """parallelized_inv1z + parallelized_cg1 (Hybrid_TPS_cg.f90:18-197)."""
    d = oop.dim
    b = np.array(b_start, dtype=np.complex128, copy=True)
    x, p, r, y = (np.zeros(d, dtype=np.complex128) for _ in range(4))
    it = lib().orc_inverse_iteration(*oop.args(), float(eig), x, b, p, r, y, max_inv, max_cg,
                                     1 if saved_loop else 0)
    return it, x


def sz_q(n, q, site_vec, qv, x):
    v = np.zeros_like(x)
    lib().orc_sz_q(n, q, np.ascontiguousarray(site_vec, dtype=np.float64).reshape(-1),
                   np.asarray(qv, dtype=np.float64), x, v)
    return v


def intalv_xy(n, site_vec, qv, tot_sz, x):
    """intalv_xy (Hybrid_TPS_DSF.f90:89-193): x in the sector with tot_sz set bits ->
    normalised S^+(q) x in the sector with tot_sz + 1, and cnorm = ||S^+(q) x||^2 / n
    (the reference's tiny-norm floor 1e-25 included)."""
    listt, _, _ = sector_tables(n, tot_sz)
    _, la, lb = sector_tables(n, tot_sz + 1)
    ldim = lib().orc_mk_list12(n, tot_sz + 1, None, None, None)
    v = np.zeros(ldim, dtype=np.complex128)
    lib().orc_intalv_xy(n, np.ascontiguousarray(site_vec, dtype=np.float64).reshape(-1),
                        np.asarray(qv, dtype=np.float64), len(listt), listt,
                        np.ascontiguousarray(x, dtype=np.complex128), ldim, la, lb, v)
    cn = float(np.vdot(v, v).real)
    if cn < 1e-25:
        cn = 1e-25
    return v / np.sqrt(cn), cn / n


def intalv_pm(n, q, site_vec, qv, ipm, x):
    """intalv_xy over the full q^n space (parity_type != 0): normalised
    sum_j e^{i q.r_j} S^{ipm}_j x (spin-S ladder factors; q = 2: bit j set or cleared)
    and cnorm = ||.||^2 / n."""
    v = np.zeros(q ** n, dtype=np.complex128)
    lib().orc_intalv_pm(n, q, np.ascontiguousarray(site_vec, dtype=np.float64).reshape(-1),
                        np.asarray(qv, dtype=np.float64), int(ipm), np.ascontiguousarray(x, dtype=np.complex128), v)
    cn = float(np.vdot(v, v).real)
    if cn < 1e-25:
        cn = 1e-25
    return v / np.sqrt(cn), cn / n


def tps(oop: OracleOp, site_vec, sq_q, lmax, nsteps, v_start):
    """parallelized_TPSz + eval_phys_val (Hybrid_TPS_TPS.f90:83-111, 325-409)."""
    d = oop.dim
    v1 = np.array(v_start, dtype=np.complex128, copy=True)
    v0, vg, vt = (np.zeros(d, dtype=np.complex128) for _ in range(3))
    sq_q = np.ascontiguousarray(np.asarray(sq_q, dtype=np.float64).reshape(-1))
    nsq = len(sq_q) // 2
    alpha = np.zeros(nsteps)
    beta = np.zeros(nsteps)
    coef = np.zeros(nsteps * (7 + nsq))
    lib().orc_tps(oop.n, oop.q, oop.ibond, oop.site_pair, oop.btype, oop.ipair, oop.xl, oop.mk,
                  np.ascontiguousarray(site_vec, dtype=np.float64).reshape(-1),
                  nsq, sq_q if nsq else np.zeros(2), float(lmax), nsteps, v1, v0, vg, vt, alpha, beta, coef)
    return alpha, beta, coef.reshape(nsteps, 7 + nsq)


def dsf_lanczos(oop: OracleOp, v_exc, cnorm, s_maxitr):
    """parallelized_dsflnc1z (Hybrid_TPS_DSF.f90:195-261)."""
    v1 = np.array(v_exc, dtype=np.complex128, copy=True)
    v0 = np.zeros_like(v1)
    sa = np.zeros(s_maxitr)
    sb = np.zeros(s_maxitr + 1)
    lib().orc_dsf_lanczos(*oop.args(), float(cnorm), s_maxitr, v1, v0, sa, sb)
    return sa, sb


def renbun(salpha, sbeta, ene, eta, i_range, split, s_maxitr):
    """exec_renbun (Hybrid_TPS_DSF.f90:263-334)."""
    w = np.zeros(i_range)
    dsf = np.zeros(i_range)
    conv = np.zeros(i_range, dtype=np.int32)
    sa = np.ascontiguousarray(salpha, dtype=np.float64)
    sb = np.ascontiguousarray(sbeta, dtype=np.float64)
    lib().orc_renbun(i_range, s_maxitr, split, ene, eta, sa, sb, w, conv, dsf)
    return w, dsf, conv


def normalize_dsf(dsf, split, sq):
    d = np.ascontiguousarray(dsf, dtype=np.float64).copy()
    lib().orc_normalize_dsf(len(d), split, float(sq), d)
    return d


def cpu_rows_sample(op, rows: int, seed: int = 20261202):
    """Time the oracle's reference-structured H.v on output rows [0, rows) of ``op``.

    Used for the CPU baseline at sizes whose full vectors do not fit in host memory
    (N = 32: 64 GiB per complex128 vector).  v1 and v0 are anonymous NORESERVE
    mappings of the full dimension; only the blocks the row slice touches are filled
    (the partner blocks i ^ mask of every flip term), so the memory actually used is a
    few GB.  Returns (seconds, rows, threads).
    """
    import mmap
    import time

    oop = OracleOp(op)
    d = oop.dim
    nbytes = d * 16
    flags = mmap.MAP_PRIVATE | mmap.MAP_ANONYMOUS | getattr(mmap, "MAP_NORESERVE", 0x4000)
    m1 = mmap.mmap(-1, nbytes, flags=flags, prot=mmap.PROT_READ | mmap.PROT_WRITE)
    m0 = mmap.mmap(-1, nbytes, flags=flags, prot=mmap.PROT_READ | mmap.PROT_WRITE)
    v1 = np.frombuffer(m1, dtype=np.complex128)
    v0 = np.frombuffer(m0, dtype=np.complex128)
    # blocks touched by rows [0, rows): every partner index i ^ m stays in the block
    # (m & ~(B-1)) of size B = next power of two >= rows
    B = 1 << max(0, (rows - 1).bit_length())
    masks = set()
    pairs = oop.site_pair.reshape(-1, 2)
    for a, b in pairs:
        for m in (1 << int(a), 1 << int(b), (1 << int(a)) | (1 << int(b)), 0):
            masks.add(m & ~(B - 1))
    for base in sorted(masks):
        lib().orc_fill_range(seed, base, base + B, v1.ctypes.data)
    pr = np.zeros(1, dtype=np.complex128)
    t = time.perf_counter()
    lib().orc_hv_rows(*oop.args(), 0.0, v1, v0, pr, 0, rows)
    dt = time.perf_counter() - t
    threads = lib().orc_num_threads()
    del v1, v0
    m1.close()
    m0.close()
    return dt, rows, threads
