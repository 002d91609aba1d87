"""Python front of libpned_gpu.so (the C ABI in include/pned_gpu.h).

Method names follow the reference modules they replace: ``p_mltply_v3``,
``calc_vec_inprdct``, ``calc_alpha_beta`` ... (S_one_half/Hybrid_TPS_parallel.f90:6-8),
``lnc1`` (Hybrid_TPS_Lanczos.f90:7-163), ``inv1`` (Hybrid_TPS_cg.f90:7-100), ``tps``
(Hybrid_TPS_TPS.f90:45-409), ``intalv_zz``/``dsflnc1z``/``renbun`` (Hybrid_TPS_DSF.f90).
Errors raise :class:`PnedError` with the library's message (the reference prints and
stops).  There is no CPU fallback: without the built library or a GPU every call fails.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpned_gpu.so")

C128, C64 = 0, 1
E_OK, E_ARG, E_NOMEM, E_HIP, E_UNSUPPORTED, E_BREAKDOWN, E_NOCONV = 0, -1, -2, -3, -4, -5, -6


class PnedError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pned error {code}: {msg}")
        self.code = code


_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    """Load libpned_gpu.so (built in-tree by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PnedError(E_HIP, f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P, i32, i64, f64, u64 = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_uint64
    L.pned_last_error.restype = C.c_char_p
    L.pned_device_count.restype = i32
    L.pned_ctx_create.argtypes = [i32, i32, i32, _ip, _ip, _dp, _ip, i32, C.POINTER(P)]
    L.pned_ctx_destroy.argtypes = [P]
    L.pned_ctx_dim.argtypes = [P]
    L.pned_ctx_dim.restype = i64
    L.pned_ctx_set_matrices.argtypes = [P, _dp, _ip]
    L.pned_ctx_tune.argtypes = [P, i32, i32, i32]
    L.pned_ctx_plan_info.argtypes = [P, C.POINTER(i32), _ip, i32]
    L.pned_ctx_spin_kernel.argtypes = [P, i32]
    L.pned_ctx_graphs.argtypes = [P, i32]
    L.pned_ctx_s12_kernel.argtypes = [P, i32]
    L.pned_ctx_sz_sector.argtypes = [P, i32]
    L.pned_io_subrecord_limit.argtypes = [C.c_uint64]
    L.pned_dump_alpha_beta_coef.argtypes = [C.c_char_p, i32, i32, _dp, _dp, _dp]
    L.pned_read_alpha_beta_coef.argtypes = [C.c_char_p, C.POINTER(i32), C.POINTER(i32), C.c_void_p, C.c_void_p,
                                            C.c_void_p, i32, i32]
    L.pned_dump_dsf_alpha_beta.argtypes = [C.c_char_p, _dp, i32, _dp, i32, _dp]
    L.pned_read_dsf_alpha_beta.argtypes = [C.c_char_p, _dp, i32, _dp, i32, _dp]
    L.pned_dump_vectors.argtypes = [C.c_char_p, i32, i32, C.POINTER(P)]
    L.pned_read_vectors.argtypes = [C.c_char_p, C.POINTER(i32), i32, C.POINTER(P)]
    L.pned_basis_write.argtypes = [C.c_char_p, C.c_double, P]
    L.pned_basis_read.argtypes = [C.c_char_p, C.POINTER(C.c_double), P]
    L.pned_interval_dump.argtypes = [C.c_char_p, i32, i32, i32, C.POINTER(C.c_void_p), C.POINTER(C.c_int64), i32,
                                     C.POINTER(P)]
    L.pned_interval_dump_read.argtypes = [C.c_char_p, C.POINTER(i32), C.POINTER(i32), i32, C.POINTER(C.c_void_p),
                                          C.POINTER(C.c_int64), i32, C.POINTER(P)]
    L.pned_dump_info_write.argtypes = [C.c_char_p, C.c_char_p, i32, i32, _ip, i32, i32, C.c_char_p, i32, i32, i32,
                                       i32, i32, _ip]
    L.pned_dump_info_read.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(i32), i32, _ip, C.POINTER(i32), C.c_char_p,
                                      C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), C.POINTER(i32),
                                      C.POINTER(i32), _ip]
    L.pned_write_dsf_res.argtypes = [C.c_char_p, i32, f64, _dp, i32, _dp, _dp, _ip]
    L.pned_ctpq_eval.argtypes = [i32, i32, i32, _dp, _dp, _dp, C.c_double, C.c_double, i32, _dp]
    L.pned_ctx_autotune.argtypes = [P, P, P, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32),
                                    C.POINTER(f64)]
    L.pned_ctx_sync.argtypes = [P]
    L.pned_ctx_stream.argtypes = [P]
    L.pned_ctx_stream.restype = P
    L.pned_vec_create.argtypes = [P, i32, i32, C.POINTER(P)]
    L.pned_vec_destroy.argtypes = [P]
    L.pned_vec_upload.argtypes = [P, P, P, i64, i64]
    L.pned_vec_download.argtypes = [P, P, P, i64, i64]
    L.pned_vec_zero.argtypes = [P, P]
    L.pned_vec_copy.argtypes = [P, P, P]
    L.pned_vec_data.argtypes = [P]
    L.pned_vec_data.restype = P
    L.pned_set_initial_vec.argtypes = [P, P, u64]
    L.pned_p_mltply_v3.argtypes = [P, P, P, f64, _dp]
    L.pned_hv_fresh.argtypes = [P, P, P, f64, _dp]
    L.pned_p_mltply_v3_noalloc.argtypes = [P, P, P, f64, P]
    L.pned_calc_vec_inprdct.argtypes = [P, P, P, _dp]
    L.pned_calc_beta.argtypes = [P, _dp, P, P, _dp]
    L.pned_calc_alpha_beta.argtypes = [P, P, P, _dp, _dp]
    L.pned_calc_alpha_beta2.argtypes = [P, P, P, _dp, _dp]
    L.pned_axpby.argtypes = [P, _dp, P, _dp, P]
    L.pned_scale.argtypes = [P, P, _dp]
    L.pned_lanczos_update.argtypes = [P, P, P, _dp, _dp]
    L.pned_lnc1.argtypes = [P, P, P, i32, i32, i32, _dp, _dp, _dp, C.POINTER(i32)]
    L.pned_lnc1_span.argtypes = [P, P, P, i32, i32, i32, i32, i32, _dp, _dp, _dp, C.POINTER(i32)]
    L.pned_tps_span.argtypes = [P, P, P, P, P, f64, i32, i32, i32, _dp, _dp, _dp]
    L.pned_lnc1_reorth.argtypes = [P, C.POINTER(P), i32, P, i32, i32, _dp, _dp, _dp, C.POINTER(i32)]
    L.pned_dsflnc1z_reorth.argtypes = [P, C.POINTER(P), i32, P, f64, i32, _dp, _dp]
    L.pned_dsflnc1z_span.argtypes = [P, P, P, f64, i32, i32, i32, _dp, _dp, C.POINTER(i32)]
    L.pned_bisec.argtypes = [_dp, _dp, i32, _dp, i32, f64]
    L.pned_vec12.argtypes = [_dp, _dp, i32, _dp, i32, _dp]
    L.pned_ctx_bisec.argtypes = [P, _dp, _dp, i32, _dp, i32, f64]
    L.pned_ctx_vec12.argtypes = [P, _dp, _dp, i32, _dp, i32, _dp]
    L.pned_ritz_vector.argtypes = [P, P, P, P, i32, _dp, _dp, _dp]
    L.pned_inv1.argtypes = [P, f64, P, P, P, P, P, i32, i32, i32, C.POINTER(i32)]
    L.pned_cocg.argtypes = [P, f64, f64, P, P, P, P, P, i32, f64, C.POINTER(i32), C.POINTER(f64)]
    L.pned_check1.argtypes = [P, P, P, _dp, _dp, i32, C.POINTER(i32)]
    L.pned_residual.argtypes = [P, P, f64, P, C.POINTER(f64)]
    L.pned_tps_set_observables.argtypes = [P, _dp, i32, _dp]
    L.pned_tps.argtypes = [P, P, P, P, P, f64, i32, _dp, _dp, _dp]
    L.pned_tps_fused.argtypes = [P, P, P, P, f64, i32, _dp, _dp, _dp]
    L.pned_tps_fused_graph.argtypes = [P, P, P, P, f64, i32, i32, _dp, _dp, _dp]
    L.pned_intalv_zz.argtypes = [P, P, P, _dp, _dp, C.POINTER(f64)]
    L.pned_intalv_xy.argtypes = [P, P, P, _dp, _dp, i32, C.POINTER(f64)]
    L.pned_dsflnc1z.argtypes = [P, P, P, f64, i32, _dp, _dp]
    L.pned_renbun.argtypes = [P, i32, i32, f64, f64, f64, _dp, _dp, f64, _dp, _dp, _ip]
    L.pned_ctx_create_batch.argtypes = [i32, i32, i32, _ip, _ip, i32, _dp, _ip, i32, C.POINTER(P)]
    L.pned_ctx_set_column_matrices.argtypes = [P, i32, _dp, _ip]
    L.pned_lanczos_batch_step.argtypes = [P, P, P, _ip, _dp, _dp, _dp, _dp]
    L.pned_lanczos_batch_steps.argtypes = [P, P, P, _ip, _dp, _dp, i32, _dp, _dp]
    L.pned_set_initial_column.argtypes = [P, P, i32, u64]
    L.pned_s12_plan_export.argtypes = [i32, i32, _ip, _ip, _dp, _ip, i32, i32, i32, C.POINTER(i32),
                                       C.POINTER(i32), C.POINTER(i32), _ip, _ip, _dp,
                                       np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS"), _dp, i32, i32,
                                       C.c_void_p]
    L.pned_spin_plan_export.argtypes = [i32, i32, i32, _ip, _ip, _dp, _ip, i32, i32, C.POINTER(i32),
                                        C.POINTER(i32), _ip, _ip, _dp, _ip, i32]
    L.pned_s12_layout_choose.argtypes = [i32, i32, _ip, _ip, _dp, _ip, _ip]
    L.pned_ctx_layout.argtypes = [P, _ip]
    L.pned_ctx_set_layout.argtypes = [P, C.c_void_p]
    _lib = L
    return L


def check(rc: int):
    if rc != 0:
        raise PnedError(rc, lib().pned_last_error().decode(errors="replace"))


def device_count() -> int:
    return int(lib().pned_device_count())


class Vector:
    """A device vector of ``dim`` states x ``nvec`` interleaved columns."""

    def __init__(self, ctx: "Context", nvec: int = 1, dtype: int = C128):
        self.ctx = ctx
        self.nvec = nvec
        self.dtype = dtype
        h = C.c_void_p()
        check(lib().pned_vec_create(ctx.h, nvec, dtype, C.byref(h)))
        self.h = h

    @property
    def np_dtype(self):
        return np.complex128 if self.dtype == C128 else np.complex64

    def upload(self, host: np.ndarray, offset: int = 0):
        a = np.ascontiguousarray(host, dtype=self.np_dtype)
        count = a.size // self.nvec
        check(lib().pned_vec_upload(self.ctx.h, self.h, a.ctypes.data, offset, count))

    def download(self, offset: int = 0, count: Optional[int] = None) -> np.ndarray:
        if count is None:
            count = self.ctx.dim - offset
        out = np.empty(count * self.nvec, dtype=self.np_dtype)
        check(lib().pned_vec_download(self.ctx.h, self.h, out.ctypes.data, offset, count))
        return out.reshape(count, self.nvec) if self.nvec > 1 else out

    def zero(self):
        check(lib().pned_vec_zero(self.ctx.h, self.h))

    def copy_from(self, other: "Vector"):
        check(lib().pned_vec_copy(self.ctx.h, other.h, self.h))

    def data_ptr(self) -> int:
        return lib().pned_vec_data(self.h)

    def free(self):
        if getattr(self, "h", None):
            lib().pned_vec_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One operator on one GPU: the reference's data_mem tables handed to the device."""

    def __init__(self, op, device: int = 0, batch_ops=None):
        """``batch_ops``: a list of operators on the same cluster (power-of-two length),
        one per column of this context's vectors (pned_ctx_create_batch)."""
        self.op = op
        self.n, self.q, self.ibond = op.n, op.q, op.ibond
        h = C.c_void_p()
        ipair = np.ascontiguousarray(op.ipair, dtype=np.int32)
        btype = np.ascontiguousarray(op.ibond_type, dtype=np.int32)
        if batch_ops is None:
            check(lib().pned_ctx_create(op.n, op.q, op.ibond, ipair, btype,
                                        np.ascontiguousarray(op.fortran_xlocal().view(np.float64)),
                                        op.fortran_llocal(), device, C.byref(h)))
            self.nop = 1
        else:
            xl = np.ascontiguousarray(np.concatenate([o.fortran_xlocal().view(np.float64) for o in batch_ops]))
            ll = np.ascontiguousarray(np.concatenate([o.fortran_llocal() for o in batch_ops]))
            check(lib().pned_ctx_create_batch(op.n, op.q, op.ibond, ipair, btype, len(batch_ops), xl, ll, device,
                                              C.byref(h)))
            self.nop = len(batch_ops)
        self.h = h
        self.dim = int(lib().pned_ctx_dim(h))

    def set_column_matrices(self, col: int, op):
        check(lib().pned_ctx_set_column_matrices(self.h, col, np.ascontiguousarray(op.fortran_xlocal().view(np.float64)),
                                                 op.fortran_llocal()))

    def set_initial_column(self, v: "Vector", col: int, seed: int):
        check(lib().pned_set_initial_column(self.h, v.h, col, seed))

    def lanczos_batch_step(self, v1: "Vector", v0: "Vector", mode, alpha_prev, beta_prev):
        nc = v1.nvec
        a = np.zeros(nc)
        b = np.zeros(nc)
        check(lib().pned_lanczos_batch_step(self.h, v1.h, v0.h, np.ascontiguousarray(mode, dtype=np.int32),
                                            np.ascontiguousarray(alpha_prev, dtype=float),
                                            np.ascontiguousarray(beta_prev, dtype=float), a, b))
        return a, b

    def lanczos_batch_steps(self, v1: "Vector", v0: "Vector", mode, alpha_prev, beta_prev, nsteps: int):
        """nsteps batch steps in one call; alpha, beta as (nsteps, nvec) arrays."""
        nc = v1.nvec
        a = np.zeros(nsteps * nc)
        b = np.zeros(nsteps * nc)
        check(lib().pned_lanczos_batch_steps(self.h, v1.h, v0.h, np.ascontiguousarray(mode, dtype=np.int32),
                                             np.ascontiguousarray(alpha_prev, dtype=float),
                                             np.ascontiguousarray(beta_prev, dtype=float), int(nsteps), a, b))
        return a.reshape(nsteps, nc), b.reshape(nsteps, nc)

    # -- management -------------------------------------------------------------------
    def vector(self, nvec: int = 1, dtype: int = C128) -> Vector:
        return Vector(self, nvec, dtype)

    def set_matrices(self, op):
        self.op = op
        check(lib().pned_ctx_set_matrices(self.h, np.ascontiguousarray(op.fortran_xlocal().view(np.float64)),
                                          op.fortran_llocal()))

    def layout(self) -> np.ndarray:
        """Memory bit of each site (identity unless the planner chose another layout)."""
        out = np.zeros(self.n, dtype=np.int32)
        check(lib().pned_ctx_layout(self.h, out))
        return out

    def set_layout(self, site_bit=None):
        """Force a memory layout (before any vector of the context exists)."""
        if site_bit is None:
            check(lib().pned_ctx_set_layout(self.h, None))
        else:
            sb = np.ascontiguousarray(site_bit, dtype=np.int32)
            check(lib().pned_ctx_set_layout(self.h, sb.ctypes.data))

    def tune(self, tile_bits=0, low_bits=0, threads=0):
        check(lib().pned_ctx_tune(self.h, tile_bits, low_bits, threads))

    def autotune(self, v1: "Vector", v0: "Vector", reps: int = 2):
        """Pick the fastest tile plan on these vectors (v0 is overwritten)."""
        k, l, t, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        check(lib().pned_ctx_autotune(self.h, v1.h, v0.h, reps, C.byref(k), C.byref(l), C.byref(t), C.byref(ms)))
        return {"tile_bits": k.value, "low_bits": l.value, "threads": t.value, "ms": ms.value}

    def bisec(self, alpha, beta, ne, eps=1e-25):
        """pned_bisec's eigenvalues, computed on this context's GPU (the same doubles)."""
        a = np.ascontiguousarray(alpha, dtype=float)
        b = np.ascontiguousarray(beta, dtype=float)
        E = np.zeros(max(ne, 1))
        check(lib().pned_ctx_bisec(self.h, a, b, len(a), E, ne, eps))
        return E[:ne]

    def vec12(self, alpha, beta, E, nvec):
        """pned_vec12's eigenvectors, computed on this context's GPU (the same doubles)."""
        a = np.ascontiguousarray(alpha, dtype=float)
        b = np.ascontiguousarray(beta, dtype=float)
        v = np.zeros(len(a) * nvec)
        check(lib().pned_ctx_vec12(self.h, a, b, len(a), np.ascontiguousarray(E, dtype=float), nvec, v))
        return v.reshape(nvec, len(a))

    def graphs(self, enable: bool):
        """HIP graph replay of the one-column S=1/2 Lanczos steps (lnc1, dsflnc1z); on by default."""
        check(lib().pned_ctx_graphs(self.h, 1 if enable else 0))

    def spin_kernel(self, tiled: bool):
        check(lib().pned_ctx_spin_kernel(self.h, 1 if tiled else 0))

    def dump_vectors(self, path, indx: int, vecs):
        """vec_q*/vec_me* file: indx, then each (device) vector as one record."""
        arr = (C.c_void_p * len(vecs))(*[v.h for v in vecs])
        check(lib().pned_dump_vectors(str(path).encode(), int(indx), len(vecs), arr))

    def read_vectors(self, path, vecs) -> int:
        arr = (C.c_void_p * len(vecs))(*[v.h for v in vecs])
        ix = C.c_int()
        check(lib().pned_read_vectors(str(path).encode(), C.byref(ix), len(vecs), arr))
        return ix.value

    def basis_write(self, path, e: float, x: "Vector"):
        check(lib().pned_basis_write(str(path).encode(), float(e), x.h))

    def basis_read(self, path, x: "Vector") -> float:
        e = C.c_double()
        check(lib().pned_basis_read(str(path).encode(), C.byref(e), x.h))
        return e.value

    def interval_dump(self, path, data_size: int, vec_size: int, reals=(), vecs=()):
        """Checkpoint vecdump_me* (interval_dump): data_size, vec_size, the given real
        arrays (sa, sb, sc, se that the caller dumps, in that order), the given vectors."""
        arrs = [np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, order="F")) for a in reals]
        ptrs = (C.c_void_p * max(len(arrs), 1))(*[a.ctypes.data for a in arrs])
        cnts = (C.c_int64 * max(len(arrs), 1))(*[a.size for a in arrs])
        vv = (C.c_void_p * max(len(vecs), 1))(*[v.h for v in vecs])
        check(lib().pned_interval_dump(str(path).encode(), int(data_size), int(vec_size), len(arrs), ptrs, cnts,
                                       len(vecs), vv))

    def interval_dump_read(self, path, real_counts=(), vecs=()):
        """interval_dump_read: returns (data_size, vec_size, [real arrays]); vectors filled."""
        arrs = [np.zeros(int(c)) for c in real_counts]
        ptrs = (C.c_void_p * max(len(arrs), 1))(*[a.ctypes.data for a in arrs])
        cnts = (C.c_int64 * max(len(arrs), 1))(*[a.size for a in arrs])
        vv = (C.c_void_p * max(len(vecs), 1))(*[v.h for v in vecs])
        ds, vs = C.c_int(), C.c_int()
        check(lib().pned_interval_dump_read(str(path).encode(), C.byref(ds), C.byref(vs), len(arrs), ptrs, cnts,
                                            len(vecs), vv))
        return ds.value, vs.value, arrs

    def sz_sector(self, n_down: int):
        """Restrict to the S^z sector with n_down set bits (n_down < 0: full space);
        vectors must be created afterwards (the dimension changes)."""
        check(lib().pned_ctx_sz_sector(self.h, n_down))
        self.dim = int(lib().pned_ctx_dim(self.h))

    def s12_kernel(self, kind: int):
        """S=1/2 pass kernel: 1 term-major (default), 0 state-major."""
        check(lib().pned_ctx_s12_kernel(self.h, kind))

    def plan_info(self):
        n = C.c_int()
        t = np.zeros(64, dtype=np.int32)
        check(lib().pned_ctx_plan_info(self.h, C.byref(n), t, 64))
        return n.value, list(t[: min(n.value, 64)])

    def sync(self):
        check(lib().pned_ctx_sync(self.h))

    def stream(self) -> int:
        return lib().pned_ctx_stream(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().pned_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- parallelized_diag ------------------------------------------------------------
    def set_initial_vec(self, v: Vector, seed: int):
        check(lib().pned_set_initial_vec(self.h, v.h, seed))

    def p_mltply_v3(self, v1: Vector, v0: Vector, eperbd: float = 0.0):
        pr = np.zeros(2 * v1.nvec)
        check(lib().pned_p_mltply_v3(self.h, v1.h, v0.h, eperbd, pr))
        z = pr[0::2] + 1j * pr[1::2]
        return z[0] if v1.nvec == 1 else z

    def p_mltply_v3_noalloc(self, v1: Vector, v0: Vector, eperbd: float, prdct_dev: int):
        """Allocation-free H.v: <v1|(H - shift)|v1> lands in device memory at prdct_dev
        (2 doubles per column); nothing is synchronised."""
        check(lib().pned_p_mltply_v3_noalloc(self.h, v1.h, v0.h, eperbd, C.c_void_p(prdct_dev)))

    def hv_fresh(self, v1: Vector, v0: Vector, eperbd: float = 0.0):
        pr = np.zeros(2 * v1.nvec)
        check(lib().pned_hv_fresh(self.h, v1.h, v0.h, eperbd, pr))
        z = pr[0::2] + 1j * pr[1::2]
        return z[0] if v1.nvec == 1 else z

    def calc_vec_inprdct(self, x: Vector, y: Vector):
        out = np.zeros(2 * x.nvec)
        check(lib().pned_calc_vec_inprdct(self.h, x.h, y.h, out))
        z = out[0::2] + 1j * out[1::2]
        return z[0] if x.nvec == 1 else z

    def calc_alpha_beta(self, v1: Vector, v0: Vector):
        a = np.zeros(v1.nvec)
        b = np.zeros(v1.nvec)
        check(lib().pned_calc_alpha_beta(self.h, v1.h, v0.h, a, b))
        return (a[0], b[0]) if v1.nvec == 1 else (a, b)

    def calc_alpha_beta2(self, v1: Vector, v0: Vector):
        a = np.zeros(v1.nvec)
        b = np.zeros(v1.nvec)
        check(lib().pned_calc_alpha_beta2(self.h, v1.h, v0.h, a, b))
        return (a[0], b[0]) if v1.nvec == 1 else (a, b)

    def axpby(self, a, x: Vector, b, y: Vector):
        a = np.broadcast_to(np.asarray(a, dtype=complex), (x.nvec,))
        b = np.broadcast_to(np.asarray(b, dtype=complex), (x.nvec,))
        check(lib().pned_axpby(self.h, np.ascontiguousarray(a.view(np.float64)), x.h,
                               np.ascontiguousarray(b.view(np.float64)), y.h))

    def scale(self, x: Vector, a):
        a = np.broadcast_to(np.asarray(a, dtype=complex), (x.nvec,))
        check(lib().pned_scale(self.h, x.h, np.ascontiguousarray(a.view(np.float64))))

    def lanczos_update(self, v1: Vector, v0: Vector, alpha, beta):
        check(lib().pned_lanczos_update(self.h, v1.h, v0.h, np.atleast_1d(np.asarray(alpha, float)),
                                        np.atleast_1d(np.asarray(beta, float))))

    # -- Krylov -----------------------------------------------------------------------
    def lnc1(self, v1: Vector, v0: Vector, max_itr=1024, nvec=8, fixed_steps=0, allow_noconv=False):
        steps = fixed_steps if fixed_steps > 0 else max_itr
        alpha = np.zeros(steps + 1)
        beta = np.zeros(steps + 1)
        ene = np.zeros(max(nvec, 1))
        itr = C.c_int()
        rc = lib().pned_lnc1(self.h, v1.h, v0.h, max_itr, nvec, fixed_steps, alpha, beta, ene, C.byref(itr))
        if rc == E_NOCONV and allow_noconv:
            rc = 0
        check(rc)
        return itr.value, alpha[: itr.value], beta[: itr.value], ene

    def lnc1_span(self, v1: Vector, v0: Vector, alpha, beta, first=1, stop=0, max_itr=1024, nvec=8, fixed_steps=0,
                  allow_noconv=False):
        """lnc1z from step ``first`` with the reference's explicit update, into the caller's
        alpha/beta arrays (steps 1..first-1 already there when first > 2); returns (itr, ene).
        stop > 0: return at the top of step ``stop`` (the checkpoint point)."""
        ene = np.zeros(max(nvec, 1))
        itr = C.c_int()
        rc = lib().pned_lnc1_span(self.h, v1.h, v0.h, first, stop, max_itr, nvec, fixed_steps, alpha, beta, ene,
                                  C.byref(itr))
        if rc == E_NOCONV and allow_noconv:
            rc = 0
        check(rc)
        return itr.value, ene

    def tps_span(self, v1: Vector, v0: Vector, vg: Vector, vt: Vector, lmax: float, alpha, beta, coef, first=1,
                 stop=0):
        """pned_tps from step ``first`` into the caller's flat arrays (nsteps*nvec,
        nsteps*nvec, nsteps*nvec*ncoef); stop > 0 returns at the top of step ``stop``."""
        nc = v1.nvec
        nsteps = len(alpha) // nc
        check(lib().pned_tps_span(self.h, v1.h, v0.h, vg.h, vt.h, lmax, first, stop, nsteps, alpha, beta, coef))

    def dsflnc1z_span(self, v1: Vector, v0: Vector, cnorm: float, salpha, sbeta, first=1, stop=0):
        """dsflnc1z from step ``first`` into the caller's salpha (s_maxitr) and sbeta
        (s_maxitr + 1); returns the last step taken."""
        icf = C.c_int()
        check(lib().pned_dsflnc1z_span(self.h, v1.h, v0.h, cnorm, first, stop, len(salpha), salpha, sbeta,
                                       C.byref(icf)))
        return icf.value

    def lnc1_reorth(self, basis, w: Vector, nvec=1, fixed_steps=0, allow_noconv=False):
        """Lanczos with full reorthogonalisation over the given basis vectors (basis[0] =
        normalised start); returns (itr, alpha, beta, ene)."""
        nb = len(basis)
        arr = (C.c_void_p * nb)(*[v.h for v in basis])
        alpha, beta, ene = np.zeros(nb), np.zeros(nb), np.zeros(max(nvec, 1))
        itr = C.c_int()
        rc = lib().pned_lnc1_reorth(self.h, arr, nb, w.h, nvec, fixed_steps, alpha, beta, ene, C.byref(itr))
        if rc == E_NOCONV and allow_noconv:
            rc = 0
        check(rc)
        return itr.value, alpha[: itr.value], beta[: itr.value], ene

    def dsflnc1z_reorth(self, basis, w: Vector, cnorm: float, s_maxitr: int):
        nb = len(basis)
        arr = (C.c_void_p * nb)(*[v.h for v in basis])
        sa, sb = np.zeros(s_maxitr), np.zeros(s_maxitr)
        check(lib().pned_dsflnc1z_reorth(self.h, arr, nb, w.h, cnorm, s_maxitr, sa, sb))
        return sa, sb

    def ritz_vector(self, v1: Vector, v0: Vector, x: Vector, alpha, beta, coef):
        n = len(coef)
        check(lib().pned_ritz_vector(self.h, v1.h, v0.h, x.h, n, np.ascontiguousarray(alpha, dtype=float),
                                     np.ascontiguousarray(beta, dtype=float), np.ascontiguousarray(coef, dtype=float)))

    def inv1(self, eig: float, x: Vector, b: Vector, p: Vector, r: Vector, y: Vector, max_inv=5000, max_cg=500,
             reference_saved_loop=True):
        it = C.c_int()
        check(lib().pned_inv1(self.h, eig, x.h, b.h, p.h, r.h, y.h, max_inv, max_cg,
                              1 if reference_saved_loop else 0, C.byref(it)))
        return it.value

    def cocg(self, e0: float, eta: float, b: Vector, x: Vector, r: Vector, p: Vector, y: Vector,
             max_it=5000, tol=1e-10):
        it = C.c_int()
        res = C.c_double()
        check(lib().pned_cocg(self.h, e0, eta, b.h, x.h, r.h, p.h, y.h, max_it, tol, C.byref(it), C.byref(res)))
        return it.value, res.value

    def check1(self, x: Vector, work: Vector, max_ratios=32):
        ex = np.zeros(2)
        ra = np.zeros(2 * max_ratios)
        k = C.c_int()
        check(lib().pned_check1(self.h, x.h, work.h, ex, ra, max_ratios, C.byref(k)))
        return complex(ex[0], ex[1]), (ra[0::2] + 1j * ra[1::2])[: k.value]

    def residual(self, x: Vector, e: float, work: Vector) -> float:
        r = C.c_double()
        check(lib().pned_residual(self.h, x.h, e, work.h, C.byref(r)))
        return r.value

    def tps_set_observables(self, site_vec, sq_q):
        sq = np.ascontiguousarray(np.asarray(sq_q, dtype=float).reshape(-1))
        check(lib().pned_tps_set_observables(self.h, np.ascontiguousarray(np.asarray(site_vec, float).reshape(-1)),
                                             len(sq) // 2, sq if len(sq) else np.zeros(2)))
        self._nsq = len(sq) // 2

    def tps(self, v1: Vector, v0: Vector, vg: Vector, vt: Vector, lmax: float, nsteps: int):
        nc = v1.nvec
        ncoef = 7 + getattr(self, "_nsq", 0)
        alpha = np.zeros(nsteps * nc)
        beta = np.zeros(nsteps * nc)
        coef = np.zeros(nsteps * nc * ncoef)
        check(lib().pned_tps(self.h, v1.h, v0.h, vg.h, vt.h, lmax, nsteps, alpha, beta, coef))
        return alpha.reshape(nsteps, nc), beta.reshape(nsteps, nc), coef.reshape(nsteps, nc, ncoef)

    def tps_fused(self, v1: Vector, va: Vector, vb: Vector, lmax: float, nsteps: int, graph_steps=None):
        """Fused mTPQ; graph_steps (a multiple of 3, or 0 for 30) replays the steps from a
        captured HIP graph (pned_tps_fused_graph), None launches them one by one."""
        nc = v1.nvec
        ncoef = 7 + getattr(self, "_nsq", 0)
        alpha = np.zeros(nsteps * nc)
        beta = np.zeros(nsteps * nc)
        coef = np.zeros(nsteps * nc * ncoef)
        if graph_steps is None:
            check(lib().pned_tps_fused(self.h, v1.h, va.h, vb.h, lmax, nsteps, alpha, beta, coef))
        else:
            check(lib().pned_tps_fused_graph(self.h, v1.h, va.h, vb.h, lmax, nsteps, int(graph_steps), alpha, beta,
                                             coef))
        return alpha.reshape(nsteps, nc), beta.reshape(nsteps, nc), coef.reshape(nsteps, nc, ncoef)

    def intalv_zz(self, x: Vector, v: Vector, q, site_vec) -> float:
        cn = C.c_double()
        check(lib().pned_intalv_zz(self.h, x.h, v.h, np.asarray(q, dtype=float),
                                   np.ascontiguousarray(np.asarray(site_vec, float).reshape(-1)), C.byref(cn)))
        return cn.value

    def intalv_xy(self, x: Vector, v: Vector, q, site_vec, ipm: int = 1) -> float:
        """intalv_xy: v (of this context) = normalised S^{ipm}(q) x; returns cnorm.  Full
        space: x of a context of the same cluster and spin.  S=1/2 sectors: x of a context
        with k set bits, this one with k + ipm."""
        cn = C.c_double()
        check(lib().pned_intalv_xy(self.h, x.h, v.h, np.asarray(q, dtype=float),
                                   np.ascontiguousarray(np.asarray(site_vec, float).reshape(-1)), int(ipm),
                                   C.byref(cn)))
        return cn.value

    def dsflnc1z(self, v1: Vector, v0: Vector, cnorm: float, s_maxitr: int):
        sa = np.zeros(s_maxitr)
        sb = np.zeros(s_maxitr + 1)
        check(lib().pned_dsflnc1z(self.h, v1.h, v0.h, cnorm, s_maxitr, sa, sb))
        return sa, sb

    def renbun(self, salpha, sbeta, ene, eta, i_range=2000, split=2.5e-3, s_maxitr=None, sq=0.0):
        if s_maxitr is None:
            s_maxitr = len(salpha)
        sa = np.zeros(max(s_maxitr, len(salpha)))
        sa[: len(salpha)] = salpha
        sb = np.zeros(len(sa) + 1)
        sb[: len(sbeta)] = sbeta
        w = np.zeros(i_range)
        dsf = np.zeros(i_range)
        conv = np.zeros(i_range, dtype=np.int32)
        check(lib().pned_renbun(self.h, i_range, s_maxitr, split, ene, eta, sa, sb, sq, w, dsf, conv))
        return w, dsf, conv


def s12_layout(op):
    """The memory layout a new context of this S=1/2 operator gets: site k -> memory bit
    (no GPU needed)."""
    out = np.zeros(op.n, dtype=np.int32)
    check(lib().pned_s12_layout_choose(op.n, op.ibond, np.ascontiguousarray(op.ipair, dtype=np.int32),
                                       np.ascontiguousarray(op.ibond_type, dtype=np.int32),
                                       np.ascontiguousarray(op.fortran_xlocal().view(np.float64)),
                                       op.fortran_llocal(), out))
    return out


def s12_plan(op, colbits=0, tile_bits=12, low_bits=4, site_bit=None):
    """The planner's decomposition of an S=1/2 operator (no GPU needed); site_bit: the
    memory layout (site k at memory bit site_bit[k]), None = identity."""
    mp, mt = 64, 4096
    sb = None if site_bit is None else np.ascontiguousarray(site_bit, dtype=np.int32)
    npass, nterm, ngroup = C.c_int(), C.c_int(), C.c_int()
    pinfo = np.zeros(4 * mp, dtype=np.int32)
    pbits = np.zeros(40 * mp, dtype=np.int32)
    terms = np.zeros(6 * mt)
    gmask = np.zeros(mt, dtype=np.uint32)
    diag = np.zeros(5 * 64 + 2)
    check(lib().pned_s12_plan_export(op.n, op.ibond, np.ascontiguousarray(op.ipair, dtype=np.int32),
                                     np.ascontiguousarray(op.ibond_type, dtype=np.int32),
                                     np.ascontiguousarray(op.fortran_xlocal().view(np.float64)),
                                     op.fortran_llocal(), colbits, tile_bits, low_bits, C.byref(npass),
                                     C.byref(nterm), C.byref(ngroup), pinfo, pbits, terms, gmask, diag, mp, mt,
                                     None if sb is None else sb.ctypes.data))
    P, T, G = npass.value, nterm.value, ngroup.value
    passes = []
    for p in range(P):
        K, ft, nt, hd = pinfo[4 * p: 4 * p + 4]
        passes.append({"K": int(K), "first": int(ft), "n": int(nt), "diag": bool(hd),
                       "tb": [int(b) for b in pbits[40 * p: 40 * p + K]]})
    tt = terms[: 6 * T].reshape(T, 6)
    return {"passes": passes, "tmask": tt[:, 0].astype(np.int64), "selmask": tt[:, 1].astype(np.int64),
            "c0": tt[:, 2] + 1j * tt[:, 3], "c1": tt[:, 4] + 1j * tt[:, 5], "gmask": gmask[:T].astype(np.int64),
            "diag": diag[: 5 * G].reshape(G, 5), "c0diag": diag[5 * G] + 1j * diag[5 * G + 1]}


def spin_plan(op, tile_sites=0, low_sites=0):
    """The tiled spin-S kernel's passes for an operator with q = 3..5 (no GPU needed): per
    pass its tile sites (lowest first) and off-diagonal classes (bond, digit shift of the
    min and max site, coefficient by output local state); the pass of each bond; the forced
    low sites."""
    mp, qq = 64, op.q * op.q
    row = 3 + 2 * qq
    npass, low = C.c_int(), C.c_int()
    sites = np.zeros(32 * mp, dtype=np.int32)
    bond_pass = np.full(op.ibond, -1, dtype=np.int32)
    cls = np.zeros(mp * 64 * row)
    ncls = np.zeros(mp, dtype=np.int32)
    check(lib().pned_spin_plan_export(op.n, op.q, op.ibond, np.ascontiguousarray(op.ipair, dtype=np.int32),
                                      np.ascontiguousarray(op.ibond_type, dtype=np.int32),
                                      np.ascontiguousarray(op.fortran_xlocal().view(np.float64)),
                                      op.fortran_llocal(), int(tile_sites), int(low_sites), C.byref(npass),
                                      C.byref(low), sites, bond_pass, cls, ncls, mp))
    passes = []
    for p in range(npass.value):
        st = [int(x) for x in sites[32 * p: 32 * p + 32] if x >= 0]
        rows = cls[p * 64 * row: (p * 64 + int(ncls[p])) * row].reshape(-1, row)
        passes.append({"sites": st, "bonds": [int(b) for b in np.flatnonzero(bond_pass == p)],
                       "classes": [{"bond": int(r[0]), "da": int(r[1]), "db": int(r[2]),
                                    "coef": r[3::2] + 1j * r[4::2]} for r in rows]})
    return {"passes": passes, "low_sites": low.value, "bond_pass": bond_pass}


def tps_arrays(alpha, beta, coef, nvec, ncoef):
    """Reshape `Context.tps` / `tps_fused` outputs (step-major, columns interleaved) into
    the reference's per-sample arrays: alpha, beta (samples, steps), coef (samples,
    ncoef, steps) -- what SalphaSbeta files hold per seed."""
    a = np.asarray(alpha, dtype=float).reshape(-1, nvec).T.copy()
    b = np.asarray(beta, dtype=float).reshape(-1, nvec).T.copy()
    c = np.asarray(coef, dtype=float).reshape(-1, nvec, ncoef).transpose(1, 2, 0).copy()
    return a, b, c


def ctpq(alpha, beta, coef, t_min, t_max, num_div):
    """cTPQ evaluation (pned_ctpq_eval): alpha, beta (samples, steps), coef (samples,
    ncoef, steps).  Returns (num_div+1, 4): T, C, variance term, dC/dT."""
    a = np.ascontiguousarray(alpha, dtype=float)
    b = np.ascontiguousarray(beta, dtype=float)
    c = np.ascontiguousarray(coef, dtype=float)
    seeds, nd1 = b.shape
    out = np.zeros(4 * (num_div + 1))
    check(lib().pned_ctpq_eval(nd1, c.shape[1], seeds, a.reshape(-1), b.reshape(-1), c.reshape(-1),
                               float(t_min), float(t_max), int(num_div), out))
    return out.reshape(num_div + 1, 4)


def io_subrecord_limit(nbytes: int = 0):
    """Largest subrecord the file writers emit (0 = 2^31-9, the 4-byte marker limit)."""
    check(lib().pned_io_subrecord_limit(int(nbytes)))


def dump_alpha_beta_coef(path, alpha, beta, coef):
    """SalphaSbeta_me*seed* file of an mTPQ sample: alpha, beta (nd1,), coef (nd1, nd2)."""
    a = np.ascontiguousarray(alpha, dtype=float)
    b = np.ascontiguousarray(beta, dtype=float)
    c = np.asarray(coef, dtype=float)
    check(lib().pned_dump_alpha_beta_coef(str(path).encode(), len(a), c.shape[1], a, b, np.ascontiguousarray(c.T)))


def read_alpha_beta_coef(path):
    n1, n2 = C.c_int(), C.c_int()
    check(lib().pned_read_alpha_beta_coef(str(path).encode(), C.byref(n1), C.byref(n2), None, None, None, 0, 0))
    a, b = np.zeros(n1.value), np.zeros(n1.value)
    c = np.zeros((n2.value, n1.value))
    check(lib().pned_read_alpha_beta_coef(str(path).encode(), C.byref(n1), C.byref(n2), a.ctypes.data,
                                          b.ctypes.data, c.ctypes.data, n1.value, n2.value))
    return a, b, c.T.copy()


def dump_dsf_alpha_beta(path, sq, salpha, sbeta):
    """SalphaSbeta_q*me* file of a DSF run: sq (complex), salpha, sbeta."""
    a = np.ascontiguousarray(salpha, dtype=float)
    b = np.ascontiguousarray(sbeta, dtype=float)
    s = np.array([complex(sq).real, complex(sq).imag])
    check(lib().pned_dump_dsf_alpha_beta(str(path).encode(), s, len(a), a, len(b), b))


def read_dsf_alpha_beta(path, n_alpha, n_beta):
    """Returns (sq as a complex number, salpha, sbeta)."""
    s = np.zeros(2)
    a, b = np.zeros(n_alpha), np.zeros(n_beta)
    check(lib().pned_read_dsf_alpha_beta(str(path).encode(), s, n_alpha, a, n_beta, b))
    return complex(s[0], s[1]), a, b


def dump_info_write(path, dump_name, np_, position, max_itr, dsf_comp, s_maxitr, n, ibond, nvec, total_sz,
                    itr_counter, true_value=1):
    """The checkpoint's .info file (interval_dump); logicals written as true_value."""
    pos = np.ascontiguousarray(np.asarray(position, dtype=np.int32))
    itc = np.ascontiguousarray(np.asarray(itr_counter, dtype=np.int32))
    assert itc.size == 12
    check(lib().pned_dump_info_write(str(path).encode(), dump_name.encode(), np_, pos.size, pos, true_value, max_itr,
                                     dsf_comp.encode(), s_maxitr, n, ibond, nvec, total_sz, itc))


def dump_info_read(path, npos=7):
    """check_interval_dump's read of the .info file: a dict of its fields."""
    name = C.create_string_buffer(129)
    comp = C.create_string_buffer(3)
    pos = np.zeros(npos, dtype=np.int32)
    itc = np.zeros(12, dtype=np.int32)
    v = [C.c_int() for _ in range(7)]
    check(lib().pned_dump_info_read(str(path).encode(), name, C.byref(v[0]), npos, pos, C.byref(v[1]), comp,
                                    C.byref(v[2]), C.byref(v[3]), C.byref(v[4]), C.byref(v[5]), C.byref(v[6]), itc))
    return {"dump_name": name.value.decode(), "np": v[0].value, "position": [bool(p) for p in pos],
            "max_itr": v[1].value, "dsf_comp": comp.value.decode(), "s_maxitr": v[2].value, "n": v[3].value,
            "ibond": v[4].value, "nvec": v[5].value, "total_sz": v[6].value, "itr_counter": itc.tolist()}


def write_dsf_res(path, sq, q, w, dsf, conv, append=True):
    """One momentum's block of the dsf_res_<comp>.dat* text file (write_dsf_res)."""
    w = np.ascontiguousarray(w, dtype=np.float64)
    d = np.ascontiguousarray(dsf, dtype=np.float64)
    c = np.ascontiguousarray(conv, dtype=np.int32)
    check(lib().pned_write_dsf_res(str(path).encode(), 1 if append else 0, float(sq), np.asarray(q, dtype=float),
                                   len(w), w, d, c))


def bisec(alpha, beta, ne, eps=1e-25):
    a = np.ascontiguousarray(alpha, dtype=float)
    b = np.ascontiguousarray(beta, dtype=float)
    E = np.zeros(max(ne, 1))
    check(lib().pned_bisec(a, b, len(a), E, ne, eps))
    return E[:ne]


def vec12(alpha, beta, E, nvec):
    a = np.ascontiguousarray(alpha, dtype=float)
    b = np.ascontiguousarray(beta, dtype=float)
    v = np.zeros(len(a) * nvec)
    check(lib().pned_vec12(a, b, len(a), np.ascontiguousarray(E, dtype=float), nvec, v))
    return v.reshape(nvec, len(a))
