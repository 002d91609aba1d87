"""MI355X-native matrix-free H.v and Krylov solvers for S=1/2 and spin-S
Kitaev-Heisenberg clusters (the hot path of pNED), behind the C ABI in
include/pned_gpu.h.

* :mod:`pned_amd.lattice` / :mod:`pned_amd.model` -- the reference's cluster and
  bond-matrix tables (host side, what the reference's data_mem module builds);
* :mod:`pned_amd.gpu` -- ctypes front of libpned_gpu.so (loaded on first use).
"""
from . import gpu, lattice, model  # noqa: F401

__all__ = ["lattice", "model", "gpu"]
