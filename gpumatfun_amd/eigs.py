"""The K smallest eigenpairs of a symmetric sparse A on the GPU
(``gm_eigs_smallest``): the deflation inputs Q, Lambda that the reference
computes on the host before ``lanczos_create`` (``eigs(K, M, 'sm')`` in
process_nonuniform.m:106-110; scipy's ARPACK ``eigsh`` in the Python
drivers).  Chebyshev-filtered subspace iteration; suited to operators whose
largest eigenvalue is not many orders above the wanted ones."""

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib


def sym_eig(A):
    """w, V of a small dense symmetric matrix by the library's host Jacobi
    routine (w ascending, V's columns the eigenvectors)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    if A.shape != (n, n):
        raise ValueError("A must be square")
    w = np.empty(n)
    V = np.empty((n, n))
    lib = _lib.load()
    dp = C.POINTER(C.c_double)
    _lib.check(lib.gm_sym_eig(n, A.ctypes.data_as(dp), w.ctypes.data_as(dp),
                              V.ctypes.data_as(dp)), eigs=True)
    return w, V


def eigs_smallest(A, k, *, block_size=None, degree=40, max_passes=200, tol=1e-10,
                  upper_bound=0.0, seed=0, device=0, want_vectors=True, info=False):
    """(lambda, Q) of the k smallest eigenpairs of the symmetric A.

    ``block_size`` (16, 32 or 64; default the smallest one above 2k) sets the
    filtered block; ``degree`` the Chebyshev degree per pass; ``tol`` the
    relative residual ``||A q - l q|| / |l|`` every returned pair meets.
    With ``info`` also returns dict(passes=, max_residual=).  Raises
    GpuMatfunError (GM_EBREAKDOWN) if tol is not met in max_passes."""
    A = sp.csr_matrix(A, dtype=np.float64)
    A.sort_indices()
    n = A.shape[0]
    if block_size is None:
        block_size = next((s for s in (16, 32, 64) if s >= 2 * k), 64)
    keep = {}

    def arr(name, a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        keep[name] = a
        return a.ctypes.data_as(C.POINTER(C.c_double if dtype == np.float64 else C.c_int32))

    d = _lib.GmEigsDesc()
    d.n = n
    d.nnz = A.nnz
    d.sparse_format = _lib.GM_SPARSE_CSR
    d.A_values = arr("val", A.data, np.float64)
    d.A_ptr = arr("ptr", A.indptr, np.int32)
    d.A_ind = arr("ind", A.indices, np.int32)
    d.num_eigs = k
    d.block_size = block_size
    d.degree = degree
    d.max_passes = max_passes
    d.tol = tol
    d.upper_bound = upper_bound
    d.seed = seed
    d.device = device
    lam = np.empty(k)
    Q = np.empty((n, k), order="F") if want_vectors else None
    passes = C.c_int32(0)
    res = C.c_double(0.0)
    dp = C.POINTER(C.c_double)
    lib = _lib.load()
    st = lib.gm_eigs_smallest(C.byref(d), lam.ctypes.data_as(dp),
                              Q.ctypes.data_as(dp) if Q is not None else None,
                              C.byref(passes), C.byref(res))
    _lib.check(st, eigs=True)
    out = (lam, Q)
    if info:
        out = out + (dict(passes=passes.value, max_residual=res.value),)
    return out
