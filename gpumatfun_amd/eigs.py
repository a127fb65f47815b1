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


def _csr_arrays(A):
    A = sp.csr_matrix(A, dtype=np.float64)
    A.sort_indices()
    return (A.shape[0], np.ascontiguousarray(A.indptr, dtype=np.int32),
            np.ascontiguousarray(A.indices, dtype=np.int32),
            np.ascontiguousarray(A.data, dtype=np.float64))


_ipt = C.POINTER(C.c_int32)
_dpt = C.POINTER(C.c_double)


def nd_order(A):
    """Nested-dissection order of A's symmetric pattern (``gm_nd_order``):
    ``perm[q]`` is the old index placed at q."""
    n, ptr, ind, _ = _csr_arrays(A)
    perm = np.empty(n, dtype=np.int32)
    _lib.check(_lib.load().gm_nd_order(n, ptr.ctypes.data_as(_ipt), ind.ctypes.data_as(_ipt),
                                       perm.ctypes.data_as(_ipt)), chol=True)
    return perm.astype(np.int64)


class SparseCholesky:
    """Cholesky factor of A - sigma I for symmetric positive definite A - sigma I
    (``gm_chol_factor``: nested dissection, supernodal multifrontal; host
    threads).  ``solve(B)`` returns (A - sigma I)^{-1} B."""

    def __init__(self, A, sigma=0.0, threads=0):
        n, ptr, ind, val = _csr_arrays(A)
        self.n = n
        h = C.c_void_p()
        _lib.check(_lib.load().gm_chol_factor(n, ptr.ctypes.data_as(_ipt),
                                              ind.ctypes.data_as(_ipt),
                                              val.ctypes.data_as(_dpt), float(sigma),
                                              int(threads), C.byref(h)), chol=True)
        self._h = h

    def solve(self, B):
        B = np.asarray(B, dtype=np.float64)
        vec = B.ndim == 1
        X = np.array(B.reshape(self.n, -1), dtype=np.float64, order="F", copy=True)
        _lib.check(_lib.load().gm_chol_solve(self._h, X.shape[1], X.ctypes.data_as(_dpt)),
                   chol=True)
        return X[:, 0] if vec else X

    def info(self):
        nnz = C.c_int64(0)
        fl = C.c_double(0.0)
        ns = C.c_int32(0)
        secs = (C.c_double * 3)()
        _lib.check(_lib.load().gm_chol_info(self._h, C.byref(nnz), C.byref(fl), C.byref(ns),
                                            secs), chol=True)
        return dict(nnz_l=nnz.value, flops=fl.value, supernodes=ns.value,
                    order_s=secs[0], symbolic_s=secs[1], numeric_s=secs[2])

    def perm(self):
        p = np.empty(self.n, dtype=np.int32)
        _lib.check(_lib.load().gm_chol_perm(self._h, p.ctypes.data_as(_ipt)), chol=True)
        return p.astype(np.int64)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().gm_chol_delete(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_shift(A):
    """A shift just below the spectrum of the positive semidefinite A:
    -1e-10 times its Gershgorin bound (the largest absolute row sum).  It is
    far above the factorisation's rounding (~1e-16 ||A||), so A - sigma I
    stays positive definite, and close enough to 0 that the smallest
    eigenvalues are well separated in (A - sigma I)^{-1}: on the 1M-node disc
    sigma = -1 took 15 block solves where sigma = -200 took 33."""
    A = sp.csr_matrix(A)
    bound = float(np.max(np.asarray(abs(A).sum(axis=1)).ravel()))
    return -1e-10 * max(bound, np.finfo(float).tiny)


def eigs_shift_invert(A, k, *, sigma=None, block_size=None, max_blocks=12, max_restarts=4,
                      tol=1e-14, seed=0, threads=0, want_vectors=True, info=False):
    """(lambda, Q) of the k smallest eigenpairs of the symmetric A by block
    Lanczos on (A - sigma I)^{-1} through ``gm_eigs_shift_invert`` (host
    factorisation; what ``eigs(K, M, k, 'sm')`` does in process_nonuniform.m).

    ``sigma`` must lie below lambda_1 (default ``default_shift(A)``);
    ``block_size`` defaults to k + 4 (at least 8); every returned pair has
    ``||A q - l q|| <= tol ||A||_inf``.  With ``info`` also returns
    dict(blocks=, max_residual=)."""
    n, ptr, ind, val = _csr_arrays(A)
    if sigma is None:
        sigma = default_shift(A)
    if block_size is None:
        block_size = min(max(k + 4, 8), n // 2)
    d = _lib.GmSieigsDesc()
    d.n = n
    d.nnz = int(ptr[-1])
    d.A_values = val.ctypes.data_as(_dpt)
    d.A_ptr = ptr.ctypes.data_as(_ipt)
    d.A_ind = ind.ctypes.data_as(_ipt)
    d.num_eigs = k
    d.block_size = block_size
    d.max_blocks = max_blocks
    d.max_restarts = max_restarts
    d.tol = tol
    d.sigma = float(sigma)
    d.seed = seed
    d.threads = threads
    lam = np.empty(k)
    Q = np.empty((n, k), order="F") if want_vectors else None
    blocks = C.c_int32(0)
    res = C.c_double(0.0)
    st = _lib.load().gm_eigs_shift_invert(C.byref(d), lam.ctypes.data_as(_dpt),
                                          Q.ctypes.data_as(_dpt) if Q is not None else None,
                                          C.byref(blocks), C.byref(res))
    _lib.check(st, chol=True)
    out = (lam, Q)
    if info:
        out = out + (dict(blocks=blocks.value, max_residual=res.value),)
    return out


def lambda_max(A, *, tol=1e-12, max_iter=300, seed=0, threads=0, info=False):
    """lambda_max of the symmetric A by host Lanczos (``gm_lambda_max``; the
    ``eigs(K, M, 1, 'lm')`` of process_nonuniform.m:40)."""
    n, ptr, ind, val = _csr_arrays(A)
    out = C.c_double(0.0)
    it = C.c_int32(0)
    _lib.check(_lib.load().gm_lambda_max(n, ptr.ctypes.data_as(_ipt), ind.ctypes.data_as(_ipt),
                                         val.ctypes.data_as(_dpt), tol, max_iter, seed, threads,
                                         C.byref(out), C.byref(it)), chol=True)
    return (out.value, dict(iterations=it.value)) if info else out.value
