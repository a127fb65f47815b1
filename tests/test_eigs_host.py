"""Host parts of the deflation eigensolver (no GPU): the Jacobi routine that
diagonalises the block's Gram and Rayleigh-quotient matrices, against numpy's
LAPACK eigh, and the descriptor checks that run before any device call."""

import ctypes as C

import numpy as np
import pytest

from gpumatfun_amd import _lib
from gpumatfun_amd.eigs import sym_eig


@pytest.mark.parametrize("n,seed", [(1, 0), (5, 1), (16, 2), (64, 3)])
def test_sym_eig_matches_lapack(n, seed):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((n, n))
    A = B + B.T
    w, V = sym_eig(A)
    w_ref = np.linalg.eigvalsh(A)
    scale = np.abs(w_ref).max()
    assert np.abs(w - w_ref).max() <= 1e-13 * scale
    assert np.abs(V.T @ V - np.eye(n)).max() <= 1e-13
    assert np.abs(A @ V - V * w).max() <= 1e-12 * scale


def test_sym_eig_graded_and_repeated_spectrum():
    """A Gram matrix of a badly scaled block (eigenvalues 1 .. 1e-12) and a
    repeated eigenvalue: the smallest come out with absolute error at
    rounding of the largest, the eigenvectors orthonormal."""
    rng = np.random.default_rng(7)
    Qm, _ = np.linalg.qr(rng.standard_normal((32, 32)))
    d = np.logspace(0, -12, 32)
    d[5] = d[6]
    A = (Qm * d) @ Qm.T
    A = 0.5 * (A + A.T)
    w, V = sym_eig(A)
    assert np.abs(w - np.sort(d)).max() <= 1e-14
    assert np.abs(V.T @ V - np.eye(32)).max() <= 1e-13


def test_sym_eig_diagonal_and_closed_form():
    """diag input: no rotation, sorted; 2 x 2 [[a, b], [b, a]]: a -+ b with
    eigenvectors (1, -+1)/sqrt(2)."""
    w, V = sym_eig(np.diag([3.0, -1.0, 2.0]))
    assert np.array_equal(w, [-1.0, 2.0, 3.0])
    assert np.array_equal(np.abs(V), np.eye(3)[:, [1, 2, 0]])
    w, V = sym_eig(np.array([[2.0, 0.5], [0.5, 2.0]]))
    assert np.allclose(w, [1.5, 2.5], rtol=0, atol=1e-15)
    assert np.allclose(np.abs(V), np.sqrt(0.5), rtol=0, atol=1e-15)
    assert V[0, 0] * V[1, 0] < 0 and V[0, 1] * V[1, 1] > 0


def test_eigs_rejects_bad_descriptors():
    lib = _lib.load()
    dp = C.POINTER(C.c_double)
    lam = np.zeros(4)
    assert lib.gm_eigs_smallest(None, lam.ctypes.data_as(dp), None, None, None) == _lib.GM_EINVAL
    ptr = np.array([0, 1, 2, 3], dtype=np.int32)
    ind = np.array([0, 1, 7], dtype=np.int32)
    val = np.ones(3)
    d = _lib.GmEigsDesc()
    d.n, d.nnz, d.sparse_format = 3, 3, _lib.GM_SPARSE_CSR
    d.A_values = val.ctypes.data_as(dp)
    d.A_ptr = ptr.ctypes.data_as(C.POINTER(C.c_int32))
    d.A_ind = ind.ctypes.data_as(C.POINTER(C.c_int32))
    d.num_eigs, d.block_size, d.degree, d.max_passes, d.tol = 2, 16, 10, 10, 1e-8
    # block larger than n
    assert lib.gm_eigs_smallest(C.byref(d), lam.ctypes.data_as(dp), None, None, None) == _lib.GM_EINVAL
    assert b"block_size" in lib.gm_eigs_last_error()
    d.block_size = 24
    assert lib.gm_eigs_smallest(C.byref(d), lam.ctypes.data_as(dp), None, None, None) == _lib.GM_EINVAL
    assert b"16, 32 or 64" in lib.gm_eigs_last_error()
