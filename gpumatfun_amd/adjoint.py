"""Differentiable f(A)b: ``FabEngine`` (C ABI ``gm_fab_*``) and a ``torch.autograd.Function``.

u = f(A(kappa)) b with f(x) = exp(-tau x^(alpha/2)), evaluated by Lanczos and
the contour rule of ``problem_startup.m:95-103`` on a fixed contour; the
backward pass returns dL/db, dL/dkappa and dL/dalpha for an upstream
gradient dL/du (see ``include/gpumatfun.h``).  A(kappa) is the
variable-coefficient Laplacian of ``problems.varcoef_laplacian_2d``.
"""

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib
from . import problems, quadrature as q


def _p(a):
    """A double* for a NumPy array or a float64 torch tensor on the GPU (the
    gm_fab_* vector arguments take host or device memory)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        if a.dtype.itemsize != 8 or not a.is_contiguous():
            raise ValueError("device vectors must be contiguous float64")
        return C.cast(C.c_void_p(a.data_ptr()), C.POINTER(C.c_double))
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _vec(x):
    """x as contiguous float64 (a GPU tensor is synchronised with torch's stream,
    since the engine runs on its own)."""
    if hasattr(x, "data_ptr"):
        import torch
        torch.cuda.current_stream().synchronize()
        return x.contiguous()
    return np.ascontiguousarray(x, dtype=np.float64)


def _empty_like(x, n):
    if hasattr(x, "data_ptr"):
        import torch
        return torch.empty(n, dtype=torch.float64, device=x.device)
    return np.empty(n)


def _cplx(z):
    z = np.asarray(z, dtype=np.complex128).ravel()
    out = np.empty(2 * z.size)
    out[0::2] = z.real
    out[1::2] = z.imag
    return out


def contour_weights(xi, dxidt, s, tau, alpha):
    """c_j = s f(xi_j) dxi/dt_j and dc_j/dalpha for f = exp(-tau x^(alpha/2))."""
    xi = np.asarray(xi, dtype=np.complex128)
    p = np.power(xi, alpha / 2.0)
    c = s * np.exp(-tau * p) * dxidt
    dc = c * (-tau * p * np.log(xi) / 2.0)
    return c, dc


class FabEngine:
    """Device-resident u = f(A) b with its adjoint (one matrix, one right-hand side)."""

    def __init__(self, A, *, m, xi, wts, dwts=None, recompute=False, device=0, Q=None, lam=None,
                 f_lambda=None, f_divdiff=None):
        """``Q`` (n x K), ``lam``, ``f_lambda`` = f(lam) deflate the forward and
        backward solves; ``f_divdiff`` (K x K divided differences of f at lam)
        is needed for ``backward_param`` with deflation."""
        lib = _lib.load()
        self._lib = lib
        self._h = C.c_void_p()
        A = sp.csr_matrix(A)
        A.sort_indices()
        self.n = A.shape[0]
        self.nnz = A.nnz
        self._ptr = np.ascontiguousarray(A.indptr, dtype=np.int32)
        self._ind = np.ascontiguousarray(A.indices, dtype=np.int32)
        val = np.ascontiguousarray(A.data, dtype=np.float64)
        xi_i = _cplx(xi)
        w_i = _cplx(wts)
        dw_i = None if dwts is None else _cplx(dwts)
        d = _lib.GmFabDesc()
        d.n = self.n
        d.sparse_format = _lib.GM_SPARSE_CSR
        d.nnz = self.nnz
        d.A_values = _p(val)
        d.A_ptr = self._ptr.ctypes.data_as(C.POINTER(C.c_int32))
        d.A_ind = self._ind.ctypes.data_as(C.POINTER(C.c_int32))
        d.subspace_size = int(m)
        d.contour_size = int(np.size(xi))
        d.xi = _p(xi_i)
        d.wts = _p(w_i)
        d.dwts = _p(dw_i)
        d.flags = _lib.GM_FAB_RECOMPUTE if recompute else 0
        d.device = int(device)
        if Q is not None and np.shape(Q)[1] > 0:
            self._Q = np.asfortranarray(Q, dtype=np.float64)
            self._lam = np.ascontiguousarray(lam, dtype=np.float64)
            self._flam = np.ascontiguousarray(f_lambda, dtype=np.float64)
            self._fdd = (None if f_divdiff is None
                         else np.asfortranarray(f_divdiff, dtype=np.float64))
            d.num_eigs = self._Q.shape[1]
            d.Q = _p(self._Q)
            d.lambda_ = _p(self._lam)
            d.f_lambda = _p(self._flam)
            d.f_divdiff = _p(self._fdd)
        _lib.check(lib.gm_fab_create(C.byref(d), C.byref(self._h)), fab=True)
        self.ns = int(np.size(xi))

    def _call(self, fn, *args):
        if not self._h:
            raise _lib.GpuMatfunError(_lib.GM_EBADHANDLE, "engine has been deleted")
        _lib.check(fn(self._h, *args), fab=True)

    def set_matrix(self, A):
        A = sp.csr_matrix(A)
        A.sort_indices()
        if A.nnz != self.nnz or not np.array_equal(A.indptr, self._ptr):
            raise ValueError("set_matrix needs the same sparsity pattern")
        self._call(self._lib.gm_fab_set_matrix, _p(np.ascontiguousarray(A.data, dtype=np.float64)))

    def set_weights(self, wts, dwts=None):
        self._call(self._lib.gm_fab_set_weights, _p(_cplx(wts)),
                   None if dwts is None else _p(_cplx(dwts)))

    def forward(self, b, want_u=True):
        """u = f(A) b; with want_u=False only the state a backward pass needs.  A
        float64 torch tensor on the engine's GPU in, one out (no host copies)."""
        b = _vec(b)
        u = _empty_like(b, self.n) if want_u else None
        self._call(self._lib.gm_fab_forward, _p(b), _p(u))
        return u

    def backward(self, g, kappa=None, want_b=True, want_kappa=True, want_alpha=True):
        g = np.ascontiguousarray(g, dtype=np.float64)
        gb = np.empty(self.n) if want_b else None
        gk = np.empty(self.n) if want_kappa else None
        ga = np.zeros(1) if want_alpha else None
        kap = None if kappa is None else np.ascontiguousarray(kappa, dtype=np.float64)
        self._call(self._lib.gm_fab_backward, _p(g), _p(kap), _p(gb), _p(gk), _p(ga))
        return gb, gk, (None if ga is None else float(ga[0]))

    def _derivs(self, dA_list):
        vals = []
        for E in dA_list:
            E = sp.csr_matrix(E)
            E.sort_indices()
            if E.nnz != self.nnz or not np.array_equal(E.indptr, self._ptr) or \
                    not np.array_equal(E.indices, self._ind):
                raise ValueError("dA/dtheta must have A's sparsity pattern")
            vals.append(np.ascontiguousarray(E.data, dtype=np.float64))
        return vals, (C.POINTER(C.c_double) * max(1, len(vals)))(*[_p(v) for v in vals])

    def set_param_derivs(self, dA_list):
        """Keep the dA/dtheta_q on the device for ``backward_param(g, None)``."""
        vals, arr = self._derivs(dA_list)
        self._call(self._lib.gm_fab_set_param_derivs, len(vals), arr)
        self._nstored = len(vals)

    def backward_param(self, g, dA_list, want_b=False):
        """dL/dtheta_q = Im sum_j c_j z_j^T (dA/dtheta_q) x_j for each sparse dA/dtheta_q
        (A's sparsity pattern) after a forward pass; returns (grad_b or None, grads).
        ``dA_list=None`` uses the derivatives kept by ``set_param_derivs``."""
        g = _vec(g)
        if dA_list is None:
            n_p, arr = getattr(self, "_nstored", 0), None
        else:
            vals, arr = self._derivs(dA_list)
            n_p = len(vals)
        out = np.zeros(max(1, n_p))
        gb = _empty_like(g, self.n) if want_b else None
        self._call(self._lib.gm_fab_backward_param, _p(g), n_p, arr, _p(gb), _p(out))
        return gb, out[:n_p].copy()

    def delete(self):
        if self._h:
            _lib.check(self._lib.gm_fab_delete(self._h), fab=True)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self._lib.gm_fab_delete(self._h)
        except Exception:
            pass


class FractionalDiffusion:
    """f(A(kappa)) b on an n x n grid with a fixed contour on [l1, lN].

    ``apply(b, kappa, alpha)`` (torch tensors, float64) is differentiable in
    all three arguments through :class:`FabFunction`.
    """

    def __init__(self, n, *, tau, l1, lN, ns=32, m=200, recompute=False, device=0):
        self.n_side = n
        self.tau = float(tau)
        c = q.contour_nodes(l1, lN, ns, tau, 1.0)
        self.xi, self.dxidt, self.s = c["xi"], c["dxidt"], c["s"]
        self.m = m
        self.recompute = recompute
        self.device = device
        self.engine = None

    def weights(self, alpha):
        return contour_weights(self.xi, self.dxidt, self.s, self.tau, alpha)

    def matrix(self, kappa):
        return problems.varcoef_laplacian_2d(kappa, self.n_side)

    def ensure(self, A, alpha):
        c, dc = self.weights(alpha)
        if self.engine is None:
            self.engine = FabEngine(A, m=self.m, xi=self.xi, wts=c, dwts=dc,
                                    recompute=self.recompute, device=self.device)
        else:
            self.engine.set_matrix(A)
            self.engine.set_weights(c, dc)
        return self.engine

    def apply(self, b, kappa, alpha):
        return FabFunction.apply(b, kappa, alpha, self)


try:
    import torch

    class FabFunction(torch.autograd.Function):
        """u = f(A(kappa)) b; backward via the C ABI adjoint."""

        @staticmethod
        def forward(ctx, b, kappa, alpha, op):
            kap = kappa.detach().cpu().double().numpy()
            a = float(alpha.detach().cpu())
            eng = op.ensure(op.matrix(kap), a)
            u = eng.forward(b.detach().cpu().double().numpy())
            ctx.op = op
            ctx.save_for_backward(b, kappa, alpha)
            return torch.from_numpy(u).to(b.device)

        @staticmethod
        def backward(ctx, grad_u):
            b, kappa, alpha = ctx.saved_tensors
            op = ctx.op
            kap = kappa.detach().cpu().double().numpy()
            eng = op.ensure(op.matrix(kap), float(alpha.detach().cpu()))
            eng.forward(b.detach().cpu().double().numpy())   # state for this (b, kappa, alpha)
            gb, gk, ga = eng.backward(grad_u.detach().cpu().double().numpy(), kap)
            dev = b.device
            return (torch.from_numpy(gb).to(dev), torch.from_numpy(gk).to(dev),
                    torch.tensor(ga, dtype=torch.float64, device=dev), None)
except ImportError:   # pragma: no cover - torch is part of the image
    FabFunction = None


class BatchFabEngine:
    """f(A) on R columns at once with the parameter adjoint (C ABI ``gm_bfab_*``).

    Columns pair like the batched engine's: even columns use ``wts1`` and
    ``f1_lambda``, odd columns ``wts2`` and ``f2_lambda``.  Vectors are N x R
    (row-major: NumPy C order, or float64 torch tensors on the GPU)."""

    def __init__(self, A, *, R, m, xi, wts1, wts2, Q=None, lam=None, f1_lambda=None,
                 f2_lambda=None, f1_divdiff=None, f2_divdiff=None, device=0):
        lib = _lib.load()
        self._lib = lib
        self._h = C.c_void_p()
        A = sp.csr_matrix(A)
        A.sort_indices()
        self.n, self.nnz, self.R = A.shape[0], A.nnz, int(R)
        self._ptr = np.ascontiguousarray(A.indptr, dtype=np.int32)
        self._ind = np.ascontiguousarray(A.indices, dtype=np.int32)
        keep = [np.ascontiguousarray(A.data, dtype=np.float64), _cplx(xi), _cplx(wts1),
                _cplx(wts2)]
        d = _lib.GmBfabDesc()
        d.n = self.n
        d.sparse_format = _lib.GM_SPARSE_CSR
        d.nnz = self.nnz
        d.A_values = _p(keep[0])
        d.A_ptr = self._ptr.ctypes.data_as(C.POINTER(C.c_int32))
        d.A_ind = self._ind.ctypes.data_as(C.POINTER(C.c_int32))
        d.columns = self.R
        d.subspace_size = int(m)
        d.contour_size = int(np.size(xi))
        d.xi, d.wts1, d.wts2 = _p(keep[1]), _p(keep[2]), _p(keep[3])
        d.device = int(device)
        if Q is not None and np.shape(Q)[1] > 0:
            arrs = [np.asfortranarray(Q, dtype=np.float64),
                    np.ascontiguousarray(lam, dtype=np.float64),
                    np.ascontiguousarray(f1_lambda, dtype=np.float64),
                    np.ascontiguousarray(f2_lambda, dtype=np.float64)]
            dds = [None if x is None else np.asfortranarray(x, dtype=np.float64)
                   for x in (f1_divdiff, f2_divdiff)]
            keep += arrs + [x for x in dds if x is not None]
            d.num_eigs = arrs[0].shape[1]
            d.Q, d.lambda_, d.f1_lambda, d.f2_lambda = [_p(x) for x in arrs]
            d.f1_divdiff, d.f2_divdiff = _p(dds[0]), _p(dds[1])
        _lib.check(lib.gm_bfab_create(C.byref(d), C.byref(self._h)), bfab=True)
        self._np = 0

    def _call(self, fn, *args):
        if not self._h:
            raise _lib.GpuMatfunError(_lib.GM_EBADHANDLE, "engine has been deleted")
        _lib.check(fn(self._h, *args), bfab=True)

    def _mat(self, B):
        if hasattr(B, "data_ptr"):
            import torch
            torch.cuda.current_stream().synchronize()
            return B.contiguous()
        return np.ascontiguousarray(B, dtype=np.float64)

    def _out(self, B):
        if hasattr(B, "data_ptr"):
            import torch
            return torch.empty((self.n, self.R), dtype=torch.float64, device=B.device)
        return np.empty((self.n, self.R))

    def forward(self, B, want_u=True):
        B = self._mat(B)
        U = self._out(B) if want_u else None
        self._call(self._lib.gm_bfab_forward, _p(B), _p(U))
        return U

    def set_param_derivs(self, dA_list):
        vals = []
        for E in dA_list:
            E = sp.csr_matrix(E)
            E.sort_indices()
            if E.nnz != self.nnz or not np.array_equal(E.indptr, self._ptr) or \
                    not np.array_equal(E.indices, self._ind):
                raise ValueError("dA/dtheta must have A's sparsity pattern")
            vals.append(np.ascontiguousarray(E.data, dtype=np.float64))
        arr = (C.POINTER(C.c_double) * max(1, len(vals)))(*[_p(v) for v in vals])
        self._call(self._lib.gm_bfab_set_param_derivs, len(vals), arr)
        self._np = len(vals)

    def backward_param(self, G, want_b=True):
        """(f(A) G or None, grad_theta [nparams][R]) after a forward pass."""
        G = self._mat(G)
        gb = self._out(G) if want_b else None
        gt = np.zeros((max(1, self._np), self.R))
        self._call(self._lib.gm_bfab_backward_param, _p(G), _p(gb), _p(gt))
        return gb, gt[:self._np]

    def delete(self):
        if self._h:
            _lib.check(self._lib.gm_bfab_delete(self._h), bfab=True)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self._lib.gm_bfab_delete(self._h)
        except Exception:
            pass
