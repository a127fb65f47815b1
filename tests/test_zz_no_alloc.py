"""gm_fab_forward_noalloc (the entry point that allocates nothing) against
gm_fab_forward, through the C ABI: two calls in a row, the second on a smaller
problem (the shared scratch array keeps the first call's contents), each equal
to the handle-based forward on the same inputs (the same kernels and the same
reduction partitions, so the same bits); device memory unchanged by a call;
and the size limit reported as GM_EINVAL before any work is enqueued."""

import ctypes as C

import numpy as np
import pytest

from conftest import rel
from gpumatfun_amd import _lib, problems, quadrature as q
from gpumatfun_amd.adjoint import FabEngine, contour_weights

pytestmark = pytest.mark.gpu
TAU, ALPHA = 1e-3, 1.5


def _problem(n, seed):
    kappa = np.exp(0.5 * problems.gaussian_random_field(n, 0.1, seed=seed))
    A = problems.varcoef_laplacian_2d(kappa, n)
    w = np.linalg.eigvalsh(A.toarray()) if A.shape[0] <= 4096 else None
    lo, hi = (w.min(), w.max()) if w is not None else (0.5, 8.5 * (n + 1) ** 2)
    c = q.contour_nodes(0.95 * lo, 1.05 * hi, 32, TAU, 1.0)
    cw, _ = contour_weights(c["xi"], c["dxidt"], c["s"], TAU, ALPHA)
    b = np.random.default_rng(seed).standard_normal(A.shape[0])
    return A, c["xi"], cw, b


def _cplx(z):
    z = np.asarray(z, dtype=np.complex128).ravel()
    out = np.empty(2 * z.size)
    out[0::2], out[1::2] = z.real, z.imag
    return out


def _noalloc(lib, A, xi, wts, b, m, torch):
    dev = torch.device("cuda:0")
    ptr = torch.as_tensor(np.asarray(A.indptr, dtype=np.int32), device=dev)
    ind = torch.as_tensor(np.asarray(A.indices, dtype=np.int32), device=dev)
    val = torch.as_tensor(np.asarray(A.data, dtype=np.float64), device=dev)
    xt = torch.as_tensor(_cplx(xi), device=dev)
    wt = torch.as_tensor(_cplx(wts), device=dev)
    bt = torch.as_tensor(b, device=dev)
    ut = torch.full_like(bt, float("nan"))
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info(0)[0]
    status = C.c_uint32(0xFFFF)
    st = lib.gm_fab_forward_noalloc(A.shape[0], A.nnz, C.c_void_p(ptr.data_ptr()),
                                    C.c_void_p(ind.data_ptr()), C.c_void_p(val.data_ptr()), m,
                                    len(xi), C.c_void_p(xt.data_ptr()), C.c_void_p(wt.data_ptr()),
                                    C.c_void_p(bt.data_ptr()), C.c_void_p(ut.data_ptr()), None, 0,
                                    C.byref(status))
    assert st == _lib.GM_OK, lib.gm_fab_last_error()
    assert status.value == 0
    free1 = torch.cuda.mem_get_info(0)[0]
    return ut.cpu().numpy(), free0, free1


def test_noalloc_forward_matches_handle_forward_twice():
    import torch
    lib = _lib.load()
    outs = []
    for n, seed, m in ((96, 1, 120), (48, 2, 60)):
        A, xi, cw, b = _problem(n, seed)
        print(f"case noalloc forward {n}^2 m={m}", flush=True)
        eng = FabEngine(A, m=m, xi=xi, wts=cw)
        u_ref = eng.forward(b)
        eng.delete()
        u, free0, free1 = _noalloc(lib, A, xi, cw, b, m, torch)
        e = rel(u, u_ref)
        print(f"  rel diff to gm_fab_forward {e:.2e}, bitwise {np.array_equal(u, u_ref)}, "
              f"free device memory {free0} -> {free1}", flush=True)
        assert e <= 1e-14
        assert free1 == free0
        outs.append(u)
    assert outs[1].size < outs[0].size


def test_noalloc_rejects_sizes_beyond_the_scratch_limit():
    lib = _lib.load()
    status = C.c_uint32(0)
    # n (m + 2) = 2e5 * 82 doubles > 64 MiB: refused before any device access
    fake = C.c_void_p(16)
    st = lib.gm_fab_forward_noalloc(200_000, 1_000_000, fake, fake, fake, 80, 32, fake, fake,
                                    fake, fake, None, 0, C.byref(status))
    assert st == _lib.GM_EINVAL
    assert b"limit" in lib.gm_fab_last_error()
