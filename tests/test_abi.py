"""C ABI contract (no GPU needed): the library loads, exports every symbol the
header declares, and rejects malformed inputs with the reference's MEX
identifiers before touching a device."""

import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from gpumatfun_amd import _lib
from gpumatfun_amd.engine import LanczosEngine, GpuMatfunError

HEADER = os.path.join(ROOT, "include", "gpumatfun.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gm_[a-zA-Z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    decl = header_functions()
    assert len(decl) >= 21
    missing = [f for f in decl if not hasattr(lib, f)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == decl


@pytest.mark.parametrize("cname,pyname", [("gm_create_desc", "GmCreateDesc"),
                                           ("gm_ms_params", "GmMsParams"),
                                           ("gm_batch_desc", "GmBatchDesc"),
                                           ("gm_fab_desc", "GmFabDesc"),
                                           ("gm_ubatch_desc", "GmUbatchDesc"),
                                           ("gm_eigs_desc", "GmEigsDesc"),
                                           ("gm_bfab_desc", "GmBfabDesc"),
                                           ("gm_sieigs_desc", "GmSieigsDesc")])
def test_struct_layout_matches_header_field_order(cname, pyname):
    text = open(HEADER).read()
    body = text[text.index("typedef struct " + cname):text.index("} " + cname + ";")]
    body = body[body.index("{") + 1:]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"\b([a-zA-Z_][a-zA-Z0-9_]*)\s*[,;]", body)
    py = [f[0].rstrip("_") for f in getattr(_lib, pyname)._fields_]
    assert names == py


def _good_desc_inputs(n=9):
    from gpumatfun_amd import problems, quadrature as q
    A = problems.laplacian_2d_dirichlet(3)
    lam = problems.dirichlet_eigs_2d(3)
    u = np.linspace(0.1, 0.9, n)
    return q.make_inputs(A, u, u * (1 - u), m=4, n=4, eta=1e-2, alpha=1.5, dt=0.1,
                         l1=float(lam.min()), lN=float(lam.max()))


def _create_raw(**over):
    """Build a gm_create_desc from good inputs, apply overrides, call create."""
    inp = _good_desc_inputs()
    A = inp.A.tocsr()
    keep = {}

    def arr(name, a, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype=dtype)
        keep[name] = a
        return a.ctypes.data_as(C.POINTER(C.c_double if dtype == np.float64 else C.c_int32))

    d = _lib.GmCreateDesc()
    d.n = A.shape[0]
    d.U = arr("U", inp.U.reshape(-1, order="F"))
    d.sparse_format = _lib.GM_SPARSE_CSR
    d.nnz = A.nnz
    d.A_values = arr("val", A.data)
    d.A_ptr = arr("ptr", A.indptr, np.int32)
    d.A_ind = arr("ind", A.indices, np.int32)
    d.num_eigs = 0
    d.subspace_size = inp.m
    d.gamma_size = inp.gamma.size
    d.gamma = arr("g", inp.gamma)
    d.contour_size = 4
    d.gamma_s_real = arr("gsr", inp.gamma_s_real.reshape(-1, order="F"))
    d.gamma_s_imag = arr("gsi", inp.gamma_s_imag.reshape(-1, order="F"))
    d.dt = 0.1
    d.wts1 = arr("w1", inp.wts1.reshape(-1, order="F"))
    d.wts2 = arr("w2", inp.wts2.reshape(-1, order="F"))
    d.shifts = arr("sh", inp.shifts)
    d.valvec = arr("vv", [-1.0, 0.0, 1.0])
    for k, v in over.items():
        if isinstance(v, np.ndarray):
            v = arr(k, v, v.dtype.type)
        setattr(d, k, v)
    lib = _lib.load()
    h = C.c_void_p()
    st = lib.gm_lanczos_create(C.byref(d), C.byref(h))
    return st, lib.gm_last_error().decode()


@pytest.mark.parametrize("over,ident", [
    (dict(subspace_size=0), "bad_N"),
    (dict(gamma_size=0), "bad_gamma"),
    (dict(gamma_size=40), "bad_gamma"),
    (dict(wts1=None), "bad_wts1"),
    (dict(shifts=None), "bad_shifts"),
    (dict(num_eigs=2), "badQ"),
    (dict(dt=float("nan")), "bad_dt"),
    (dict(valvec=np.array([1.0, 0.0, 1.0])), "bad_valvec"),
    (dict(A_ind=np.array([0, 99, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                          0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0], dtype=np.int32)), "inputNotSquare"),
    (dict(sparse_format=7), "inputNotSparse"),
])
def test_create_rejects_bad_inputs(over, ident):
    st, msg = _create_raw(**over)
    assert st == _lib.GM_EINVAL
    assert f"MATLAB:lanczos_create:{ident}" in msg


def test_create_without_gpu_reports_device_error_not_crash(gpu_available):
    if gpu_available:
        pytest.skip("a GPU is present")
    st, msg = _create_raw()
    assert st == _lib.GM_EDEVICE
    assert "device" in msg.lower()


def test_null_handle_is_rejected():
    lib = _lib.load()
    assert lib.gm_build_subspace(None) == _lib.GM_EBADHANDLE
    assert lib.gm_lanczos_delete(C.c_void_p(12345)) == _lib.GM_EBADHANDLE
    assert b"handle" in lib.gm_last_error()


def test_engine_python_side_validation():
    inp = _good_desc_inputs()
    with pytest.raises(GpuMatfunError, match="bad_shifts"):
        LanczosEngine(inp.U, inp.A, inp.Q, inp.lam, inp.m, inp.gamma, inp.gamma_s_real,
                      inp.gamma_s_imag, inp.exp_lambda, inp.phi_lambda, inp.dt, inp.ode, 0, 0,
                      0, 0, 0, 0, 0, 0, 0, inp.mass, inp.wts1, inp.wts2, inp.shifts[:-2])
    with pytest.raises(GpuMatfunError, match="inputNotSparse"):
        LanczosEngine(inp.U, inp.A.toarray(), inp.Q, inp.lam, inp.m, inp.gamma,
                      inp.gamma_s_real, inp.gamma_s_imag, inp.exp_lambda, inp.phi_lambda,
                      inp.dt, inp.ode, 0, 0, 0, 0, 0, 0, 0, 0, 0, inp.mass, inp.wts1,
                      inp.wts2, inp.shifts)
    with pytest.raises(GpuMatfunError, match="bad_wts1"):
        LanczosEngine(inp.U, inp.A, inp.Q, inp.lam, inp.m, inp.gamma, inp.gamma_s_real,
                      inp.gamma_s_imag, inp.exp_lambda, inp.phi_lambda, inp.dt, inp.ode, 0, 0,
                      0, 0, 0, 0, 0, 0, 0, inp.mass, inp.wts1[:2], inp.wts2, inp.shifts)


def test_ubatch_create_rejects_bad_descriptors():
    """Validation runs before any device call, so this needs no GPU."""
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.gm_ubatch_create(None, C.byref(h)) == _lib.GM_EINVAL
    d = _lib.GmUbatchDesc()
    d.num_problems = 0
    assert lib.gm_ubatch_create(C.byref(d), C.byref(h)) == _lib.GM_EINVAL
    assert b"problem count" in lib.gm_ubatch_last_error()
    # one 2-row problem whose column index is out of range
    n = np.array([2], dtype=np.int32)
    nnz = np.array([2], dtype=np.int64)
    ptr = np.array([0, 1, 2], dtype=np.int32)
    ind = np.array([0, 5], dtype=np.int32)
    val = np.array([1.0, 1.0])
    U = np.zeros(4)
    g = np.array([1.0])
    sm = np.zeros(64)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    d.num_problems = 1
    d.n = n.ctypes.data_as(ip)
    d.nnz = nnz.ctypes.data_as(C.POINTER(C.c_int64))
    d.sparse_format = _lib.GM_SPARSE_CSR
    pv = (dp * 1)(val.ctypes.data_as(dp))
    pp = (ip * 1)(ptr.ctypes.data_as(ip))
    pi = (ip * 1)(ind.ctypes.data_as(ip))
    pu = (dp * 1)(U.ctypes.data_as(dp))
    d.A_values, d.A_ptr, d.A_ind, d.U = pv, pp, pi, pu
    d.subspace_size, d.gamma_size, d.contour_size = 2, 1, 1
    d.gamma = g.ctypes.data_as(dp)
    for f in ("gamma_s_real", "gamma_s_imag", "wts1", "wts2", "shifts"):
        setattr(d, f, sm.ctypes.data_as(dp))
    assert lib.gm_ubatch_create(C.byref(d), C.byref(h)) == _lib.GM_EINVAL
    assert b"out of range" in lib.gm_ubatch_last_error()
    assert lib.gm_ubatch_step(None, 1, 0) == _lib.GM_EBADHANDLE


def test_bfab_create_rejects_bad_descriptors():
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.gm_bfab_create(None, C.byref(h)) == _lib.GM_EINVAL
    d = _lib.GmBfabDesc()
    ptr = np.array([0, 1, 2], dtype=np.int32)
    ind = np.array([0, 1], dtype=np.int32)
    val = np.ones(2)
    z = np.zeros(8)
    dp = C.POINTER(C.c_double)
    d.n, d.nnz, d.sparse_format = 2, 2, _lib.GM_SPARSE_CSR
    d.A_values = val.ctypes.data_as(dp)
    d.A_ptr = ptr.ctypes.data_as(C.POINTER(C.c_int32))
    d.A_ind = ind.ctypes.data_as(C.POINTER(C.c_int32))
    d.subspace_size, d.contour_size = 2, 2
    d.xi = d.wts1 = d.wts2 = z.ctypes.data_as(dp)
    d.columns = 3
    assert lib.gm_bfab_create(C.byref(d), C.byref(h)) == _lib.GM_EINVAL
    assert b"columns" in lib.gm_bfab_last_error()
    d.columns = 2
    d.num_eigs = 1
    assert lib.gm_bfab_create(C.byref(d), C.byref(h)) == _lib.GM_EINVAL
    assert b"deflation" in lib.gm_bfab_last_error()
    assert lib.gm_bfab_forward(None, None, None) == _lib.GM_EBADHANDLE
