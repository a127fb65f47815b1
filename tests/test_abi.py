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


def test_struct_layout_matches_header_field_order():
    text = open(HEADER).read()
    body = text[text.index("typedef struct gm_create_desc"):text.index("} gm_create_desc;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"\b([a-zA-Z_][a-zA-Z0-9_]*)\s*[,;]", body)
    py = [f[0].rstrip("_") for f in _lib.GmCreateDesc._fields_]
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
