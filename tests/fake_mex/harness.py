"""Build and drive the MEX gateways (gpumatfun_amd/mex/gm_mex.cpp) without MATLAB.

Each gateway is compiled with g++ against the mex.h stand-in in this folder,
its ``mexFunction`` renamed ``gmx_<name>``, and linked with fake_mex.cpp and
libgpumatfun_amd.so into one test library.  ``Mex`` calls a gateway with
mxArrays built from NumPy arrays, as MATLAB would.
"""

import ctypes as C
import os
import subprocess

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(ROOT, "gpumatfun_amd", "mex", "gm_mex.cpp")
LIBDIR = os.path.join(ROOT, "gpumatfun_amd")
OUT = os.path.join(HERE, "_build", "libgm_mex_harness.so")

GATEWAYS = {
    "lanczos_create": 1, "lanczos_delete": 2, "build_subspace_mex": 3,
    "lin_sys_solves_mex": 4, "apply_V_mex": 5, "update_U_mex": 6,
    "update_U_monodomain_mex": 7, "get_U_mex": 8, "get_xsol_mex": 9,
    "get_aabb_mex": 10, "set_y_mex": 11, "eigs_smallest_mex": 12,
    "eigs_shift_invert_mex": 13, "lambda_max_mex": 14,
}


def build():
    """Compile the gateways and the fake MEX runtime into OUT.  Several pytest
    workers may call this at once: the build runs under an exclusive lock,
    from per-process object files, and OUT is replaced atomically."""
    import fcntl
    bdir = os.path.dirname(OUT)
    os.makedirs(bdir, exist_ok=True)
    with open(os.path.join(bdir, ".lock"), "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        tag = str(os.getpid())
        objs = []
        inc = ["-I" + HERE, "-I" + os.path.join(ROOT, "include")]
        for name, gid in GATEWAYS.items():
            o = os.path.join(bdir, f"{name}.{tag}.o")
            subprocess.run(["g++", "-O2", "-fPIC", "-std=c++17", "-c", SRC, "-o", o,
                            f"-DGM_GATEWAY={gid}", f"-DmexFunction=gmx_{name}"] + inc, check=True)
            objs.append(o)
        fo = os.path.join(bdir, f"fake_mex.{tag}.o")
        subprocess.run(["g++", "-O2", "-fPIC", "-std=c++17", "-c",
                        os.path.join(HERE, "fake_mex.cpp"), "-o", fo] + inc, check=True)
        tmp = f"{OUT}.{tag}.tmp"
        subprocess.run(["g++", "-shared", "-o", tmp] + objs + [fo] +
                       ["-L" + LIBDIR, "-lgpumatfun_amd", "-Wl,-rpath," + LIBDIR], check=True)
        os.replace(tmp, OUT)
        for o in objs + [fo]:
            os.remove(o)
    return OUT


class MexError(RuntimeError):
    def __init__(self, ident, msg):
        self.ident = ident
        super().__init__(f"{ident}: {msg}")


class Mex:
    """Call the gateways as MATLAB would: ``Mex()('get_U_mex', h)``."""

    def __init__(self):
        if not os.path.exists(OUT):
            build()
        self.lib = C.CDLL(OUT)
        L = self.lib
        vp = C.c_void_p
        L.fm_double.restype = vp
        L.fm_double.argtypes = [C.c_size_t, C.c_size_t, C.POINTER(C.c_double)]
        L.fm_complex.restype = vp
        L.fm_complex.argtypes = [C.c_size_t, C.c_size_t, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]
        L.fm_sparse.restype = vp
        L.fm_sparse.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t, C.POINTER(C.c_double),
                                C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.fm_uint64.restype = vp
        L.fm_uint64.argtypes = [C.c_uint64]
        L.fm_get_uint64.restype = C.c_uint64
        L.fm_get_uint64.argtypes = [vp]
        L.fm_get_m.restype = C.c_size_t
        L.fm_get_m.argtypes = [vp]
        L.fm_get_n.restype = C.c_size_t
        L.fm_get_n.argtypes = [vp]
        L.fm_copy_pr.argtypes = [vp, C.POINTER(C.c_double)]
        L.fm_destroy.argtypes = [vp]
        L.fm_call.restype = C.c_int
        L.fm_call.argtypes = [vp, C.c_int, C.POINTER(vp), C.c_int, C.POINTER(vp),
                              C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t]

    # -- mxArray construction -------------------------------------------------
    def to_mx(self, x):
        L = self.lib
        if isinstance(x, np.uint64):
            return L.fm_uint64(int(x))
        if sp.issparse(x):
            A = sp.csc_matrix(x)
            A.sort_indices()
            pr = np.ascontiguousarray(A.data, dtype=np.float64)
            ir = np.ascontiguousarray(A.indices, dtype=np.int64)
            jc = np.ascontiguousarray(A.indptr, dtype=np.int64)
            return L.fm_sparse(A.shape[0], A.shape[1], A.nnz,
                               pr.ctypes.data_as(C.POINTER(C.c_double)),
                               ir.ctypes.data_as(C.POINTER(C.c_int64)),
                               jc.ctypes.data_as(C.POINTER(C.c_int64)))
        a = np.asarray(x)
        if a.ndim == 0:
            a = a.reshape(1, 1)
        elif a.ndim == 1:
            a = a.reshape(-1, 1)
        if np.iscomplexobj(a):
            re = np.ascontiguousarray(a.real.reshape(-1, order="F"), dtype=np.float64)
            im = np.ascontiguousarray(a.imag.reshape(-1, order="F"), dtype=np.float64)
            return L.fm_complex(a.shape[0], a.shape[1], re.ctypes.data_as(C.POINTER(C.c_double)),
                                im.ctypes.data_as(C.POINTER(C.c_double)))
        f = np.ascontiguousarray(a.reshape(-1, order="F"), dtype=np.float64)
        return L.fm_double(a.shape[0], a.shape[1], f.ctypes.data_as(C.POINTER(C.c_double)))

    def from_mx(self, p, keep_uint64=False):
        L = self.lib
        m, n = L.fm_get_m(p), L.fm_get_n(p)
        if keep_uint64:
            return np.uint64(L.fm_get_uint64(p))
        out = np.empty(m * n)
        L.fm_copy_pr(p, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out.reshape(m, n, order="F")

    def __call__(self, name, *args, nlhs=None):
        L = self.lib
        if nlhs is None:
            nlhs = 1 if name in ("lanczos_create", "get_U_mex", "get_xsol_mex",
                                 "get_aabb_mex") else 0
        fn = C.cast(getattr(L, f"gmx_{name}"), C.c_void_p)
        prhs = [self.to_mx(a) for a in args]
        arr_in = (C.c_void_p * max(1, len(prhs)))(*prhs)
        arr_out = (C.c_void_p * max(1, nlhs))()
        idb = C.create_string_buffer(256)
        msgb = C.create_string_buffer(2048)
        rc = L.fm_call(fn, nlhs, arr_out, len(prhs), arr_in, idb, 256, msgb, 2048)
        for p in prhs:
            L.fm_destroy(p)
        if rc:
            raise MexError(idb.value.decode(), msgb.value.decode())
        outs = []
        for i in range(nlhs):
            p = arr_out[i]
            outs.append(self.from_mx(p, keep_uint64=(name == "lanczos_create")))
            L.fm_destroy(p)
        return outs[0] if nlhs == 1 else (tuple(outs) if outs else None)
