"""ctypes binding of the C ABI in ``include/gpumatfun.h``.

The shared library ``libgpumatfun_amd.so`` is built in-tree by
``gpumatfun_amd/build.py``.  There is no fallback: if the library is missing
or cannot be loaded, every entry point raises.
"""

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpumatfun_amd.so")

GM_OK, GM_EINVAL, GM_ENOMEM, GM_EDEVICE, GM_EBREAKDOWN, GM_EBADHANDLE, GM_ESTATE = range(7)
GM_SPARSE_CSC, GM_SPARSE_CSR = 0, 1
GM_FLAG_NO_CLAMP = 1
GM_FLAG_AUTOTUNE = 2
GM_FLAG_ROW_GROUPS = 4
GM_FLAG_ROW_BLOCKS = 8
GM_FLAG_INDEX32 = 16
GM_FLAG_INDEX16 = 32
GM_FLAG_EARLY_LOADS = 64
GM_FLAG_WAVE_SOLVES = 128
GM_FLAG_COLUMN_SLABS = 256
GM_FLAG_LOCALITY_ORDER = 512
GM_REACTION_FISHER, GM_REACTION_MS = 0, 2
# Mitchell-Schaeffer defaults (the library's own, gm_solver.hip)
MS_DEFAULTS = dict(tau_in=0.3, tau_out=6.0, tau_open=120.0, tau_close=150.0, v_gate=0.13)

STATUS_NAMES = {0: "GM_OK", 1: "GM_EINVAL", 2: "GM_ENOMEM", 3: "GM_EDEVICE",
                4: "GM_EBREAKDOWN", 5: "GM_EBADHANDLE", 6: "GM_ESTATE"}

# Every symbol declared in include/gpumatfun.h.
EXPORTS = (
    "gm_last_error", "gm_lanczos_create", "gm_lanczos_delete", "gm_build_subspace",
    "gm_lin_sys_solves", "gm_apply_V", "gm_update_U", "gm_update_U_monodomain",
    "gm_get_U", "gm_get_xsol", "gm_get_aabb", "gm_set_y", "gm_get_y", "gm_get_state",
    "gm_set_state", "gm_step", "gm_synchronize", "gm_get_info", "gm_get_schedule",
    "gm_get_stream", "gm_get_solve_form",
    "gm_profile_spmm", "gm_profile_spmm_flushed", "gm_profile_copy", "gm_profile_barrier", "gm_set_ms_params", "gm_update_U_ms",
    "gm_batch_last_error", "gm_batch_create", "gm_batch_delete", "gm_batch_step",
    "gm_batch_get_U", "gm_batch_synchronize", "gm_batch_info", "gm_batch_get_stream",
    "gm_batch_get_ode", "gm_batch_get_column_slabs",
    "gm_fab_last_error", "gm_fab_create", "gm_fab_delete", "gm_fab_set_matrix",
    "gm_fab_set_weights", "gm_fab_forward", "gm_fab_forward_noalloc", "gm_fab_backward",
    "gm_fab_backward_param",
    "gm_fab_set_param_derivs",
    "gm_ubatch_last_error", "gm_ubatch_create", "gm_ubatch_delete", "gm_ubatch_step",
    "gm_ubatch_get_U", "gm_ubatch_synchronize", "gm_ubatch_info", "gm_ubatch_get_schedule", "gm_ubatch_get_solve_form",
    "gm_polynomial_coeffs", "gm_contour_setup", "gm_shift_setup", "gm_eig_weights",
    "gm_subspace_size", "gm_rcm_order", "gm_locality_order",
    "gm_eigs_last_error", "gm_eigs_smallest", "gm_sym_eig",
    "gm_chol_last_error", "gm_nd_order", "gm_chol_factor", "gm_chol_solve", "gm_chol_info",
    "gm_chol_perm", "gm_chol_delete", "gm_eigs_shift_invert", "gm_lambda_max",
    "gm_bfab_last_error", "gm_bfab_create", "gm_bfab_delete", "gm_bfab_forward",
    "gm_bfab_set_param_derivs", "gm_bfab_backward_param", "gm_bfab_info",
)
GM_FAB_RECOMPUTE = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class GmCreateDesc(C.Structure):
    """Mirror of ``gm_create_desc`` (field order and types must match)."""
    _fields_ = [
        ("n", C.c_int32),
        ("U", _dp),
        ("sparse_format", C.c_int32),
        ("nnz", C.c_int64),
        ("A_values", _dp),
        ("A_ptr", _ip),
        ("A_ind", _ip),
        ("num_eigs", C.c_int32),
        ("Q", _dp),
        ("lambda_", _dp),
        ("subspace_size", C.c_int32),
        ("gamma_size", C.c_int32),
        ("gamma", _dp),
        ("contour_size", C.c_int32),
        ("gamma_s_real", _dp),
        ("gamma_s_imag", _dp),
        ("exp_lambda", _dp),
        ("phi_lambda", _dp),
        ("dt", C.c_double),
        ("ode", _dp),
        ("aode", C.c_double), ("bode", C.c_double), ("code", C.c_double),
        ("vrest", C.c_double), ("vamp", C.c_double), ("vth", C.c_double),
        ("vpeak", C.c_double), ("c1ion", C.c_double), ("c2ion", C.c_double),
        ("mass", _dp),
        ("wts1", _dp),
        ("wts2", _dp),
        ("shifts", _dp),
        ("valvec", _dp),
        ("flags", C.c_uint32),
        ("device", C.c_int32),
    ]


class GmMsParams(C.Structure):
    """Mirror of ``gm_ms_params`` (Mitchell-Schaeffer time constants and gate threshold)."""
    _fields_ = [("tau_in", C.c_double), ("tau_out", C.c_double), ("tau_open", C.c_double),
                ("tau_close", C.c_double), ("v_gate", C.c_double)]


class GmBatchDesc(C.Structure):
    """Mirror of ``gm_batch_desc``."""
    _fields_ = [
        ("n", C.c_int32),
        ("sparse_format", C.c_int32),
        ("nnz", C.c_int64),
        ("A_values", _dp),
        ("A_ptr", _ip),
        ("A_ind", _ip),
        ("num_problems", C.c_int32),
        ("U", _dp),
        ("subspace_size", C.c_int32),
        ("gamma_size", C.c_int32),
        ("gamma", _dp),
        ("contour_size", C.c_int32),
        ("gamma_s_real", _dp),
        ("gamma_s_imag", _dp),
        ("wts1", _dp),
        ("wts2", _dp),
        ("etaxi", _dp),
        ("dt", C.c_double),
        ("mass", _dp),
        ("chunk_columns", C.c_int32),
        ("flags", C.c_uint32),
        ("device", C.c_int32),
        ("num_eigs", C.c_int32),
        ("Q", _dp),
        ("exp_lambda", _dp),
        ("phi_lambda", _dp),
        ("reaction", C.c_int32),
        ("ode", _dp),
        ("ms", GmMsParams),
    ]


class GmFabDesc(C.Structure):
    """Mirror of ``gm_fab_desc``."""
    _fields_ = [
        ("n", C.c_int32),
        ("sparse_format", C.c_int32),
        ("nnz", C.c_int64),
        ("A_values", _dp),
        ("A_ptr", _ip),
        ("A_ind", _ip),
        ("subspace_size", C.c_int32),
        ("contour_size", C.c_int32),
        ("xi", _dp),
        ("wts", _dp),
        ("dwts", _dp),
        ("flags", C.c_uint32),
        ("device", C.c_int32),
        ("num_eigs", C.c_int32),
        ("Q", _dp),
        ("lambda_", _dp),
        ("f_lambda", _dp),
        ("f_divdiff", _dp),
    ]


class GmUbatchDesc(C.Structure):
    """Mirror of ``gm_ubatch_desc``."""
    _fields_ = [
        ("num_problems", C.c_int32),
        ("n", _ip),
        ("nnz", C.POINTER(C.c_int64)),
        ("sparse_format", C.c_int32),
        ("A_values", C.POINTER(_dp)),
        ("A_ptr", C.POINTER(_ip)),
        ("A_ind", C.POINTER(_ip)),
        ("U", C.POINTER(_dp)),
        ("mass", C.POINTER(_dp)),
        ("subspace_size", C.c_int32),
        ("gamma_size", C.c_int32),
        ("gamma", _dp),
        ("contour_size", C.c_int32),
        ("gamma_s_real", _dp),
        ("gamma_s_imag", _dp),
        ("wts1", _dp),
        ("wts2", _dp),
        ("shifts", _dp),
        ("dt", C.c_double),
        ("flags", C.c_uint32),
        ("device", C.c_int32),
    ]


class GmEigsDesc(C.Structure):
    """Mirror of ``gm_eigs_desc``."""
    _fields_ = [
        ("n", C.c_int32),
        ("nnz", C.c_int64),
        ("sparse_format", C.c_int32),
        ("A_values", _dp),
        ("A_ptr", _ip),
        ("A_ind", _ip),
        ("num_eigs", C.c_int32),
        ("block_size", C.c_int32),
        ("degree", C.c_int32),
        ("max_passes", C.c_int32),
        ("tol", C.c_double),
        ("upper_bound", C.c_double),
        ("seed", C.c_uint64),
        ("device", C.c_int32),
    ]


class GmSieigsDesc(C.Structure):
    """Mirror of ``gm_sieigs_desc``."""
    _fields_ = [
        ("n", C.c_int32),
        ("nnz", C.c_int64),
        ("A_values", _dp),
        ("A_ptr", _ip),
        ("A_ind", _ip),
        ("num_eigs", C.c_int32),
        ("block_size", C.c_int32),
        ("max_blocks", C.c_int32),
        ("max_restarts", C.c_int32),
        ("tol", C.c_double),
        ("sigma", C.c_double),
        ("seed", C.c_uint64),
        ("threads", C.c_int32),
    ]


class GmBfabDesc(C.Structure):
    """Mirror of ``gm_bfab_desc``."""
    _fields_ = [
        ("n", C.c_int32),
        ("sparse_format", C.c_int32),
        ("nnz", C.c_int64),
        ("A_values", _dp),
        ("A_ptr", _ip),
        ("A_ind", _ip),
        ("columns", C.c_int32),
        ("subspace_size", C.c_int32),
        ("contour_size", C.c_int32),
        ("xi", _dp),
        ("wts1", _dp),
        ("wts2", _dp),
        ("num_eigs", C.c_int32),
        ("Q", _dp),
        ("lambda_", _dp),
        ("f1_lambda", _dp),
        ("f2_lambda", _dp),
        ("f1_divdiff", _dp),
        ("f2_divdiff", _dp),
        ("device", C.c_int32),
    ]


class GpuMatfunError(RuntimeError):
    """An entry point returned a non-OK gm_status."""

    def __init__(self, status, message):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")


_lib = None


def load():
    """Load the shared library (once); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run gpumatfun_amd/build.py (hipcc, gfx950) first")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "gm_last_error": (C.c_char_p, []),
        "gm_lanczos_create": (C.c_int, [C.POINTER(GmCreateDesc), C.POINTER(vp)]),
        "gm_lanczos_delete": (C.c_int, [vp]),
        "gm_build_subspace": (C.c_int, [vp]),
        "gm_lin_sys_solves": (C.c_int, [vp]),
        "gm_apply_V": (C.c_int, [vp]),
        "gm_update_U": (C.c_int, [vp]),
        "gm_update_U_monodomain": (C.c_int, [vp]),
        "gm_get_U": (C.c_int, [vp, _dp]),
        "gm_get_xsol": (C.c_int, [vp, _dp]),
        "gm_get_aabb": (C.c_int, [vp, _dp]),
        "gm_set_y": (C.c_int, [vp, _dp, _dp]),
        "gm_get_y": (C.c_int, [vp, _dp, _dp]),
        "gm_get_state": (C.c_int, [vp, _dp, _dp, _dp]),
        "gm_set_state": (C.c_int, [vp, _dp, _dp, _dp]),
        "gm_step": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32]),
        "gm_synchronize": (C.c_int, [vp]),
        "gm_get_info": (C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32)]),
        "gm_get_stream": (C.c_int, [vp, C.POINTER(vp)]),
        "gm_get_solve_form": (C.c_int, [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "gm_get_schedule": (C.c_int, [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                      C.POINTER(C.c_int32)]),
        "gm_profile_spmm": (C.c_int, [vp, C.c_int32, C.c_int32, _dp]),
        "gm_profile_spmm_flushed": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int64, _dp]),
        "gm_profile_barrier": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, _dp, _dp]),
        "gm_profile_copy": (C.c_int, [vp, C.c_int64, C.c_int32, _dp]),
        "gm_set_ms_params": (C.c_int, [vp, C.POINTER(GmMsParams)]),
        "gm_update_U_ms": (C.c_int, [vp]),
        "gm_batch_last_error": (C.c_char_p, []),
        "gm_batch_create": (C.c_int, [C.POINTER(GmBatchDesc), C.POINTER(vp)]),
        "gm_batch_delete": (C.c_int, [vp]),
        "gm_batch_step": (C.c_int, [vp, C.c_int32, C.c_int32]),
        "gm_batch_get_U": (C.c_int, [vp, _dp]),
        "gm_batch_synchronize": (C.c_int, [vp]),
        "gm_batch_info": (C.c_int, [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "gm_batch_get_stream": (C.c_int, [vp, C.POINTER(vp)]),
        "gm_batch_get_ode": (C.c_int, [vp, _dp]),
        "gm_batch_get_column_slabs": (C.c_int, [vp, C.POINTER(C.c_int32)]),
        "gm_fab_last_error": (C.c_char_p, []),
        "gm_fab_create": (C.c_int, [C.POINTER(GmFabDesc), C.POINTER(vp)]),
        "gm_fab_delete": (C.c_int, [vp]),
        "gm_fab_set_matrix": (C.c_int, [vp, _dp]),
        "gm_fab_set_weights": (C.c_int, [vp, _dp, _dp]),
        "gm_fab_forward": (C.c_int, [vp, _dp, _dp]),
        "gm_fab_forward_noalloc": (C.c_int, [C.c_int32, C.c_int64, vp, vp, vp, C.c_int32,
                                             C.c_int32, vp, vp, vp, vp, vp, C.c_int32,
                                             C.POINTER(C.c_uint32)]),
        "gm_fab_backward": (C.c_int, [vp, _dp, _dp, _dp, _dp, _dp]),
        "gm_fab_backward_param": (C.c_int, [vp, _dp, C.c_int32, C.POINTER(_dp), _dp, _dp]),
        "gm_fab_set_param_derivs": (C.c_int, [vp, C.c_int32, C.POINTER(_dp)]),
        "gm_ubatch_last_error": (C.c_char_p, []),
        "gm_ubatch_create": (C.c_int, [C.POINTER(GmUbatchDesc), C.POINTER(vp)]),
        "gm_ubatch_delete": (C.c_int, [vp]),
        "gm_ubatch_step": (C.c_int, [vp, C.c_int32, C.c_int32]),
        "gm_ubatch_get_U": (C.c_int, [vp, C.c_int32, _dp]),
        "gm_ubatch_synchronize": (C.c_int, [vp]),
        "gm_ubatch_info": (C.c_int, [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                     C.POINTER(C.c_int32)]),
        "gm_ubatch_get_schedule": (C.c_int, [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                             C.POINTER(C.c_int32)]),
        "gm_ubatch_get_solve_form": (C.c_int, [vp, C.POINTER(C.c_int32)]),
        "gm_polynomial_coeffs": (C.c_int, [C.c_int32, C.c_double, _dp]),
        "gm_contour_setup": (C.c_int, [C.c_double, C.c_double, C.c_int32, C.c_double,
                                       C.c_double, _dp, _dp, _dp, _dp, _dp]),
        "gm_shift_setup": (C.c_int, [_dp, C.c_int32, _dp, C.c_int32, C.c_int32, _dp, _dp,
                                     _dp, _dp]),
        "gm_eig_weights": (C.c_int, [_dp, C.c_int32, C.c_double, C.c_double, _dp, _dp]),
        "gm_subspace_size": (C.c_int, [_dp, C.c_int32, C.c_double, C.c_double, C.c_double,
                                       C.POINTER(C.c_int32)]),
        "gm_rcm_order": (C.c_int, [C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32)]),
        "gm_locality_order": (C.c_int, [C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                        C.c_int32, C.POINTER(C.c_int32)]),
        "gm_eigs_last_error": (C.c_char_p, []),
        "gm_eigs_smallest": (C.c_int, [C.POINTER(GmEigsDesc), _dp, _dp, C.POINTER(C.c_int32),
                                       _dp]),
        "gm_sym_eig": (C.c_int, [C.c_int32, _dp, _dp, _dp]),
        "gm_chol_last_error": (C.c_char_p, []),
        "gm_nd_order": (C.c_int, [C.c_int32, _ip, _ip, _ip]),
        "gm_chol_factor": (C.c_int, [C.c_int32, _ip, _ip, _dp, C.c_double, C.c_int32,
                                     C.POINTER(vp)]),
        "gm_chol_solve": (C.c_int, [vp, C.c_int32, _dp]),
        "gm_chol_info": (C.c_int, [vp, C.POINTER(C.c_int64), _dp, C.POINTER(C.c_int32), _dp]),
        "gm_chol_perm": (C.c_int, [vp, _ip]),
        "gm_chol_delete": (C.c_int, [vp]),
        "gm_eigs_shift_invert": (C.c_int, [C.POINTER(GmSieigsDesc), _dp, _dp,
                                           C.POINTER(C.c_int32), _dp]),
        "gm_lambda_max": (C.c_int, [C.c_int32, _ip, _ip, _dp, C.c_double, C.c_int32,
                                    C.c_uint64, C.c_int32, _dp, C.POINTER(C.c_int32)]),
        "gm_bfab_last_error": (C.c_char_p, []),
        "gm_bfab_create": (C.c_int, [C.POINTER(GmBfabDesc), C.POINTER(vp)]),
        "gm_bfab_delete": (C.c_int, [vp]),
        "gm_bfab_forward": (C.c_int, [vp, _dp, _dp]),
        "gm_bfab_set_param_derivs": (C.c_int, [vp, C.c_int32, C.POINTER(_dp)]),
        "gm_bfab_backward_param": (C.c_int, [vp, _dp, _dp, _dp]),
        "gm_bfab_info": (C.c_int, [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(status, batch=False, fab=False, ubatch=False, eigs=False, bfab=False, chol=False):
    if status != GM_OK:
        lib = load()
        if chol:
            raw = lib.gm_chol_last_error()
        elif bfab:
            raw = lib.gm_bfab_last_error()
        elif eigs:
            raw = lib.gm_eigs_last_error()
        elif ubatch:
            raw = lib.gm_ubatch_last_error()
        elif fab:
            raw = lib.gm_fab_last_error()
        elif batch:
            raw = lib.gm_batch_last_error()
        else:
            raw = lib.gm_last_error()
        raise GpuMatfunError(status, raw.decode("utf-8", "replace"))
