"""``LanczosEngine``: the host-side mirror of the reference's ``lanczosEngine.m``.

Same constructor argument order as ``lanczosEngine.m:8-10`` (which forwards
them to ``lanczos_create`` with ``valvec = -1:1``) and the same methods
(``lanczosEngine.m:21-55``), each a direct call of the C ABI in
``include/gpumatfun.h``.  Errors raise :class:`GpuMatfunError` carrying the
library's message, which for argument errors starts with the reference's MEX
identifier (e.g. ``MATLAB:lanczos_create:bad_shifts``).
"""

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import GpuMatfunError  # noqa: F401  (re-export)


def _f64(a, size=None, name="array"):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if size is not None and a.size != size:
        raise ValueError(f"{name}: expected {size} values, got {a.size}")
    return a


def _fcol(a):
    """Column-major (Fortran order) flat float64 copy, as MATLAB hands arrays over."""
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, order="F")
                                if np.ndim(a) > 1 else np.asarray(a, dtype=np.float64))


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


class LanczosEngine:
    """One device-resident f(A)b problem (``lanczosEngine.m``).

    Keyword options beyond the reference's 26 inputs: ``clamp`` (the u >= 0
    clamp of ``update_U.cu:9``), ``autotune`` (time every SpMM schedule on A
    at create and keep the fastest), ``schedule`` (without autotune: None
    picks row blocks when rows average <= 16 nonzeros and row groups
    otherwise; ``"rows"`` or ``"blocks"`` forces one), ``index16`` (without
    autotune: store the row-block SpMM's column indices as 16-bit offsets from
    the block's first row where they fit), ``index32`` (never use those
    offsets, not even when autotuning), ``early_loads`` (without autotune: the
    row-block SpMM loads each row's own operands before its gathers),
    ``wave_solves`` (without autotune: the shifted solves one wavefront per
    system by cyclic reduction, m <= 128), ``device``.
    """

    def __init__(self, U, A, Q, lambda_, N, gamma, gamma_s_real, gamma_s_imag,
                 exp_lambda, phi_lambda, dt, ode, aode, bode, code, vrest, vamp, vth,
                 vpeak, c1ion, c2ion, mass, wts1, wts2, shifts, valvec=(-1.0, 0.0, 1.0),
                 *, clamp=True, autotune=False, schedule=None, index32=False, index16=False,
                 early_loads=False, wave_solves=False, device=0):
        lib = _lib.load()
        self._lib = lib
        self._h = C.c_void_p()
        U = np.asarray(U, dtype=np.float64)
        n = U.shape[0]
        if U.ndim != 2 or U.shape[1] != 2:
            raise GpuMatfunError(_lib.GM_EINVAL,
                                 "MATLAB:lanczos_create:badU: Invalid U input argument.")
        if not sp.issparse(A):
            raise GpuMatfunError(_lib.GM_EINVAL, "MATLAB:lanczos_create:inputNotSparse: "
                                 "Invalid sparse matrix input argument.")
        if A.shape[0] != A.shape[1]:
            raise GpuMatfunError(_lib.GM_EINVAL, "MATLAB:lanczos_create:inputNotSquare: "
                                 "Sparse matrix must be square.")
        if sp.isspmatrix_csc(A) or getattr(A, "format", "") == "csc":
            fmt = _lib.GM_SPARSE_CSC
        else:
            A = sp.csr_matrix(A)
            fmt = _lib.GM_SPARSE_CSR
        self._A_val = np.ascontiguousarray(A.data, dtype=np.float64)
        self._A_ptr = np.ascontiguousarray(A.indptr, dtype=np.int32)
        self._A_ind = np.ascontiguousarray(A.indices, dtype=np.int32)
        Q = np.asarray(Q, dtype=np.float64).reshape(n, -1) if np.size(Q) else np.zeros((n, 0))
        K = Q.shape[1]
        gamma = _f64(gamma).ravel()
        nd = gamma.size
        gsr = np.asarray(gamma_s_real, dtype=np.float64).reshape(nd, -1)
        gsi = np.asarray(gamma_s_imag, dtype=np.float64).reshape(nd, -1)
        ns = gsr.shape[1]
        if gsi.shape != gsr.shape:
            raise GpuMatfunError(_lib.GM_EINVAL, "MATLAB:lanczos_create:bad_gamma_s_imag: "
                                 "Invalid gamma_s_imag input argument.")
        m = int(np.asarray(N).ravel()[0])
        w1 = np.asarray(wts1, dtype=np.float64)
        w2 = np.asarray(wts2, dtype=np.float64)
        if w1.shape != (ns, 2):
            raise GpuMatfunError(_lib.GM_EINVAL, "MATLAB:lanczos_create:bad_wts1: "
                                 "Invalid wts1 input argument.")
        if w2.shape != (ns, 2):
            raise GpuMatfunError(_lib.GM_EINVAL, "MATLAB:lanczos_create:bad_wts2: "
                                 "Invalid wts2 input argument.")
        sh = _f64(shifts).ravel()
        if sh.size != 4 * m * ns:
            raise GpuMatfunError(_lib.GM_EINVAL, "MATLAB:lanczos_create:bad_shifts: "
                                 "Invalid shifts input argument c.")
        lam = _f64(lambda_).ravel()
        el = _f64(exp_lambda).ravel()
        pl = _f64(phi_lambda).ravel()
        if lam.size != K:
            raise GpuMatfunError(_lib.GM_EINVAL, "MATLAB:lanczos_create:bad_lambda: "
                                 "Invalid lambda input argument.")
        if el.size != K:
            raise GpuMatfunError(_lib.GM_EINVAL, "MATLAB:lanczos_create:bad_exp_lambda: "
                                 "Invalid exp_lambda input argument.")
        if pl.size != K:
            raise GpuMatfunError(_lib.GM_EINVAL, "MATLAB:lanczos_create:bad_phi_lambda: "
                                 "Invalid phi_lambda input argument.")
        ode = _f64(ode).ravel()
        mass = _f64(mass).ravel()
        if ode.size != n:
            raise GpuMatfunError(_lib.GM_EINVAL, "MATLAB:lanczos_create:bad_ode: "
                                 "Invalid ode input argument.")
        if mass.size != n:
            raise GpuMatfunError(_lib.GM_EINVAL, "MATLAB:lanczos_create:bad_mass: "
                                 "Invalid mass input argument.")
        # keep every buffer alive for the duration of the create call
        self._keep = [_fcol(U), _fcol(Q), lam, gamma, _fcol(gsr), _fcol(gsi), el, pl, ode,
                      mass, _fcol(w1), _fcol(w2), sh,
                      None if valvec is None else _f64(valvec, 3, "valvec")]
        (Uf, Qf, lam, gamma, gsrf, gsif, el, pl, ode, mass, w1f, w2f, sh, vv) = self._keep
        d = _lib.GmCreateDesc()
        d.n = n
        d.U = _ptr(Uf)
        d.sparse_format = fmt
        d.nnz = self._A_val.size
        d.A_values = _ptr(self._A_val)
        d.A_ptr = self._A_ptr.ctypes.data_as(C.POINTER(C.c_int32))
        d.A_ind = self._A_ind.ctypes.data_as(C.POINTER(C.c_int32))
        d.num_eigs = K
        d.Q = _ptr(Qf)
        d.lambda_ = _ptr(lam)
        d.subspace_size = m
        d.gamma_size = nd
        d.gamma = _ptr(gamma)
        d.contour_size = ns
        d.gamma_s_real = _ptr(gsrf)
        d.gamma_s_imag = _ptr(gsif)
        d.exp_lambda = _ptr(el)
        d.phi_lambda = _ptr(pl)
        d.dt = float(dt)
        d.ode = _ptr(ode)
        d.aode, d.bode, d.code = float(aode), float(bode), float(code)
        d.vrest, d.vamp, d.vth = float(vrest), float(vamp), float(vth)
        d.vpeak, d.c1ion, d.c2ion = float(vpeak), float(c1ion), float(c2ion)
        d.mass = _ptr(mass)
        d.wts1 = _ptr(w1f)
        d.wts2 = _ptr(w2f)
        d.shifts = _ptr(sh)
        d.valvec = _ptr(vv)
        sched = {None: 0, "rows": _lib.GM_FLAG_ROW_GROUPS, "blocks": _lib.GM_FLAG_ROW_BLOCKS}
        if schedule not in sched:
            raise GpuMatfunError(_lib.GM_EINVAL, "schedule must be None, 'rows' or 'blocks'")
        d.flags = ((0 if clamp else _lib.GM_FLAG_NO_CLAMP)
                   | (_lib.GM_FLAG_AUTOTUNE if autotune else 0) | sched[schedule]
                   | (_lib.GM_FLAG_INDEX32 if index32 else 0)
                   | (_lib.GM_FLAG_INDEX16 if index16 else 0)
                   | (_lib.GM_FLAG_EARLY_LOADS if early_loads else 0)
                   | (_lib.GM_FLAG_WAVE_SOLVES if wave_solves else 0))
        d.device = int(device)
        _lib.check(lib.gm_lanczos_create(C.byref(d), C.byref(self._h)))
        self._keep = None
        self.n, self.m, self.ns, self.nd, self.K = n, m, ns, nd, K

    # -- construction from the set-up dataclass --------------------------------
    @classmethod
    def from_inputs(cls, inp, **kw):
        """Create from a :class:`gpumatfun_amd.quadrature.CreateInputs`."""
        return cls(inp.U, inp.A, inp.Q, inp.lam, inp.m, inp.gamma, inp.gamma_s_real,
                   inp.gamma_s_imag, inp.exp_lambda, inp.phi_lambda, inp.dt, inp.ode,
                   inp.aode, inp.bode, inp.code, inp.vrest, inp.vamp, inp.vth, inp.vpeak,
                   inp.c1ion, inp.c2ion, inp.mass, inp.wts1, inp.wts2, inp.shifts, **kw)

    # -- lanczosEngine.m methods ------------------------------------------------
    def _call(self, fn, *args):
        if not self._h:
            raise GpuMatfunError(_lib.GM_EBADHANDLE, "engine has been deleted")
        _lib.check(fn(self._h, *args))

    def delete(self):
        """``lanczosEngine.m:12-15``."""
        if self._h:
            _lib.check(self._lib.gm_lanczos_delete(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self._lib.gm_lanczos_delete(self._h)
        except Exception:
            pass

    def build_subspace(self):
        self._call(self._lib.gm_build_subspace)

    def lin_sys_solves(self):
        self._call(self._lib.gm_lin_sys_solves)

    def apply_V(self):
        self._call(self._lib.gm_apply_V)

    def update_U(self):
        self._call(self._lib.gm_update_U)

    def update_U_monodomain(self):
        self._call(self._lib.gm_update_U_monodomain)

    def update_U_ms(self):
        """Mitchell-Schaeffer reaction (gate kept in the ode vector)."""
        self._call(self._lib.gm_update_U_ms)

    def set_ms_params(self, tau_in=0.3, tau_out=6.0, tau_open=120.0, tau_close=150.0,
                      v_gate=0.13):
        p = _lib.GmMsParams(tau_in, tau_out, tau_open, tau_close, v_gate)
        self._call(self._lib.gm_set_ms_params, C.byref(p))

    def _get_pair(self, fn):
        out = np.empty(2 * self.n)
        self._call(fn, _ptr(out))
        return out.reshape(2, self.n).T.copy()

    def get_U(self):
        """N x 2 [u, dt*g(u)] (``get_U.cpp``)."""
        return self._get_pair(self._lib.gm_get_U)

    def get_xsol(self):
        """N x 2 [f1(A) ubar, f2(A) dt gbar] (``get_xsol.cpp``)."""
        return self._get_pair(self._lib.gm_get_xsol)

    def get_aabb(self):
        out = np.empty(4 * (self.m + 1))
        self._call(self._lib.gm_get_aabb, _ptr(out))
        return out.reshape(4, self.m + 1).T.copy()

    def set_y(self, yreal, yimag):
        yr = _fcol(np.asarray(yreal, dtype=np.float64).reshape(self.m, 2 * self.ns))
        yi = _fcol(np.asarray(yimag, dtype=np.float64).reshape(self.m, 2 * self.ns))
        self._call(self._lib.gm_set_y, _ptr(yr), _ptr(yi))

    def get_y(self):
        yr = np.empty(self.m * 2 * self.ns)
        yi = np.empty(self.m * 2 * self.ns)
        self._call(self._lib.gm_get_y, _ptr(yr), _ptr(yi))
        return (yr.reshape(2 * self.ns, self.m).T.copy(),
                yi.reshape(2 * self.ns, self.m).T.copy())

    # -- extensions ------------------------------------------------------------
    def get_state(self):
        u = np.empty(self.n)
        dtg = np.empty(self.n)
        ode = np.empty(self.n)
        self._call(self._lib.gm_get_state, _ptr(u), _ptr(dtg), _ptr(ode))
        return u, dtg, ode

    def set_state(self, u=None, dtg=None, ode=None):
        arrs = [None if a is None else _f64(a, self.n) for a in (u, dtg, ode)]
        self._call(self._lib.gm_set_state, *[_ptr(a) for a in arrs])

    def step(self, nsteps=1, monodomain=False, use_graph=False):
        """monodomain: False/0 Fisher-KPP, True/1 FitzHugh-Nagumo, 2 Mitchell-Schaeffer."""
        self._call(self._lib.gm_step, int(nsteps), int(monodomain), int(bool(use_graph)))

    def synchronize(self):
        self._call(self._lib.gm_synchronize)

    def info(self):
        n = C.c_int64(); nnz = C.c_int64()
        m = C.c_int32(); nd = C.c_int32(); ns = C.c_int32(); K = C.c_int32(); lanes = C.c_int32()
        self._call(self._lib.gm_get_info, C.byref(n), C.byref(nnz), C.byref(m), C.byref(nd),
                   C.byref(ns), C.byref(K), C.byref(lanes))
        kind = C.c_int32(); param = C.c_int32(); ib = C.c_int32()
        self._call(self._lib.gm_get_schedule, C.byref(kind), C.byref(param), C.byref(ib))
        if kind.value == 0:
            sched = f"k_spmm<{param.value}>"
        else:
            sched = (f"k_spmm_stream<{param.value}{',c16' if ib.value == 2 else ''}"
                     f"{',pre' if kind.value == 2 else ''}>")
        form = C.c_int32(); spb = C.c_int32()
        self._call(self._lib.gm_get_solve_form, C.byref(form), C.byref(spb))
        return {"n": n.value, "nnz": nnz.value, "m": m.value, "gamma_size": nd.value,
                "contour_size": ns.value, "num_eigs": K.value, "spmm_lanes": lanes.value,
                "spmm_schedule": sched, "spmm_kind": kind.value, "spmm_param": param.value,
                "index_bytes": ib.value,
                "solve_form": ("lane_global", "lane_lds", "wave_pcr")[form.value]}

    def stream(self):
        p = C.c_void_p()
        self._call(self._lib.gm_get_stream, C.byref(p))
        return p.value

    def profile_spmm(self, mode=0, reps=50):
        """Average duration (ms) of one SpMM launch, timed with HIP events on the solver stream."""
        out = C.c_double()
        self._call(self._lib.gm_profile_spmm, int(mode), int(reps), C.byref(out))
        return out.value

    def profile_spmm_flushed(self, mode=0, reps=20, flush_bytes=1 << 30):
        """As profile_spmm, with a streaming copy of ``flush_bytes`` before every launch
        (cold L2 and Infinity Cache); each launch timed alone, latency included."""
        out = C.c_double()
        self._call(self._lib.gm_profile_spmm_flushed, int(mode), int(reps), int(flush_bytes),
                   C.byref(out))
        return out.value

    def profile_copy(self, nbytes, reps=20):
        """Average duration (ms) of a plain device copy of ``nbytes`` bytes."""
        out = C.c_double()
        self._call(self._lib.gm_profile_copy, int(nbytes), int(reps), C.byref(out))
        return out.value
