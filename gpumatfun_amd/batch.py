"""``BatchEngine``: many problems sharing one matrix, stepped together (C ABI ``gm_batch_*``).

Problem p has the column pair [u_p, g_p] and follows exactly the plain path's
step without deflation: u' = f1(A) u + dt f2(A) g(u) with g(u) = mass*ub*(1-ub),
ub = u/mass (``update_U.cu:8-16``, the clamp of ``update_U.cu:9`` switchable).
"""

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class BatchEngine:
    """P problems sharing A, stepped together (``gm_batch_*``).

    Optional: deflation by shared eigenpairs (``Q`` n x K, ``exp_lambda``,
    ``phi_lambda``), and ``reaction="ms"`` (Mitchell-Schaeffer, gate ``ode``
    n x P, parameters ``ms`` as for ``LanczosEngine.set_ms_params``) instead
    of Fisher-KPP.

    Schedule: ``chunk_columns`` (columns per pass, 0: the widest that fits),
    ``autotune`` (time the chunk widths and the two SpMM forms at create),
    ``column_slabs`` (without autotune: the SpMM in slabs of 32 columns),
    ``locality_order`` (renumber the nodes inside the engine by the library's
    locality order; inputs and results stay in the caller's order).
    """

    def __init__(self, A, U, *, m, gamma, gamma_s_real, gamma_s_imag, wts1, wts2, etaxi, dt,
                 mass=None, chunk_columns=0, clamp=True, autotune=False, device=0,
                 Q=None, exp_lambda=None, phi_lambda=None, reaction="fisher", ode=None,
                 ms=None, column_slabs=False, locality_order=False):
        lib = _lib.load()
        self._lib = lib
        self._locality = bool(locality_order)
        self._h = C.c_void_p()
        A = sp.csr_matrix(A)
        A.sort_indices()
        U = np.asarray(U, dtype=np.float64)
        n, twoP = U.shape
        if twoP % 2:
            raise ValueError("U must have an even number of columns (pairs [u_p, g_p])")
        gamma = np.ascontiguousarray(gamma, dtype=np.float64).ravel()
        nd = gamma.size
        gsr = np.asarray(gamma_s_real, dtype=np.float64).reshape(nd, -1)
        ns = gsr.shape[1]
        keep = dict(
            val=np.ascontiguousarray(A.data, dtype=np.float64),
            ptr=np.ascontiguousarray(A.indptr, dtype=np.int32),
            ind=np.ascontiguousarray(A.indices, dtype=np.int32),
            U=np.ascontiguousarray(U.reshape(-1, order="F")),
            gamma=gamma,
            gsr=np.ascontiguousarray(gsr.reshape(-1, order="F")),
            gsi=np.ascontiguousarray(np.asarray(gamma_s_imag, dtype=np.float64)
                                     .reshape(nd, -1).reshape(-1, order="F")),
            w1=np.ascontiguousarray(np.asarray(wts1, dtype=np.float64).reshape(ns, 2)
                                    .reshape(-1, order="F")),
            w2=np.ascontiguousarray(np.asarray(wts2, dtype=np.float64).reshape(ns, 2)
                                    .reshape(-1, order="F")),
            ex=np.empty(2 * ns),
            mass=None if mass is None else np.ascontiguousarray(mass, dtype=np.float64),
        )
        et = np.asarray(etaxi, dtype=np.complex128).ravel()
        keep["ex"][0::2] = et.real
        keep["ex"][1::2] = et.imag
        d = _lib.GmBatchDesc()
        d.n = n
        d.sparse_format = _lib.GM_SPARSE_CSR
        d.nnz = A.nnz
        d.A_values = _p(keep["val"])
        d.A_ptr = keep["ptr"].ctypes.data_as(C.POINTER(C.c_int32))
        d.A_ind = keep["ind"].ctypes.data_as(C.POINTER(C.c_int32))
        d.num_problems = twoP // 2
        d.U = _p(keep["U"])
        d.subspace_size = int(m)
        d.gamma_size = nd
        d.gamma = _p(keep["gamma"])
        d.contour_size = ns
        d.gamma_s_real = _p(keep["gsr"])
        d.gamma_s_imag = _p(keep["gsi"])
        d.wts1 = _p(keep["w1"])
        d.wts2 = _p(keep["w2"])
        d.etaxi = _p(keep["ex"])
        d.dt = float(dt)
        d.mass = _p(keep["mass"]) if keep["mass"] is not None else None
        d.chunk_columns = int(chunk_columns)
        d.flags = ((0 if clamp else _lib.GM_FLAG_NO_CLAMP) | (_lib.GM_FLAG_AUTOTUNE if autotune else 0)
                   | (_lib.GM_FLAG_COLUMN_SLABS if column_slabs else 0)
                   | (_lib.GM_FLAG_LOCALITY_ORDER if locality_order else 0))
        d.device = int(device)
        K = 0 if Q is None else np.asarray(Q).reshape(n, -1).shape[1]
        if K:
            keep["Q"] = np.ascontiguousarray(np.asarray(Q, dtype=np.float64).reshape(n, K)
                                             .reshape(-1, order="F"))
            keep["el"] = np.ascontiguousarray(exp_lambda, dtype=np.float64).ravel()
            keep["pl"] = np.ascontiguousarray(phi_lambda, dtype=np.float64).ravel()
            if keep["el"].size != K or keep["pl"].size != K:
                raise ValueError("exp_lambda and phi_lambda need K values")
            d.Q, d.exp_lambda, d.phi_lambda = _p(keep["Q"]), _p(keep["el"]), _p(keep["pl"])
        d.num_eigs = K
        if reaction not in ("fisher", "ms"):
            raise ValueError("reaction must be 'fisher' or 'ms'")
        d.reaction = _lib.GM_REACTION_MS if reaction == "ms" else _lib.GM_REACTION_FISHER
        if reaction == "ms":
            prm = dict(_lib.MS_DEFAULTS)
            prm.update(ms or {})
            d.ms = _lib.GmMsParams(**prm)
            if ode is not None:
                keep["ode"] = np.ascontiguousarray(np.asarray(ode, dtype=np.float64)
                                                   .reshape(n, twoP // 2).reshape(-1, order="F"))
                d.ode = _p(keep["ode"])
        _lib.check(lib.gm_batch_create(C.byref(d), C.byref(self._h)), batch=True)
        self.reaction = reaction
        self.n, self.P = n, twoP // 2

    @classmethod
    def from_inputs(cls, inp, U, **kw):
        """Share A, contour, polynomial and deflation of a plain-path ``CreateInputs``;
        U is n x 2P."""
        if inp.Q.shape[1]:
            kw.setdefault("Q", inp.Q)
            kw.setdefault("exp_lambda", inp.exp_lambda)
            kw.setdefault("phi_lambda", inp.phi_lambda)
        return cls(inp.A, U, m=inp.m, gamma=inp.gamma, gamma_s_real=inp.gamma_s_real,
                   gamma_s_imag=inp.gamma_s_imag, wts1=inp.wts1, wts2=inp.wts2,
                   etaxi=inp.extra["etaxi"], dt=inp.dt, mass=inp.mass, **kw)

    def get_ode(self):
        """n x P Mitchell-Schaeffer gates."""
        out = np.empty(self.n * self.P)
        self._call(self._lib.gm_batch_get_ode, _p(out))
        return out.reshape(self.P, self.n).T.copy()

    def _call(self, fn, *args):
        if not self._h:
            raise _lib.GpuMatfunError(_lib.GM_EBADHANDLE, "batch has been deleted")
        _lib.check(fn(self._h, *args), batch=True)

    def step(self, nsteps=1, use_graph=False):
        self._call(self._lib.gm_batch_step, int(nsteps), int(bool(use_graph)))

    def get_U(self):
        out = np.empty(self.n * 2 * self.P)
        self._call(self._lib.gm_batch_get_U, _p(out))
        return out.reshape(2 * self.P, self.n).T.copy()

    def synchronize(self):
        self._call(self._lib.gm_batch_synchronize)

    def info(self):
        P = C.c_int32()
        rc = C.c_int32()
        self._call(self._lib.gm_batch_info, C.byref(P), C.byref(rc))
        xs = C.c_int32()
        self._call(self._lib.gm_batch_get_column_slabs, C.byref(xs))
        return {"num_problems": P.value, "chunk_columns": rc.value, "column_slabs": bool(xs.value),
                "locality_order": self._locality}

    def stream(self):
        p = C.c_void_p()
        self._call(self._lib.gm_batch_get_stream, C.byref(p))
        return p.value

    def delete(self):
        if self._h:
            _lib.check(self._lib.gm_batch_delete(self._h), batch=True)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self._lib.gm_batch_delete(self._h)
        except Exception:
            pass
