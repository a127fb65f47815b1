"""``UnevenBatch``: independent problems of different sizes stepped together
(C ABI ``gm_ubatch_*``).

Each problem is what one ``lanczos_create`` call takes (its own matrix, its
own [u, g] pair, contour weights, shifted gammas and shifts); the subspace
size m, the polynomial gamma, the number of contour nodes and dt are shared.
Every problem follows the plain path's step without deflation
(``nicksmethodgpu_problem.m:7-16`` with K = 0, ``update_U.cu:8-16``).
"""

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


def _f(a):
    """Column-major flat float64 copy (MATLAB's layout)."""
    a = np.asarray(a, dtype=np.float64)
    return np.ascontiguousarray(a.reshape(-1, order="F") if a.ndim > 1 else a)


class UnevenBatch:
    def __init__(self, problems, *, m, gamma, dt, clamp=True, autotune=False, wave_solves=False,
                 device=0):
        """``problems``: sequence of dicts with keys A (n x n, symmetric), U
        (n x 2 [u, g(u)]), mass (n or None), wts1, wts2 (n_s x 2 [Re Im]),
        shifts (4 m n_s), gamma_s_real, gamma_s_imag ((d+1) x n_s)."""
        lib = _lib.load()
        self._lib = lib
        self._h = C.c_void_p()
        gamma = np.ascontiguousarray(gamma, dtype=np.float64).ravel()
        nd = gamma.size
        P = len(problems)
        if P < 1:
            raise ValueError("at least one problem")
        ns = np.asarray(problems[0]["gamma_s_real"]).reshape(nd, -1).shape[1]
        keep = {"gamma": gamma}
        vals, ptrs, inds, Us, masses, ns_list = [], [], [], [], [], []
        gsr, gsi, w1, w2, sh = [], [], [], [], []
        sizes, nnzs = [], []
        for pr in problems:
            A = sp.csr_matrix(pr["A"])
            A.sort_indices()
            n = A.shape[0]
            U = np.asarray(pr["U"], dtype=np.float64)
            if U.shape != (n, 2):
                raise ValueError("each U must be n x 2")
            sizes.append(n)
            nnzs.append(A.nnz)
            vals.append(np.ascontiguousarray(A.data, dtype=np.float64))
            ptrs.append(np.ascontiguousarray(A.indptr, dtype=np.int32))
            inds.append(np.ascontiguousarray(A.indices, dtype=np.int32))
            Us.append(_f(U))
            mass = pr.get("mass")
            masses.append(None if mass is None else np.ascontiguousarray(mass, dtype=np.float64))
            g_r = np.asarray(pr["gamma_s_real"], dtype=np.float64).reshape(nd, -1)
            if g_r.shape[1] != ns:
                raise ValueError("every problem needs the same number of contour nodes")
            gsr.append(_f(g_r))
            gsi.append(_f(np.asarray(pr["gamma_s_imag"], dtype=np.float64).reshape(nd, -1)))
            w1.append(_f(np.asarray(pr["wts1"], dtype=np.float64).reshape(ns, 2)))
            w2.append(_f(np.asarray(pr["wts2"], dtype=np.float64).reshape(ns, 2)))
            s = np.ascontiguousarray(pr["shifts"], dtype=np.float64).ravel()
            if s.size != 4 * m * ns:
                raise ValueError("shifts must hold 4 m n_s values")
            sh.append(s)
        keep.update(vals=vals, ptrs=ptrs, inds=inds, Us=Us, masses=masses,
                    n=np.asarray(sizes, dtype=np.int32), nnz=np.asarray(nnzs, dtype=np.int64),
                    gsr=np.concatenate(gsr), gsi=np.concatenate(gsi), w1=np.concatenate(w1),
                    w2=np.concatenate(w2), sh=np.concatenate(sh))
        keep["pv"] = (_dp * P)(*[v.ctypes.data_as(_dp) for v in vals])
        keep["pp"] = (_ip * P)(*[v.ctypes.data_as(_ip) for v in ptrs])
        keep["pi"] = (_ip * P)(*[v.ctypes.data_as(_ip) for v in inds])
        keep["pu"] = (_dp * P)(*[v.ctypes.data_as(_dp) for v in Us])
        have_mass = any(mm is not None for mm in masses)
        if have_mass:
            masses = [np.ones(n) if mm is None else mm for mm, n in zip(masses, sizes)]
            keep["masses"] = masses
            keep["pm"] = (_dp * P)(*[v.ctypes.data_as(_dp) for v in masses])
        d = _lib.GmUbatchDesc()
        d.num_problems = P
        d.n = keep["n"].ctypes.data_as(_ip)
        d.nnz = keep["nnz"].ctypes.data_as(C.POINTER(C.c_int64))
        d.sparse_format = _lib.GM_SPARSE_CSR
        d.A_values = keep["pv"]
        d.A_ptr = keep["pp"]
        d.A_ind = keep["pi"]
        d.U = keep["pu"]
        d.mass = keep["pm"] if have_mass else None
        d.subspace_size = int(m)
        d.gamma_size = nd
        d.gamma = gamma.ctypes.data_as(_dp)
        d.contour_size = ns
        d.gamma_s_real = keep["gsr"].ctypes.data_as(_dp)
        d.gamma_s_imag = keep["gsi"].ctypes.data_as(_dp)
        d.wts1 = keep["w1"].ctypes.data_as(_dp)
        d.wts2 = keep["w2"].ctypes.data_as(_dp)
        d.shifts = keep["sh"].ctypes.data_as(_dp)
        d.dt = float(dt)
        d.flags = ((0 if clamp else _lib.GM_FLAG_NO_CLAMP)
                   | (_lib.GM_FLAG_AUTOTUNE if autotune else 0)
                   | (_lib.GM_FLAG_WAVE_SOLVES if wave_solves else 0))
        d.device = int(device)
        _lib.check(lib.gm_ubatch_create(C.byref(d), C.byref(self._h)), ubatch=True)
        self.sizes = list(sizes)
        self.P = P

    @classmethod
    def from_inputs(cls, inps, **kw):
        """From plain-path ``CreateInputs`` (one per problem) sharing m, gamma and dt."""
        m, gamma, dt = inps[0].m, inps[0].gamma, inps[0].dt
        for inp in inps:
            if inp.m != m or inp.dt != dt or not np.array_equal(inp.gamma, gamma):
                raise ValueError("problems must share m, gamma and dt")
            if inp.Q.shape[1]:
                raise ValueError("the uneven batch has no deflation (K must be 0)")
        probs = [dict(A=inp.A, U=inp.U, mass=inp.mass, wts1=inp.wts1, wts2=inp.wts2,
                      shifts=inp.shifts, gamma_s_real=inp.gamma_s_real,
                      gamma_s_imag=inp.gamma_s_imag) for inp in inps]
        return cls(probs, m=m, gamma=gamma, dt=dt, **kw)

    def _call(self, fn, *args):
        if not self._h:
            raise _lib.GpuMatfunError(_lib.GM_EBADHANDLE, "uneven batch has been deleted")
        _lib.check(fn(self._h, *args), ubatch=True)

    def step(self, nsteps=1, use_graph=False):
        self._call(self._lib.gm_ubatch_step, int(nsteps), int(bool(use_graph)))

    def get_U(self, p):
        """n_p x 2 [u_p, dt g_p] of problem p (as ``get_U.cpp``)."""
        n = self.sizes[p]
        out = np.empty(2 * n)
        self._call(self._lib.gm_ubatch_get_U, int(p), out.ctypes.data_as(_dp))
        return out.reshape(2, n).T.copy()

    def synchronize(self):
        self._call(self._lib.gm_ubatch_synchronize)

    def info(self):
        P = C.c_int32()
        N = C.c_int64()
        G = C.c_int32()
        self._call(self._lib.gm_ubatch_info, C.byref(P), C.byref(N), C.byref(G))
        tile, ib, pre = C.c_int32(), C.c_int32(), C.c_int32()
        self._call(self._lib.gm_ubatch_get_schedule, C.byref(tile), C.byref(ib), C.byref(pre))
        form = C.c_int32()
        self._call(self._lib.gm_ubatch_get_solve_form, C.byref(form))
        return {"num_problems": P.value, "total_rows": N.value, "row_blocks": G.value,
                "spmm_tile": tile.value, "index_bytes": ib.value,
                "early_loads": bool(pre.value),
                "solve_form": ("lane_global", "lane_lds", "wave_pcr")[form.value]}

    def delete(self):
        if self._h:
            _lib.check(self._lib.gm_ubatch_delete(self._h), ubatch=True)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self._lib.gm_ubatch_delete(self._h)
        except Exception:
            pass
