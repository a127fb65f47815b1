"""Gradient of a monodomain Mitchell-Schaeffer run with respect to a parameter
of the diffusion operator (e.g. the conductivity ratio), by the discrete
adjoint of the exponential-integrator time stepping (no reference
counterpart).

Forward scheme, per stimulus -- the plain path's step with the
Mitchell-Schaeffer reaction (``gm_step`` mode 2), f(A) applied without
deflation:

    u_{n+1} = f1(A) u_n + dt f2(A) g_n,           g_n = G(u_n, h_n)
    h_{n+1} = h_n + dt D(u_{n+1}, h_n)
    G(u, h) = h u^2 (1 - u)/tau_in - u/tau_out
    D(u, h) = (1 - h)/tau_open if u < v_gate else -h/tau_close

with f1 = exp(-eta x^(alpha/2)), f2 = phi(-eta x^(alpha/2)) on the contour of
``problem_startup.m:85-103``.  For L = ||u_T - y||^2 (or any loss of the
states, through per-step sources, e.g. a smooth activation-time map) the
adjoint runs the steps backwards (A symmetric, so f(A)^T = f(A)):

    w       = ubar_{n+1}                       (D is piecewise constant in u)
    ubar_n  = f1(A) w + dG/du(u_n, h_n) * (dt f2(A) w)
    hbar_n  = hbar_{n+1} (1 + dt dD/dh) + dG/dh(u_n, h_n) * (dt f2(A) w)
    dL/dtheta += w^T (d f1(A)/dtheta) u_n + dt w^T (d f2(A)/dtheta) g_n

where each w^T (d f(A)/dtheta) v is the contour sum Im sum_j c_j z_j^T
(dA/dtheta) x_j of ``gm_fab_backward_param``.  Each backward step
re-runs the two forward Krylov solves it needs, from the stored states.

With deflation (K eigenpairs Q, Lambda of A) f(A) v = Q f(Lambda) Q^T v +
f(A)(I - QQ^T) v as in the forward engines, and the derivative of f(A) is
taken exactly over that split (divided differences of f on the Q-Q block,
real-shift resolvents (lambda_k - A)^{-1} on the cross blocks, the contour on
the rest), so dQ/dtheta and dLambda/dtheta never enter.
"""

import numpy as np

from . import _lib
from .adjoint import FabEngine


def _xp(a):
    """The array module of a: numpy, or torch for tensors."""
    if hasattr(a, "data_ptr"):
        import torch
        return torch
    return np


def _where(cond, a, b, like):
    """where(cond, a, b) with a, b arrays or Python floats, in like's dtype (torch
    would otherwise turn float scalars into its default float32)."""
    if _xp(like) is np:
        return np.where(cond, a, b)
    import torch

    def t(v):
        return v if hasattr(v, "data_ptr") else torch.tensor(v, dtype=like.dtype,
                                                            device=like.device)
    return torch.where(cond, t(a), t(b))


def ms_g(u, h, p):
    return (h * (u * u) * (1.0 - u)) / p["tau_in"] - u / p["tau_out"]


def ms_gate(u, h, dt, p):
    dh = _where(u < p["v_gate"], (1.0 - h) / p["tau_open"], -h / p["tau_close"], u)
    return h + dt * dh


def activation_times(us, dt, v_thr=0.5, eps=0.05):
    """Smooth activation-time map of a run u_1 .. u_T: t(x) = dt sum_n s(u_n(x)),
    s(u) = 1 - sigma((u - v_thr) / eps), the time a node spends below v_thr (its
    activation time when it activates once within the window).  Returns t and, for
    dL/dt = r, the direct sources dL/du_n = r * dt * s'(u_n) via ``sources(r)``."""
    xp = _xp(us[0])

    def sig(z):
        return 0.5 * (1.0 + xp.tanh(0.5 * z))

    t = xp.zeros_like(us[0])
    for u in us:
        t += dt * (1.0 - sig((u - v_thr) / eps))

    def one(u, r):
        sg = sig((u - v_thr) / eps)
        return -r * dt * sg * (1.0 - sg) / eps

    def sources(r):
        return _Lazy(us, lambda u: one(u, r))

    return t, sources


class _Lazy:
    """A read-only sequence f(items[i]) evaluated on access (slices stay lazy), so
    a loss's per-step sources take no memory until the backward pass reads them."""

    def __init__(self, items, f):
        self.items, self.f = items, f

    def __len__(self):
        return len(self.items)

    def __getitem__(self, i):
        if isinstance(i, slice):
            return _Lazy(self.items[i], self.f)
        return self.f(self.items[i])


class MonodomainAdjoint:
    """Forward steps and the adjoint for one stimulus on a fixed contour.

    ``xi`` (complex nodes), ``wts1``/``wts2`` (complex s f1(xi) dxi/dt and
    s f2(xi) dxi/dt), subspace size ``m`` for the Krylov resolvents.
    """

    def __init__(self, A, *, m, xi, wts1, wts2, dt, ms=None, device=0, Q=None, lam=None,
                 eta=None, alpha=None, resident=False):
        """With ``Q``, ``lam`` (K eigenpairs of A) the solves are deflated as in
        the forward engines; ``eta`` and ``alpha`` then give the exact f1, f2 and
        their divided differences at ``lam`` (phi is taken exactly at every
        eigenvalue, without the phi_lambda(1) = 1 guard of the set-up).  With
        ``resident`` the states, the trajectory and the reaction arithmetic stay
        in GPU memory (float64 torch tensors), and the engines read and write
        them there."""
        self.resident = bool(resident)
        self.device = int(device)
        self.p = dict(_lib.MS_DEFAULTS)
        self.p.update(ms or {})
        self.dt = float(dt)
        kw1, kw2 = {}, {}
        if Q is not None and np.shape(Q)[1] > 0:
            if eta is None or alpha is None:
                raise ValueError("deflation needs eta and alpha")
            from . import quadrature as qd
            lam = np.asarray(lam, dtype=np.float64)
            fa = (lambda x: qd.f1(x, eta, alpha), lambda x: qd.df1(x, eta, alpha))
            fb = (lambda x: qd.f2(x, eta, alpha), lambda x: qd.df2(x, eta, alpha))
            kw1 = dict(Q=Q, lam=lam, f_lambda=fa[0](lam),
                       f_divdiff=qd.divided_differences(fa[0], fa[1], lam))
            kw2 = dict(Q=Q, lam=lam, f_lambda=fb[0](lam),
                       f_divdiff=qd.divided_differences(fb[0], fb[1], lam))
        self.f1 = FabEngine(A, m=m, xi=xi, wts=wts1, recompute=False, device=device, **kw1)
        self.f2 = FabEngine(A, m=m, xi=xi, wts=wts2, recompute=False, device=device, **kw2)
        self.traj = None

    def set_matrix(self, A):
        self.f1.set_matrix(A)
        self.f2.set_matrix(A)

    def _arr(self, x):
        if self.resident:
            import torch
            return torch.as_tensor(np.asarray(x, dtype=np.float64) if not hasattr(x, "data_ptr")
                                   else x, dtype=torch.float64,
                                   device=f"cuda:{self.device}").clone()
        return np.array(x, dtype=np.float64)

    def forward(self, u0, h0, steps):
        """Runs the steps; returns u_T and keeps (u_n, h_n, g_n) for the adjoint."""
        u = self._arr(u0)
        h = self._arr(h0)
        traj = []
        for _ in range(steps):
            g = ms_g(u, h, self.p)
            traj.append((u, h, g))
            u = self.f1.forward(u) + self.dt * self.f2.forward(g)
            h = ms_gate(u, h, self.dt, self.p)
        self.traj = traj
        self.uT, self.hT = u, h
        return u

    def states(self):
        """u_1 .. u_T of the stored run (views, not copies)."""
        if self.traj is None:
            raise RuntimeError("no stored run")
        return [t[0] for t in self.traj[1:]] + [self.uT]

    def backward(self, ubar_T, dA_list=(), sources=None):
        """Adjoint of the stored run for dL/du_T = ubar_T: returns (dL/du_0,
        dL/dh_0, [dL/dtheta_q]) for the given dA/dtheta_q.  ``sources`` (a list
        of T - 1 arrays, or None) adds a loss's direct dL/du_n for n = 1 .. T-1,
        e.g. of a loss over the whole trajectory such as activation times."""
        if self.traj is None:
            raise RuntimeError("backward needs a forward run first")
        p, dt = self.p, self.dt
        w = self._arr(ubar_T)
        xp = _xp(w)
        hbar = xp.zeros_like(w)
        gtheta = np.zeros(len(dA_list))
        for eng in (self.f1, self.f2):      # dA/dtheta kept on the device for every step
            eng.set_param_derivs(list(dA_list))
        for n in range(len(self.traj) - 1, -1, -1):
            u, h, g = self.traj[n]
            u_next = self.traj[n + 1][0] if n + 1 < len(self.traj) else self.uT
            # h_{n+1} = h_n + dt D(u_{n+1}, h_n): hbar flows to h_n, nothing to u_{n+1}
            dDdh = _where(u_next < p["v_gate"], -1.0 / p["tau_open"], -1.0 / p["tau_close"], w)
            hbar_prev = hbar * (1.0 + dt * dDdh)
            # u_{n+1} = f1(A) u_n + dt f2(A) g_n
            self.f1.forward(u, want_u=False)
            f1w, gt1 = self.f1.backward_param(w, None, want_b=True)
            self.f2.forward(g, want_u=False)
            f2w, gt2 = self.f2.backward_param(w, None, want_b=True)
            gtheta += gt1 + dt * gt2
            dGdu = h * (2.0 * u - 3.0 * u * u) / p["tau_in"] - 1.0 / p["tau_out"]
            dGdh = (u * u) * (1.0 - u) / p["tau_in"]
            w = f1w + dGdu * (dt * f2w)
            hbar = hbar_prev + dGdh * (dt * f2w)
            if sources is not None and n >= 1:
                w = w + sources[n - 1]
        return w, hbar, gtheta

    def delete(self):
        self.f1.delete()
        self.f2.delete()
        self.traj = None


class StimulusSetAdjoint:
    """Several stimuli of one operator: one device-resident MonodomainAdjoint per
    stimulus, run concurrently on ``workers`` host threads, each issuing on its
    own torch stream and its engines' streams, so that the small single-column
    kernels of different stimuli overlap on the GPU.  The loss is the sum over
    stimuli, so the parameter gradient is the sum of the per-stimulus ones."""

    def __init__(self, A, n_stimuli, *, workers=4, **kw):
        import torch
        kw = dict(kw)
        kw["resident"] = True
        self.runs = [MonodomainAdjoint(A, **kw) for _ in range(n_stimuli)]
        self.workers = max(1, min(int(workers), n_stimuli))
        dev = f"cuda:{int(kw.get('device', 0))}"
        self.streams = [torch.cuda.Stream(device=dev) for _ in range(self.workers)]

    def _map(self, fn):
        """fn(p, run) for every stimulus, stimulus p on worker p % workers (in
        order within a worker); returns the results in stimulus order."""
        import torch
        from concurrent.futures import ThreadPoolExecutor
        torch.cuda.current_stream().synchronize()   # the caller's inputs are ready
        out = [None] * len(self.runs)

        def work(w):
            with torch.cuda.stream(self.streams[w]):
                for p in range(w, len(self.runs), self.workers):
                    out[p] = fn(p, self.runs[p])
                torch.cuda.current_stream().synchronize()

        if self.workers == 1:
            work(0)
        else:
            with ThreadPoolExecutor(self.workers) as ex:
                for f in [ex.submit(work, w) for w in range(self.workers)]:
                    f.result()
        return out

    def forward(self, u0s, h0, steps):
        """u_T of every stimulus (u0s: one initial state per stimulus)."""
        return self._map(lambda p, r: r.forward(u0s[p], h0, steps))

    def states(self, p):
        return self.runs[p].states()

    def backward(self, ubar_Ts, dA_list=(), sources=None):
        """Sum over stimuli of dL/dtheta (and the per-stimulus dL/du_0, dL/dh_0)."""
        res = self._map(lambda p, r: r.backward(ubar_Ts[p], dA_list,
                                                None if sources is None else sources[p]))
        return [r[0] for r in res], [r[1] for r in res], sum(r[2] for r in res)

    def delete(self):
        for r in self.runs:
            r.delete()


class MonodomainBatchAdjoint:
    """All stimuli in one multi-column engine (``gm_bfab_*``): the columns are
    [u_1, g_1, u_2, g_2, ...] (even columns f1, odd f2), so each stage of a
    forward step -- and of a backward step, on [w_p, w_p] -- is one launch over
    every stimulus.  States, trajectory and reaction arithmetic are float64
    torch tensors in GPU memory, N x P.  Same scheme and adjoint as
    :class:`MonodomainAdjoint`; P in {1, 2, 4, 8, 16, 32}."""

    def __init__(self, A, n_stimuli, *, m, xi, wts1, wts2, dt, ms=None, device=0, Q=None,
                 lam=None, eta=None, alpha=None):
        from .adjoint import BatchFabEngine
        self.P = int(n_stimuli)
        if 2 * self.P not in (2, 4, 8, 16, 32, 64):
            raise ValueError("n_stimuli must be 1, 2, 4, 8, 16 or 32")
        self.p = dict(_lib.MS_DEFAULTS)
        self.p.update(ms or {})
        self.dt = float(dt)
        self.device = int(device)
        kw = {}
        if Q is not None and np.shape(Q)[1] > 0:
            if eta is None or alpha is None:
                raise ValueError("deflation needs eta and alpha")
            from . import quadrature as qd
            lam = np.asarray(lam, dtype=np.float64)
            f1 = (lambda x: qd.f1(x, eta, alpha))
            f2 = (lambda x: qd.f2(x, eta, alpha))
            kw = dict(Q=Q, lam=lam, f1_lambda=f1(lam), f2_lambda=f2(lam),
                      f1_divdiff=qd.divided_differences(f1, lambda x: qd.df1(x, eta, alpha), lam),
                      f2_divdiff=qd.divided_differences(f2, lambda x: qd.df2(x, eta, alpha), lam))
        self.eng = BatchFabEngine(A, R=2 * self.P, m=m, xi=xi, wts1=wts1, wts2=wts2,
                                  device=device, **kw)
        self.traj = None

    def _pairs(self, a, b):
        import torch
        return torch.stack((a, b), dim=2).reshape(a.shape[0], 2 * self.P)

    def _cols(self, xs):
        import torch
        return torch.stack([torch.as_tensor(np.asarray(x, dtype=np.float64) if not
                                            hasattr(x, "data_ptr") else x, dtype=torch.float64,
                                            device=f"cuda:{self.device}") for x in xs], dim=1)

    def forward(self, u0s, h0, steps):
        """u_T (N x P) from the P initial states u0s and the common gate h0."""
        U = self._cols(u0s).contiguous()
        H = self._cols([h0] * self.P).contiguous()
        traj = []
        for _ in range(steps):
            G = ms_g(U, H, self.p)
            traj.append((U, H, G))
            out = self.eng.forward(self._pairs(U, G))
            U = (out[:, 0::2] + self.dt * out[:, 1::2]).contiguous()
            H = ms_gate(U, H, self.dt, self.p)
        self.traj = traj
        self.uT, self.hT = U, H
        return U

    def states(self, p):
        """u_1 .. u_T of stimulus p."""
        if self.traj is None:
            raise RuntimeError("no stored run")
        return [t[0][:, p] for t in self.traj[1:]] + [self.uT[:, p]]

    def backward(self, ubar_T, dA_list=(), sources=None):
        """(dL/du_0, dL/dh_0 (N x P), sum over stimuli of dL/dtheta_q); ``sources``
        a list over stimuli of per-step sources (as in MonodomainAdjoint)."""
        import torch
        if self.traj is None:
            raise RuntimeError("backward needs a forward run first")
        p, dt = self.p, self.dt
        w = torch.as_tensor(ubar_T, dtype=torch.float64,
                            device=f"cuda:{self.device}").clone().contiguous()
        hbar = torch.zeros_like(w)
        gtheta = np.zeros(len(dA_list))
        self.eng.set_param_derivs(list(dA_list))
        for n in range(len(self.traj) - 1, -1, -1):
            U, H, G = self.traj[n]
            u_next = self.traj[n + 1][0] if n + 1 < len(self.traj) else self.uT
            dDdh = _where(u_next < p["v_gate"], -1.0 / p["tau_open"], -1.0 / p["tau_close"], w)
            hbar_prev = hbar * (1.0 + dt * dDdh)
            self.eng.forward(self._pairs(U, G), want_u=False)
            fw, gt = self.eng.backward_param(self._pairs(w, w), want_b=True)
            if len(dA_list):
                gtheta += gt[:, 0::2].sum(axis=1) + dt * gt[:, 1::2].sum(axis=1)
            f1w, f2w = fw[:, 0::2], fw[:, 1::2]
            dGdu = H * (2.0 * U - 3.0 * U * U) / p["tau_in"] - 1.0 / p["tau_out"]
            dGdh = (U * U) * (1.0 - U) / p["tau_in"]
            w = (f1w + dGdu * (dt * f2w)).contiguous()
            hbar = hbar_prev + dGdh * (dt * f2w)
            if sources is not None and n >= 1:
                w = w + torch.stack([src[n - 1] for src in sources], dim=1)
        return w, hbar, gtheta

    def delete(self):
        self.eng.delete()
        self.traj = None
