"""Gradient of a monodomain Mitchell-Schaeffer run with respect to the
conductivity ratio (and the initial state), through the time steps, against
central finite differences of the oracle's exact-resolvent forward.

Tolerances: the HIP forward uses m-step Lanczos resolvents (m = 160 on the
24^2 grid, converged to ~1e-11 here), so its end state is held to 1e-9 of the
oracle's exact one; the gradients to 1e-6 relative (finite differences with
step 1e-5 agree with steps 1e-4 and 1e-6 to ~1e-9 on this problem)."""

import numpy as np
import pytest

from oracle import adjoint as oadj
from conftest import rel
from gpumatfun_amd import problems, quadrature as q
from gpumatfun_amd._lib import MS_DEFAULTS
from gpumatfun_amd.monodomain import MonodomainAdjoint

pytestmark = pytest.mark.gpu


def test_ratio_and_initial_state_gradients_vs_finite_differences():
    n, steps, dt, r, m, h = 24, 5, 0.05, 4.0, 160, 1e-5
    A = problems.anisotropic_9pt(n, ratio=r)
    E = problems.anisotropic_9pt_ratio_derivative(n)
    w = np.linalg.eigvalsh(A.toarray())
    c = q.contour_nodes(0.95 * w.min(), 1.05 * w.max(), 24, dt * 1e-3, 1.7)
    X, Y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    u0 = ((X >= 3) & (X < 9) & (Y >= 4) & (Y < 10)).astype(float).reshape(-1, order="F")
    h0 = np.ones(n * n)
    y = 0.3 * np.random.default_rng(0).random(n * n)
    p = dict(MS_DEFAULTS)
    print(f"case monodomain adjoint n={n} steps={steps} m={m}", flush=True)
    md = MonodomainAdjoint(A, m=m, xi=c["xi"], wts1=c["wts1"], wts2=c["wts2"], dt=dt, ms=p)
    uT = md.forward(u0, h0, steps)
    ubar0, hbar0, (gr,) = md.backward(2 * (uT - y), [E])
    md.delete()

    def loss(rr, uu=u0):
        Ar = problems.anisotropic_9pt(n, ratio=rr)
        u, _ = oadj.ms_forward(Ar, uu, h0, steps, c["xi"], c["wts1"], c["wts2"], dt, p)
        return float(np.sum((u - y) ** 2)), u

    L0, u_ref = loss(r)
    print("forward rel err", rel(uT, u_ref))
    assert rel(uT, u_ref) < 1e-9
    fd = (loss(r + h)[0] - loss(r - h)[0]) / (2 * h)
    print("dL/dratio", gr, "fd", fd)
    assert abs(gr - fd) <= 1e-6 * abs(fd)
    for i in np.random.default_rng(1).choice(n * n, 4, replace=False):
        up = u0.copy(); up[i] += h
        um = u0.copy(); um[i] -= h
        fdu = (loss(r, up)[0] - loss(r, um)[0]) / (2 * h)
        print("dL/du0", i, ubar0[i], fdu)
        assert abs(ubar0[i] - fdu) <= 1e-6 * max(abs(fdu), 1e-3 * np.abs(ubar0).max())


def test_deflated_ratio_gradient_vs_finite_differences():
    """The same run with K = 6 deflation pairs and the contour on the rest of
    the spectrum (what the 512^2 workload does): dL/dratio against central
    differences of the oracle's deflated exact-resolvent forward, with the
    eigenpairs recomputed at each perturbed ratio (the derivative of f(A), not
    of a frozen eigenbasis).  m = 60 (the deflated problem converges faster)."""
    n, steps, dt, r, m, h, K = 24, 5, 0.05, 4.0, 60, 1e-5, 6
    eta, alpha = dt * 1e-3, 1.7
    A = problems.anisotropic_9pt(n, ratio=r)
    E = problems.anisotropic_9pt_ratio_derivative(n)
    w, U = np.linalg.eigh(A.toarray())
    c = q.contour_nodes(0.95 * w[K], 1.05 * w.max(), 24, eta, alpha)
    X, Y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    u0 = ((X >= 3) & (X < 9) & (Y >= 4) & (Y < 10)).astype(float).reshape(-1, order="F")
    h0 = np.ones(n * n)
    y = 0.3 * np.random.default_rng(0).random(n * n)
    p = dict(MS_DEFAULTS)
    print(f"case deflated monodomain adjoint n={n} K={K} steps={steps} m={m}", flush=True)
    md = MonodomainAdjoint(A, m=m, xi=c["xi"], wts1=c["wts1"], wts2=c["wts2"], dt=dt, ms=p,
                           Q=U[:, :K], lam=w[:K], eta=eta, alpha=alpha)
    uT = md.forward(u0, h0, steps)
    _, _, (gr,) = md.backward(2 * (uT - y), [E])
    md.delete()

    def loss(rr):
        Ar = problems.anisotropic_9pt(n, ratio=rr)
        wr, Ur = np.linalg.eigh(Ar.toarray())
        defl = (Ur[:, :K], q.f1(wr[:K], eta, alpha), q.f2(wr[:K], eta, alpha))
        u, _ = oadj.ms_forward(Ar, u0, h0, steps, c["xi"], c["wts1"], c["wts2"], dt, p, defl=defl)
        return float(np.sum((u - y) ** 2)), u

    L0, u_ref = loss(r)
    print("forward rel err", rel(uT, u_ref))
    assert rel(uT, u_ref) < 1e-9
    fd = (loss(r + h)[0] - loss(r - h)[0]) / (2 * h)
    print("dL/dratio", gr, "fd", fd)
    assert abs(gr - fd) <= 1e-6 * abs(fd)


def test_activation_time_loss_gradient_vs_finite_differences():
    """The activation-map mismatch L = ||t - t_obs||^2 of the smooth activation
    times (t = dt sum_n s(u_n)), differentiated through the deflated run by
    per-step sources, against central differences of the oracle's deflated
    exact-resolvent forward (eigenpairs recomputed per ratio).  t_obs is the
    map at ratio 3."""
    from gpumatfun_amd.monodomain import activation_times
    n, steps, dt, r, m, h, K = 24, 10, 0.05, 4.0, 60, 1e-5, 6
    eta, alpha = dt * 1e-3, 1.7
    E = problems.anisotropic_9pt_ratio_derivative(n)
    A = problems.anisotropic_9pt(n, ratio=r)
    w, U = np.linalg.eigh(A.toarray())
    c = q.contour_nodes(0.9 * w[K], 1.1 * w.max(), 24, eta, alpha)
    X, Y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    u0 = ((X >= 3) & (X < 9) & (Y >= 4) & (Y < 10)).astype(float).reshape(-1, order="F")
    h0 = np.ones(n * n)
    p = dict(MS_DEFAULTS)

    def oracle_states(rr):
        Ar = problems.anisotropic_9pt(n, ratio=rr)
        wr, Ur = np.linalg.eigh(Ar.toarray())
        defl = (Ur[:, :K], q.f1(wr[:K], eta, alpha), q.f2(wr[:K], eta, alpha))
        return oadj.ms_forward(Ar, u0, h0, steps, c["xi"], c["wts1"], c["wts2"], dt, p,
                               defl=defl, states=True)[2]

    t_obs, _ = activation_times(oracle_states(3.0), dt)
    print(f"case activation-time adjoint n={n} K={K} steps={steps}", flush=True)
    md = MonodomainAdjoint(A, m=m, xi=c["xi"], wts1=c["wts1"], wts2=c["wts2"], dt=dt, ms=p,
                           Q=U[:, :K], lam=w[:K], eta=eta, alpha=alpha)
    md.forward(u0, h0, steps)
    t, sources = activation_times(md.states(), dt)
    srcs = sources(2 * (t - t_obs))
    _, _, (gr,) = md.backward(srcs[-1], [E], sources=srcs[:-1])
    md.delete()

    def loss(rr):
        return float(np.sum((activation_times(oracle_states(rr), dt)[0] - t_obs) ** 2))

    fd = (loss(r + h) - loss(r - h)) / (2 * h)
    print("dL/dratio", gr, "fd", fd, "L", float(np.sum((t - t_obs) ** 2)))
    assert abs(gr - fd) <= 1e-6 * abs(fd)


def test_resident_run_matches_host_run():
    """The device-resident mode (states, trajectory and reaction arithmetic as
    float64 torch tensors on the GPU, engines reading and writing GPU memory)
    against the host mode: the same engines and arithmetic, so the end state
    and the gradients agree to rounding of the elementwise reaction terms."""
    import torch
    n, steps, dt, r, m, K = 24, 6, 0.05, 4.0, 60, 6
    eta, alpha = dt * 1e-3, 1.7
    E = problems.anisotropic_9pt_ratio_derivative(n)
    A = problems.anisotropic_9pt(n, ratio=r)
    w, U = np.linalg.eigh(A.toarray())
    c = q.contour_nodes(0.9 * w[K], 1.1 * w.max(), 24, eta, alpha)
    X, Y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    u0 = ((X >= 3) & (X < 9) & (Y >= 4) & (Y < 10)).astype(float).reshape(-1, order="F")
    h0 = np.ones(n * n)
    y = 0.3 * np.random.default_rng(0).random(n * n)
    out = {}
    for resident in (False, True):
        print(f"case resident={resident}", flush=True)
        md = MonodomainAdjoint(A, m=m, xi=c["xi"], wts1=c["wts1"], wts2=c["wts2"], dt=dt,
                               Q=U[:, :K], lam=w[:K], eta=eta, alpha=alpha, resident=resident)
        uT = md.forward(u0, h0, steps)
        if resident:
            assert isinstance(uT, torch.Tensor) and uT.is_cuda and uT.dtype == torch.float64
            uT = uT.cpu().numpy()
        ub0, hb0, (gr,) = md.backward(2 * (uT - y), [E])
        if resident:
            ub0, hb0 = ub0.cpu().numpy(), hb0.cpu().numpy()
        md.delete()
        out[resident] = (uT, ub0, hb0, gr)
    a, b = out[False], out[True]
    print("uT", rel(b[0], a[0]), "ubar0", rel(b[1], a[1]), "dL/dr", a[3], b[3])
    assert rel(b[0], a[0]) < 1e-13 and rel(b[1], a[1]) < 1e-12 and rel(b[2], a[2]) < 1e-12
    assert abs(b[3] - a[3]) <= 1e-12 * abs(a[3])


def test_stimulus_set_concurrent_matches_sequential():
    """Three stimuli run concurrently on two host threads (own torch and engine
    streams) give bitwise the per-stimulus results of sequential runs, and the
    set's gradient is their sum."""
    import torch
    from gpumatfun_amd.monodomain import StimulusSetAdjoint
    n, steps, dt, r, m, K = 24, 4, 0.05, 4.0, 60, 6
    eta, alpha = dt * 1e-3, 1.7
    E = problems.anisotropic_9pt_ratio_derivative(n)
    A = problems.anisotropic_9pt(n, ratio=r)
    w, U = np.linalg.eigh(A.toarray())
    c = q.contour_nodes(0.9 * w[K], 1.1 * w.max(), 24, eta, alpha)
    X, Y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    u0s = [((X >= a) & (X < a + 5) & (Y >= b) & (Y < b + 5)).astype(float).reshape(-1, order="F")
           for a, b in ((2, 3), (12, 5), (7, 15))]
    h0 = np.ones(n * n)
    y = 0.3 * np.random.default_rng(0).random(n * n)
    kw = dict(m=m, xi=c["xi"], wts1=c["wts1"], wts2=c["wts2"], dt=dt, Q=U[:, :K], lam=w[:K],
              eta=eta, alpha=alpha)
    print("case stimulus set P=3 workers=2", flush=True)
    seq = []
    for u0 in u0s:
        md = MonodomainAdjoint(A, resident=True, **kw)
        uT = md.forward(u0, h0, steps)
        ub, hb, (gr,) = md.backward(2 * (uT - torch.as_tensor(y, device=uT.device)), [E])
        seq.append((uT.cpu().numpy(), ub.cpu().numpy(), gr))
        md.delete()
    st = StimulusSetAdjoint(A, len(u0s), workers=2, **kw)
    uTs = st.forward(u0s, h0, steps)
    yt = torch.as_tensor(y, device=uTs[0].device)
    ubs, hbs, (gsum,) = st.backward([2 * (u - yt) for u in uTs], [E])
    st.delete()
    for p in range(len(u0s)):
        assert np.array_equal(uTs[p].cpu().numpy(), seq[p][0])
        assert np.array_equal(ubs[p].cpu().numpy(), seq[p][1])
    assert gsum == sum(s_[2] for s_ in seq) or abs(gsum - sum(s_[2] for s_ in seq)) <= 1e-15 * abs(gsum)


@pytest.mark.parametrize("P", [1, 2, 4, 16])
def test_batch_adjoint_matches_per_stimulus(P):
    """The one-launch multi-stimulus engine (gm_bfab_*, columns [u_p, g_p]) against
    the single-column engine run stimulus by stimulus (itself pinned by central
    differences above): end states, dL/du_0 and the summed dL/dratio, with K = 6
    deflation pairs.  The two differ only in the reduction partitions of the
    Lanczos sums, so they agree to rounding amplified by m = 60 steps."""
    import torch
    from gpumatfun_amd.monodomain import MonodomainBatchAdjoint
    n, steps, dt, r, m, K = 24, 5, 0.05, 4.0, 60, 6
    eta, alpha = dt * 1e-3, 1.7
    E = problems.anisotropic_9pt_ratio_derivative(n)
    A = problems.anisotropic_9pt(n, ratio=r)
    w, U = np.linalg.eigh(A.toarray())
    c = q.contour_nodes(0.9 * w[K], 1.1 * w.max(), 24, eta, alpha)
    X, Y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    spots = ([(2, 3), (12, 5), (7, 15), (15, 14)] if P <= 4 else
             [(1 + 5 * (k % 4), 1 + 5 * (k // 4)) for k in range(P)])[:P]
    u0s = [((X >= a) & (X < a + 5) & (Y >= b) & (Y < b + 5)).astype(float).reshape(-1, order="F")
           for a, b in spots]
    h0 = np.ones(n * n)
    y = 0.3 * np.random.default_rng(0).random(n * n)
    kw = dict(m=m, xi=c["xi"], wts1=c["wts1"], wts2=c["wts2"], dt=dt, Q=U[:, :K], lam=w[:K],
              eta=eta, alpha=alpha)
    print(f"case batch adjoint P={P}", flush=True)
    ref_u, ref_ub, ref_g = [], [], 0.0
    for u0 in u0s:
        md = MonodomainAdjoint(A, resident=True, **kw)
        uT = md.forward(u0, h0, steps)
        ub, _, (gr,) = md.backward(2 * (uT - torch.as_tensor(y, device=uT.device)), [E])
        ref_u.append(uT.cpu().numpy())
        ref_ub.append(ub.cpu().numpy())
        ref_g += gr
        md.delete()
    bm = MonodomainBatchAdjoint(A, P, **kw)
    UT = bm.forward(u0s, h0, steps)
    yt = torch.as_tensor(y, device=UT.device)[:, None]
    UB, _, (gb,) = bm.backward(2 * (UT - yt), [E])
    bm.delete()
    for p_ in range(P):
        eu = rel(UT[:, p_].cpu().numpy(), ref_u[p_])
        eb = rel(UB[:, p_].cpu().numpy(), ref_ub[p_])
        print(f"stimulus {p_}: u_T {eu:.2e}, dL/du0 {eb:.2e}")
        assert eu < 1e-10 and eb < 1e-9
    print("dL/dratio", gb, ref_g)
    assert abs(gb - ref_g) <= 1e-9 * abs(ref_g)
