"""The adjoint (gm_fab_*) against the CPU oracle and central finite differences.

Tolerances:
* same Krylov approximation on both sides (oracle/adjoint.krylov_gradients),
  small well-conditioned grid: relative 2-norm <= 1e-10;
* against exact resolvents (sparse LU), 128^2 and 256^2 GRF-kappa problems: the Krylov
  error of the m used (measured on the CPU: 7.5e-7 at m = 600 on the 256^2
  grid), so <= 1e-5 for the forward and the gradients;
* central differences, step 1e-6 on kappa_i, of the exact (sparse LU)
  forward on a 48^2 grid: <= 1e-4 relative per entry (floor 1e-2 max|grad|);
* central differences of the HIP forward itself on the 256^2 problem: step
  1e-2, because that forward (Lanczos far past the loss of orthogonality)
  carries ~2e-12 relative rounding noise between neighbouring kappa, which at
  a step of 1e-6 would exceed the difference being measured; <= 1e-3.
"""

import numpy as np
import pytest

from oracle import adjoint as oadj
from conftest import rel
from gpumatfun_amd import problems, quadrature as q
from gpumatfun_amd.adjoint import FabEngine, FractionalDiffusion, contour_weights

pytestmark = pytest.mark.gpu
TAU, ALPHA = 1e-3, 1.5


def problem(n, seed=1):
    kappa = np.exp(0.5 * problems.gaussian_random_field(n, 0.1, seed=seed))
    A = problems.varcoef_laplacian_2d(kappa, n)
    return kappa, A


def bounds(A):
    import scipy.sparse.linalg as sla
    if A.shape[0] <= 2000:
        w = np.linalg.eigvalsh(A.toarray())
        return w.min(), w.max()
    lo = sla.eigsh(A, k=1, sigma=0, which="LM", return_eigenvectors=False)[0]
    hi = sla.eigsh(A, k=1, which="LA", return_eigenvectors=False)[0]
    return lo, hi


def contour(A, ns=32, alpha=ALPHA):
    lo, hi = bounds(A)
    c = q.contour_nodes(0.95 * lo, 1.05 * hi, ns, TAU, 1.0)
    cw, dc = contour_weights(c["xi"], c["dxidt"], c["s"], TAU, alpha)
    return c, cw, dc


def test_adjoint_matches_krylov_consistent_oracle():
    n, m = 16, 24
    kappa, A = problem(n)
    c, cw, dc = contour(A, ns=24)
    rng = np.random.default_rng(0)
    b = rng.standard_normal(n * n)
    y = rng.standard_normal(n * n)
    print(f"case adjoint n={n} m={m}", flush=True)
    eng = FabEngine(A, m=m, xi=c["xi"], wts=cw, dwts=dc)
    u = eng.forward(b)
    X = oadj.krylov_resolvents(A, c["xi"], b, m)
    assert rel(u, np.imag(X @ cw)) < 1e-10
    g = 2 * (u - y)
    gb, gk, ga = eng.backward(g, kappa)
    rb, rk, ra = oadj.krylov_gradients(A, kappa, b, g, c["xi"], cw, dc, m)
    print("rel", rel(gb, rb), rel(gk, rk), abs(ga - ra) / abs(ra), flush=True)
    assert rel(gb, rb) < 1e-10
    assert rel(gk, rk) < 1e-10
    assert abs(ga - ra) <= 1e-10 * abs(ra)
    eng.delete()


def test_recompute_basis_is_bitwise_identical():
    n, m = 16, 24
    kappa, A = problem(n)
    c, cw, dc = contour(A, ns=24)
    b = np.random.default_rng(2).standard_normal(n * n)
    g = np.random.default_rng(3).standard_normal(n * n)
    e1 = FabEngine(A, m=m, xi=c["xi"], wts=cw, dwts=dc)
    e2 = FabEngine(A, m=m, xi=c["xi"], wts=cw, dwts=dc, recompute=True)
    e1.forward(b)
    e2.forward(b)
    r1 = e1.backward(g, kappa)
    r2 = e2.backward(g, kappa)
    assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1]) and r1[2] == r2[2]
    e1.delete()
    e2.delete()


def test_adjoint_against_exact_resolvents_128():
    n, m = 128, 400
    kappa, A = problem(n)
    c, cw, dc = contour(A, ns=32)
    rng = np.random.default_rng(4)
    b = rng.standard_normal(n * n)
    y = rng.standard_normal(n * n)
    eng = FabEngine(A, m=m, xi=c["xi"], wts=cw, dwts=dc)
    u = eng.forward(b)
    ue, _ = oadj.forward(A, b, c["xi"], cw)
    g = 2 * (u - y)
    gb, gk, ga = eng.backward(g, kappa)
    rb, rk, ra = oadj.gradients(A, kappa, b, g, c["xi"], cw, dc)
    e = (rel(u, ue), rel(gb, rb), rel(gk, rk), abs(ga - ra) / abs(ra))
    print("rel vs exact resolvents (u, b, kappa, alpha):", e, flush=True)
    assert max(e) < 1e-5
    eng.delete()


def test_kappa_gradient_central_differences_step_1e6():
    """Central differences with step 1e-6 on 64 random kappa entries of
    L = ||f(A)b - y||^2, differencing the exact forward (sparse LU per node,
    smooth to rounding) against the HIP gradient.  48^2 grid so that the
    128 perturbed forwards stay cheap; m = 300 converges the Krylov side to
    ~1e-10 here."""
    import scipy.sparse.linalg  # noqa: F401  (splu inside oracle.adjoint)
    n, m = 48, 300
    kappa, A = problem(n)
    c, cw, dc = contour(A, ns=24)
    rng = np.random.default_rng(1)
    b = rng.standard_normal(n * n)
    y = rng.standard_normal(n * n)
    print(f"case FD step 1e-6 n={n} m={m}", flush=True)
    eng = FabEngine(A, m=m, xi=c["xi"], wts=cw, dwts=dc)
    u = eng.forward(b)
    gb, gk, ga = eng.backward(2 * (u - y), kappa)
    eng.delete()
    gmax = np.abs(gk).max()
    worst = 0.0
    h = 1e-6
    for i in rng.choice(n * n, 64, replace=False):
        kp = kappa.copy(); kp[i] += h
        km = kappa.copy(); km[i] -= h
        up, _ = oadj.forward(problems.varcoef_laplacian_2d(kp, n), b, c["xi"], cw)
        um, _ = oadj.forward(problems.varcoef_laplacian_2d(km, n), b, c["xi"], cw)
        # L+ - L- = (u+ - u-) . (u+ + u- - 2y), formed without cancelling L
        fd = np.dot(up - um, up + um - 2 * y) / (2 * h)
        worst = max(worst, abs(fd - gk[i]) / max(abs(gk[i]), 1e-2 * gmax))
    print(f"worst FD relative error over 64 entries: {worst:.3e} (max |grad| {gmax:.3e})", flush=True)
    assert worst < 1e-4


def test_grf256_gradients_vs_exact_reference_and_fd():
    """The 256^2 GRF-conductivity problem (kappa = exp(0.5 GRF), m = 600, 32 nodes):
    all three gradients against the FP64 CPU reference gradient built from
    exact resolvents (tolerance: the Krylov error of m = 600, measured
    7.5e-7 on the CPU), and kappa-gradient central differences of the HIP
    forward itself on 64 random entries.  That forward runs Lanczos far past
    the loss of orthogonality, so it is reproducible but carries ~2e-12
    relative rounding noise from one kappa to the next; a step of 1e-2 keeps
    that noise (and the O(h^2) truncation) under 1e-3 of the gradient."""
    n, m = 256, 600
    kappa, A = problem(n)
    c, cw, dc = contour(A, ns=32)
    rng = np.random.default_rng(1)
    b = rng.standard_normal(n * n)
    y = rng.standard_normal(n * n)
    print(f"case grf256 n={n} m={m}", flush=True)
    eng = FabEngine(A, m=m, xi=c["xi"], wts=cw, dwts=dc)
    u = eng.forward(b)
    g = 2 * (u - y)
    gb, gk, ga = eng.backward(g, kappa)
    rb, rk, ra = oadj.gradients(A, kappa, b, g, c["xi"], cw, dc)
    e = (rel(gb, rb), rel(gk, rk), abs(ga - ra) / abs(ra))
    print("rel vs exact reference (b, kappa, alpha):", e, flush=True)
    assert max(e) < 1e-5
    gmax = np.abs(gk).max()
    worst = 0.0
    h = 1e-2
    for i in rng.choice(n * n, 64, replace=False):
        kp = kappa.copy(); kp[i] += h
        km = kappa.copy(); km[i] -= h
        eng.set_matrix(problems.varcoef_laplacian_2d(kp, n))
        up = eng.forward(b)
        eng.set_matrix(problems.varcoef_laplacian_2d(km, n))
        um = eng.forward(b)
        fd = np.dot(up - um, up + um - 2 * y) / (2 * h)
        worst = max(worst, abs(fd - gk[i]) / max(abs(gk[i]), 1e-1 * gmax))
    print(f"worst FD (step 1e-2) relative error: {worst:.3e} (max |grad| {gmax:.3e})", flush=True)
    assert worst < 1e-3
    eng.delete()


def test_torch_autograd_function_gradcheck():
    import torch
    n = 6
    kappa, A = problem(n)
    lo, hi = bounds(A)
    op = FractionalDiffusion(n, tau=TAU, l1=0.8 * lo, lN=1.3 * hi, ns=40, m=n * n)
    b = torch.tensor(np.random.default_rng(0).standard_normal(n * n), requires_grad=True)
    k = torch.tensor(kappa, requires_grad=True)
    a = torch.tensor(ALPHA, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda b_, k_, a_: op.apply(b_, k_, a_), (b, k, a),
                                    eps=1e-6, atol=1e-6, rtol=1e-4)


def test_param_gradient_conductivity_ratio_vs_exact_reference():
    """dL/dr for the conductivity ratio of the anisotropic 9-point operator
    (gm_fab_backward_param) against the exact-resolvent oracle, which central
    differences pin (tests/test_adjoint_oracle.py)."""
    n, m, r = 40, 300, 4.0
    A = problems.anisotropic_9pt(n, ratio=r)
    E = problems.anisotropic_9pt_ratio_derivative(n)
    c, cw, dc = contour(A, ns=24)
    rng = np.random.default_rng(5)
    b = rng.standard_normal(n * n)
    y = rng.standard_normal(n * n)
    print(f"case param gradient n={n} m={m}", flush=True)
    eng = FabEngine(A, m=m, xi=c["xi"], wts=cw, dwts=dc)
    u = eng.forward(b)
    g = 2 * (u - y)
    gb, (gr,) = eng.backward_param(g, [E], want_b=True)
    eng.delete()
    ref = oadj.param_gradient(A, E, b, g, c["xi"], cw)
    rb = np.imag(oadj.resolvents(A, c["xi"], g) @ cw)
    print("dL/dr", gr, ref, "grad_b rel", rel(gb, rb))
    assert abs(gr - ref) <= 1e-7 * abs(ref)
    assert rel(gb, rb) < 1e-7


def test_deflated_param_gradient_vs_exact_reference():
    """gm_fab_* with K = 8 deflation pairs: the forward against the oracle's
    deflated split with exact resolvents, dL/db (the same split applied to g),
    and dL/dr through the block split of the Frechet derivative
    (oracle/adjoint.deflated_param_gradient, pinned by central differences in
    tests/test_adjoint_oracle.py).  m = 120 resolves the deflated Krylov
    problem to ~1e-15 on the CPU; tolerances 1e-10 (vectors) and 1e-8 (dL/dr,
    whose cross blocks go through real-shift tridiagonal solves)."""
    n, K, m, r = 40, 8, 120, 4.0
    A = problems.anisotropic_9pt(n, ratio=r)
    E = problems.anisotropic_9pt_ratio_derivative(n)
    w, U = np.linalg.eigh(A.toarray())
    Q, lam = U[:, :K], w[:K]
    c = q.contour_nodes(0.95 * w[K], 1.05 * w.max(), 24, TAU, 1.0)
    cw, dc = contour_weights(c["xi"], c["dxidt"], c["s"], TAU, ALPHA)
    f = (lambda x: q.f1(x, TAU, ALPHA))
    F = q.divided_differences(f, lambda x: q.df1(x, TAU, ALPHA), lam)
    rng = np.random.default_rng(5)
    b = rng.standard_normal(n * n) + 5.0 * Q @ rng.standard_normal(K)
    y = rng.standard_normal(n * n)
    print(f"case deflated param gradient n={n} K={K} m={m}", flush=True)
    eng = FabEngine(A, m=m, xi=c["xi"], wts=cw, Q=Q, lam=lam, f_lambda=f(lam), f_divdiff=F)
    u = eng.forward(b)
    ref_u = oadj.deflated_forward(A, b, c["xi"], cw, Q, f(lam))
    g = 2 * (u - y)
    gb, (gr,) = eng.backward_param(g, [E], want_b=True)
    # the same with the derivative kept on the device: bitwise equal
    eng.set_param_derivs([E])
    eng.forward(b, want_u=False)
    gb2, (gr2,) = eng.backward_param(g, None, want_b=True)
    assert gr2 == gr and np.array_equal(gb2, gb)
    eng.delete()
    ref_gb = oadj.deflated_forward(A, g, c["xi"], cw, Q, f(lam))
    ref = oadj.deflated_param_gradient(A, E, b, g, c["xi"], cw, Q, lam, f(lam), F)
    print("forward", rel(u, ref_u), "grad_b", rel(gb, ref_gb), "dL/dr", gr, ref)
    assert rel(u, ref_u) < 1e-10
    assert rel(gb, ref_gb) < 1e-10
    assert abs(gr - ref) <= 1e-8 * abs(ref)


@pytest.mark.parametrize("m,K,R", [(30, 0, 4), (70, 0, 4), (70, 4, 4), (40, 6, 32), (40, 0, 64),
                                   (16, 0, 16), (30, 4, 16), (70, 0, 16)])
def test_batch_param_gradient_matches_single_column(m, K, R):
    """gm_bfab_backward_param (the P-P block contracted in the Krylov bases)
    against gm_fab_backward_param column by column (resolvent blocks, pinned
    against exact resolvents above), with and without deflation.  The P-P
    block takes two forms: S = V_g^T E V_b on MFMA contracted with M per
    column (16 columns up, m <= 48: 1, 2 or 3 tiles of 16 a side), else
    T = V_g M on MFMA (l in tiles of 16 up to 64 a pass, so m = 70 takes two
    passes) and a one-pass dot.  Both engines run the same Krylov
    approximation; they differ in reduction partitions, so they agree to
    rounding amplified by m Lanczos steps (1e-9)."""
    from gpumatfun_amd.adjoint import BatchFabEngine
    n, r = 24, 4.0
    A = problems.anisotropic_9pt(n, ratio=r)
    E = problems.anisotropic_9pt_ratio_derivative(n)
    w, U = np.linalg.eigh(A.toarray())
    c = q.contour_nodes(0.95 * w[K], 1.05 * w.max(), 24, TAU, 1.0)
    cw, _ = contour_weights(c["xi"], c["dxidt"], c["s"], TAU, ALPHA)
    kw1, kwb = {}, {}
    if K:
        f = (lambda x: q.f1(x, TAU, ALPHA))
        F = q.divided_differences(f, lambda x: q.df1(x, TAU, ALPHA), w[:K])
        kw1 = dict(Q=U[:, :K], lam=w[:K], f_lambda=f(w[:K]), f_divdiff=F)
        kwb = dict(Q=U[:, :K], lam=w[:K], f1_lambda=f(w[:K]), f2_lambda=f(w[:K]),
                   f1_divdiff=F, f2_divdiff=F)
    rng = np.random.default_rng(11)
    B = rng.standard_normal((n * n, R))
    G = rng.standard_normal((n * n, R))
    print(f"case batch param gradient m={m} K={K} R={R}", flush=True)
    be = BatchFabEngine(A, R=R, m=m, xi=c["xi"], wts1=cw, wts2=cw, **kwb)
    be.set_param_derivs([E])
    UB = be.forward(B)
    GB, gt = be.backward_param(G)
    be.delete()
    for col in range(R):
        eng = FabEngine(A, m=m, xi=c["xi"], wts=cw, **kw1)
        u = eng.forward(B[:, col])
        gb, (gr,) = eng.backward_param(G[:, col], [E], want_b=True)
        eng.delete()
        print(f"column {col}: u {rel(UB[:, col], u):.2e} grad_b {rel(GB[:, col], gb):.2e} "
              f"dL/dr {gt[0, col]:.15e} {gr:.15e}")
        assert rel(UB[:, col], u) < 1e-10 and rel(GB[:, col], gb) < 1e-10
        assert abs(gt[0, col] - gr) <= 1e-9 * abs(gr)


@pytest.mark.parametrize("R", [4, 16])
def test_batch_param_gradient_long_rows(R):
    """A 24^2 Laplacian with a hub row (node 0 coupled to 30 others, 31
    nonzeros): the P-P block takes its long-row path there (the one-pass dot
    at 4 columns, rows over 12 nonzeros; the S = V_g^T E V_b form at 16, rows
    over 9).  dL/dtheta for a random symmetric E on A's pattern, the
    multi-column engine against the single-column one, column by column."""
    import scipy.sparse as sp
    from gpumatfun_amd.adjoint import BatchFabEngine
    n, m = 24, 40
    A = problems.laplacian_2d_dirichlet(n).tolil()
    for j in range(1, 31):
        A[0, j] = A[j, 0] = -50.0
    A[0, 0] += 50.0 * 30 + 1.0
    for j in range(1, 31):
        A[j, j] += 50.0 + 1.0
    A = sp.csr_matrix(A)
    A.sort_indices()
    assert np.diff(A.indptr).max() > 12
    rng = np.random.default_rng(21)
    Eu = sp.triu(sp.csr_matrix((rng.standard_normal(A.nnz), A.indices, A.indptr), shape=A.shape))
    E = sp.csr_matrix(Eu + sp.triu(Eu, 1).T)
    E.sort_indices()
    assert np.array_equal(E.indptr, A.indptr) and np.array_equal(E.indices, A.indices)
    w = np.linalg.eigvalsh(A.toarray())
    c = q.contour_nodes(0.95 * w[0], 1.05 * w.max(), 24, TAU, 1.0)
    cw, _ = contour_weights(c["xi"], c["dxidt"], c["s"], TAU, ALPHA)
    B = rng.standard_normal((n * n, R))
    G = rng.standard_normal((n * n, R))
    print(f"case batch param gradient long rows R={R}", flush=True)
    be = BatchFabEngine(A, R=R, m=m, xi=c["xi"], wts1=cw, wts2=cw)
    be.set_param_derivs([E])
    be.forward(B)
    _, gt = be.backward_param(G, want_b=False)
    be.delete()
    for col in range(R):
        eng = FabEngine(A, m=m, xi=c["xi"], wts=cw)
        eng.forward(B[:, col])
        _, (gr,) = eng.backward_param(G[:, col], [E])
        eng.delete()
        print(f"column {col}: dL/dtheta {gt[0, col]:.15e} {gr:.15e}")
        assert abs(gt[0, col] - gr) <= 1e-9 * abs(gr)
