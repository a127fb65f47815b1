"""Batched form (many problems sharing A) against the oracle, problem by problem.

Tolerance: relative 2-norm error <= 1e-10 per column, FP64.
"""

import numpy as np
import pytest

import oracle
from cases import dirichlet_case, oracle_setup
from conftest import rel
from gpumatfun_amd.batch import BatchEngine
from gpumatfun_amd.engine import LanczosEngine

pytestmark = pytest.mark.gpu
TOL = 1e-10


def batch_U(inp, P, seed0=0):
    N = inp.A.shape[0]
    cols = []
    for p in range(P):
        u = np.random.default_rng(seed0 + p).standard_normal(N)
        cols += [u, u * (1 - u)]
    return np.column_stack(cols)


def oracle_pair(inp, u0, g0, steps):
    S = oracle_setup(inp)
    return oracle.cpu_steps_nonuniform(S, u0, g0, steps)


@pytest.mark.parametrize("P,chunk,degree", [(8, 0, 0), (8, 4, 0), (4, 0, 2), (40, 16, 0)])
def test_batch_matches_oracle_per_problem(P, chunk, degree):
    inp = dirichlet_case(side=32, m=24, n=16, degree=degree)
    U = batch_U(inp, P)
    print(f"case batch P={P} chunk={chunk} degree={degree}", flush=True)
    b = BatchEngine.from_inputs(inp, U, chunk_columns=chunk, clamp=False)
    if chunk:
        assert b.info()["chunk_columns"] == chunk
    b.step(3)
    out = b.get_U()
    b.synchronize()
    for p in (0, P // 2, P - 1):
        u, gu = oracle_pair(inp, U[:, 2 * p], U[:, 2 * p + 1], 3)
        assert rel(out[:, 2 * p], u) < TOL
        assert rel(out[:, 2 * p + 1], inp.dt * gu) < TOL
    b.delete()


_L3D = {}


@pytest.mark.parametrize("chunk", [32, 64, 128, 256])
def test_batch_laplace3d_chunk_widths(chunk):
    """The multi-RHS workload's shape at a test size: 128 problems (256 columns) on
    the 3D 7-point Laplacian 20^3, 64 shifts, m = 30, in chunks of 32, 64, 128 or
    256 columns (the column kernels' widths).  Problems 0, 63 and 127 against the
    oracle to 1e-10, and every column against the 32-column chunking to 1e-13."""
    from cases import laplace3d_case
    inp = laplace3d_case()
    P = 128
    U = batch_U(inp, P, seed0=0)
    print(f"case batch laplace3d 20^3 P={P} chunk={chunk}", flush=True)
    b = BatchEngine.from_inputs(inp, U, chunk_columns=chunk, clamp=False)
    assert b.info()["chunk_columns"] == chunk
    b.step(2)
    out = b.get_U()
    b.synchronize()
    b.delete()
    for p in (0, 63, 127):
        u, gu = oracle_pair(inp, U[:, 2 * p], U[:, 2 * p + 1], 2)
        e = (rel(out[:, 2 * p], u), rel(out[:, 2 * p + 1], inp.dt * gu))
        print(p, e, flush=True)
        assert max(e) < TOL
    if chunk == 32:
        _L3D["ref"] = out
    elif "ref" in _L3D:
        errs = [rel(out[:, c], _L3D["ref"][:, c]) for c in range(2 * P)]
        print("max column difference to chunk 32:", max(errs))
        assert max(errs) < 1e-13


def _hub_laplace3d_case(side=20, hub=600, w=1.0):
    """The 20^3 Laplacian case plus a hub: node 0 coupled to `hub` other nodes
    with weight w (a weighted graph Laplacian added, so still SPD), which puts
    a row of hub + 7 nonzeros (longer than one LDS tile) into the matrix."""
    import scipy.sparse as sp
    from cases import laplace3d_case
    from gpumatfun_amd import quadrature as q
    base = laplace3d_case(side=side)
    A = sp.lil_matrix(base.A)
    nodes = np.random.default_rng(5).choice(np.arange(1, A.shape[0]), hub, replace=False)
    for j in nodes:
        A[0, j] -= w
        A[j, 0] -= w
        A[j, j] += w
    A[0, 0] += hub * w
    A = sp.csr_matrix(A)
    h = 1.0 / (side + 1)
    lo = 3 * 4.0 * np.sin(np.pi * h / 2) ** 2 / h ** 2
    hi = float(abs(A).sum(axis=1).max())            # Gershgorin
    u0 = np.random.default_rng(0).standard_normal(A.shape[0])
    return q.make_inputs(A, u0, u0 * (1 - u0), m=30, n=64, eta=1e-3, alpha=1.5, dt=1e-3,
                         l1=lo, lN=hi, degree=0)


@pytest.mark.parametrize("chunk,degree,hub", [(256, 0, False), (256, 2, False), (256, 0, True),
                                              (32, 0, False), (64, 2, False), (128, 0, True)])
def test_batch_column_slabs_match_whole_rows(chunk, degree, hub):
    """Chunks with the SpMM in column slabs of 32 against the whole-row SpMM: 128
    problems on the 20^3 Laplacian (with the Jacobi polynomial at degree 2: the
    kFirst/kMid/kLast and Horner modes; with a 607-nonzero hub row: the row read
    from global memory), chunks of 32 to 256 columns.  Every column agrees to
    1e-13 (the per-row sums are the same; only the column sums' block order
    differs), and problems 0 and 127 match the oracle to 1e-10."""
    from cases import laplace3d_case
    from gpumatfun_amd import quadrature as q
    if hub:
        inp = _hub_laplace3d_case()
    else:
        inp = laplace3d_case()
        if degree:
            inp = q.make_inputs(inp.A, inp.U[:, 0], inp.U[:, 1], m=30, n=64, eta=1e-3,
                                alpha=1.5, dt=1e-3, l1=inp.extra["l1"], lN=inp.extra["lN"],
                                degree=degree)
    P = 128
    U = batch_U(inp, P, seed0=0)
    outs = {}
    for xs in (False, True):
        print(f"case batch chunk={chunk} column_slabs={xs} degree={degree} hub={hub}", flush=True)
        b = BatchEngine.from_inputs(inp, U, chunk_columns=chunk, clamp=False, column_slabs=xs)
        info = b.info()
        assert info["chunk_columns"] == chunk and info["column_slabs"] == xs
        b.step(2)
        outs[xs] = b.get_U()
        b.synchronize()
        b.delete()
    errs = [rel(outs[True][:, c], outs[False][:, c]) for c in range(2 * P)]
    print("max column difference slabs vs whole rows:", max(errs), flush=True)
    assert max(errs) < 1e-13
    for p in (0, P - 1):
        u, gu = oracle_pair(inp, U[:, 2 * p], U[:, 2 * p + 1], 2)
        e = (rel(outs[True][:, 2 * p], u), rel(outs[True][:, 2 * p + 1], inp.dt * gu))
        print(p, e, flush=True)
        assert max(e) < TOL


@pytest.mark.parametrize("case", ["fisher_deflated", "ms"])
def test_batch_locality_order_matches_caller_order(case):
    """The engine renumbering its nodes (GM_FLAG_LOCALITY_ORDER) against the same
    batch in the caller's order: results come back in the caller's order and
    agree to 1e-13 per column (each row's sums are unchanged; the column sums
    run over another row partition).  Fisher with K = 6 shared eigenpairs and a
    lumped mass on the FEM disc, and Mitchell-Schaeffer (gate state) on the
    20^3 Laplacian."""
    from gpumatfun_amd._lib import MS_DEFAULTS as prm
    if case == "ms":
        from cases import laplace3d_case
        inp = laplace3d_case()
        P = 16
        U = batch_U(inp, P, seed0=3)
        N = U.shape[0]
        ode = np.random.default_rng(4).uniform(0.5, 1.0, (N, P))
        kw = dict(reaction="ms", ode=ode, ms=prm)
    else:
        from cases import fem_case
        inp = fem_case(n_nodes=3000, K=6, m=30, degree=2)
        P = 8
        U = batch_U(inp, P, seed0=5)
        kw = {}
    outs, odes = {}, {}
    for loc in (False, True):
        print(f"case batch locality_order={loc} {case}", flush=True)
        b = BatchEngine.from_inputs(inp, U, clamp=False, locality_order=loc, **kw)
        assert b.info()["locality_order"] == loc
        b.step(3)
        outs[loc] = b.get_U()
        if case == "ms":
            odes[loc] = b.get_ode()
        b.synchronize()
        b.delete()
    errs = [rel(outs[True][:, c], outs[False][:, c]) for c in range(2 * P)]
    print("max column difference, locality vs caller order:", max(errs), flush=True)
    assert max(errs) < 1e-13
    if case == "ms":
        assert rel(odes[True], odes[False]) < 1e-13


def test_batch_graph_matches_eager_bitwise():
    inp = dirichlet_case(side=32, m=24, n=16)
    U = batch_U(inp, 8, seed0=10)
    a = BatchEngine.from_inputs(inp, U, clamp=False)
    g = BatchEngine.from_inputs(inp, U, clamp=False)
    a.step(2)
    g.step(2, use_graph=True)
    assert np.array_equal(a.get_U(), g.get_U())
    for _ in range(2):
        a.step(1)
        g.step(1, use_graph=True)
        a.synchronize()
        g.synchronize()
        assert np.array_equal(a.get_U(), g.get_U())
    a.delete()
    g.delete()


def test_batch_graph_matches_eager_bitwise_slabs():
    """Graph replays of the slab-SpMM dataflow (64-column chunks, unnormalised basis,
    nodes renumbered by the locality order) against eager steps, bit for bit, with
    eager work and synchronisations between the replays."""
    from cases import laplace3d_case
    inp = laplace3d_case(side=16, m=24, n=32)
    U = batch_U(inp, 32, seed0=20)
    kw = dict(clamp=False, chunk_columns=64, column_slabs=True, locality_order=True)
    a = BatchEngine.from_inputs(inp, U, **kw)
    g = BatchEngine.from_inputs(inp, U, **kw)
    assert g.info()["column_slabs"]
    print("case batch graph vs eager, slabs", flush=True)
    a.step(2)
    g.step(2, use_graph=True)
    assert np.array_equal(a.get_U(), g.get_U())
    for _ in range(3):
        a.step(1)
        g.step(1, use_graph=True)
        a.synchronize()
        g.synchronize()
        assert np.array_equal(a.get_U(), g.get_U())
    a.delete()
    g.delete()


def test_batch_agrees_with_plain_engine():
    inp = dirichlet_case(side=48, m=30, n=16)
    U = batch_U(inp, 2)
    b = BatchEngine.from_inputs(inp, U, clamp=True)
    b.step(2)
    out = b.get_U()
    for p in range(2):
        inp.U = U[:, 2 * p:2 * p + 2].copy()
        e = LanczosEngine.from_inputs(inp, clamp=True)
        e.step(2)
        ref = e.get_U()
        assert rel(out[:, 2 * p], ref[:, 0]) < TOL
        assert rel(out[:, 2 * p + 1], ref[:, 1]) < TOL
        e.delete()
    b.delete()


def test_batch_deflation_fisher_vs_oracle():
    """Shared deflation (K = 6) and lumped mass on the FEM disc (m = 30, where the
    oracle agrees with itself to 5e-12): each problem against
    problem_startup_CPU_nonuniform.m's loop, and against the plain engine."""
    from cases import fem_case
    inp = fem_case(m=30, n=16)
    N = inp.A.shape[0]
    U = np.empty((N, 8))
    for p in range(4):
        w = np.random.default_rng(40 + p).random(N)
        U[:, 2 * p] = inp.mass * w
        U[:, 2 * p + 1] = inp.mass * w * (1 - w)
    print("case batch deflation K=6 P=4", flush=True)
    b = BatchEngine.from_inputs(inp, U, clamp=False)
    b.step(3)
    out = b.get_U()
    b.synchronize()
    b.delete()
    for p in range(4):
        u, gu = oracle_pair(inp, U[:, 2 * p], U[:, 2 * p + 1], 3)
        print(p, rel(out[:, 2 * p], u), rel(out[:, 2 * p + 1], inp.dt * gu))
        assert rel(out[:, 2 * p], u) < TOL
        assert rel(out[:, 2 * p + 1], inp.dt * gu) < TOL
    inp.U = U[:, 0:2].copy()
    e = LanczosEngine.from_inputs(inp, clamp=False)
    e.step(3)
    assert rel(out[:, 0:2], e.get_U()) < 1e-12
    e.delete()


def _ms_case(n=24, K=0, m=30, ns=16, dt=0.05):
    from gpumatfun_amd import problems, quadrature as q
    A = problems.anisotropic_9pt(n, ratio=9.0)
    w, V = np.linalg.eigh(A.toarray())
    N = n * n
    p = oracle.stepping.MS_DEFAULT
    u0 = np.zeros(N)
    h0 = np.ones(N)
    l1 = float(w[K])
    inp = q.make_inputs(A, u0, u0, m=m, n=ns, eta=dt * 1e-3, alpha=1.7, dt=dt, l1=l1,
                        lN=float(w.max()), degree=0, ode=h0, Q=V[:, :K] if K else None,
                        lam=w[:K] if K else None)
    return inp, p


def _stimuli(n, P, size=5):
    X, Y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    out = []
    for k in range(P):
        cx, cy = (3 * k) % (n - size), (7 * k) % (n - size)
        out.append(((X >= cx) & (X < cx + size) & (Y >= cy) & (Y < cy + size))
                   .astype(float).reshape(-1, order="F"))
    return out


def test_batch_deflation_80_columns():
    """40 problems (80 columns) sharing K = 6 deflation pairs: QtU spans two
    groups of up to 64 columns, and the deflation apply packs four rows into a
    wavefront in the 16-column group.  Every problem against the plain engine,
    and two against the oracle."""
    from cases import fem_case
    inp = fem_case(m=30, n=16)
    N = inp.A.shape[0]
    P = 40
    U = np.empty((N, 2 * P))
    for p in range(P):
        w = np.random.default_rng(200 + p).random(N)
        U[:, 2 * p] = inp.mass * w
        U[:, 2 * p + 1] = inp.mass * w * (1 - w)
    print(f"case batch deflation K=6 P={P}", flush=True)
    b = BatchEngine.from_inputs(inp, U, clamp=False)
    b.step(3)
    out = b.get_U()
    b.synchronize()
    b.delete()
    for p in range(P):
        inp.U = U[:, 2 * p:2 * p + 2].copy()
        e = LanczosEngine.from_inputs(inp, clamp=False)
        e.step(3)
        err = rel(out[:, 2 * p:2 * p + 2], e.get_U())
        e.delete()
        assert err < 1e-12, (p, err)
    for p in (0, P - 1):
        u, gu = oracle_pair(inp, U[:, 2 * p], U[:, 2 * p + 1], 3)
        print(p, rel(out[:, 2 * p], u), rel(out[:, 2 * p + 1], inp.dt * gu))
        assert rel(out[:, 2 * p], u) < TOL
        assert rel(out[:, 2 * p + 1], inp.dt * gu) < TOL


@pytest.mark.parametrize("K", [0, 8])
def test_batch_mitchell_schaeffer_vs_oracle(K):
    """Mitchell-Schaeffer with P = 4 stimulus patches on the anisotropic 9-point
    operator, without and with shared deflation; u, dt g and the gate h of every
    problem against the oracle's step loop."""
    n, P, steps = 24, 4, 5
    inp, prm = _ms_case(n, K=K)
    N = n * n
    stims = _stimuli(n, P)
    U = np.empty((N, 2 * P))
    for k, u0 in enumerate(stims):
        U[:, 2 * k] = u0
        U[:, 2 * k + 1] = (u0 ** 2 * (1 - u0)) / prm["tau_in"] - u0 / prm["tau_out"]
    print(f"case batch MS K={K} P={P}", flush=True)
    b = BatchEngine.from_inputs(inp, U, clamp=False, reaction="ms", ode=np.ones((N, P)), ms=prm)
    b.step(steps)
    out = b.get_U()
    h = b.get_ode()
    b.synchronize()
    b.delete()
    S = oracle_setup(inp)
    for k, u0 in enumerate(stims):
        u, gu, ho = oracle.cpu_steps_ms(S, u0, np.ones(N), steps, prm)
        e = (rel(out[:, 2 * k], u), rel(out[:, 2 * k + 1], inp.dt * gu), rel(h[:, k], ho))
        print(k, e)
        assert max(e) < TOL
