"""The uneven batch (gm_ubatch_*): independent problems of different sizes in
one launch per stage, each problem against the CPU oracle (relative 2-norm
<= 1e-10) and against the plain engine run on it alone (<= 1e-12: the same
arithmetic per row, reductions over differently cut row ranges)."""

import numpy as np
import pytest

import oracle
from cases import oracle_setup
from conftest import rel
from gpumatfun_amd import problems, quadrature as q
from gpumatfun_amd.engine import LanczosEngine
from gpumatfun_amd.ubatch import UnevenBatch

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _dirichlet(side, seed, m, n, degree, prob_dim, mass=False):
    A = problems.laplacian_2d_dirichlet(side)
    lam = problems.dirichlet_eigs_2d(side)
    rng = np.random.default_rng(seed)
    u0 = rng.standard_normal(A.shape[0])
    ms = 0.5 + rng.random(A.shape[0]) if mass else None
    return q.make_inputs(A, u0, u0 * (1 - u0), m=m, n=n, eta=1e-3, alpha=1.5, dt=1e-3,
                         l1=float(lam.min()), lN=float(lam.max()), degree=degree,
                         prob_dim=prob_dim, mass=ms)


def _cpu(inp, steps):
    S = oracle_setup(inp)
    if np.all(inp.mass == 1.0):
        return oracle.cpu_steps(S, inp.U[:, 0], inp.U[:, 1], steps)
    return oracle.cpu_steps_nonuniform(S, inp.U[:, 0], inp.U[:, 1], steps)


@pytest.mark.parametrize("degree,wave", [(0, False), (2, False), (2, True)])
def test_uneven_batch_vs_oracle_and_plain(degree, wave):
    """Six 2D Dirichlet Laplacians (7^2 .. 64^2 rows, one polynomial shared at
    the largest problem's scale), one of them with a non-uniform mass in
    [0.5, 1.5) (the update of problem_startup_CPU_nonuniform.m).  A FEM disc
    without deflation is no parity case: at m = 30 its Lanczos basis has
    lost orthogonality and the oracle disagrees with itself under a row
    permutation by 1e-2."""
    m, n, steps = 30, 16, 3
    prob_dim = float(problems.dirichlet_eigs_2d(64).max()) / 4.0
    inps = [_dirichlet(s, i, m, n, degree, prob_dim) for i, s in enumerate([16, 40, 7, 64, 23])]
    inps.insert(2, _dirichlet(50, 99, m, n, degree, prob_dim, mass=True))
    print(f"case uneven batch degree={degree} wave_solves={wave} "
          f"sizes={[i.A.shape[0] for i in inps]}", flush=True)
    ub = UnevenBatch.from_inputs(inps, clamp=False, wave_solves=wave)
    info = ub.info()
    assert info["solve_form"] == ("wave_pcr" if wave else "lane_global")
    assert info["num_problems"] == len(inps)
    assert info["total_rows"] == sum(i.A.shape[0] for i in inps)
    ub.step(steps)
    ub.synchronize()
    for i, inp in enumerate(inps):
        U = ub.get_U(i)
        u, gu = _cpu(inp, steps)
        e0, e1 = rel(U[:, 0], u), rel(U[:, 1], inp.dt * gu)
        eng = LanczosEngine.from_inputs(inp, clamp=False)
        eng.step(steps)
        Up = eng.get_U()
        eng.delete()
        ep = rel(U, Up)
        print(f"problem {i} n={inp.A.shape[0]}: vs oracle {e0:.2e} {e1:.2e}, vs plain {ep:.2e}")
        assert e0 < TOL and e1 < TOL
        assert ep < 1e-12
    ub.delete()


def test_uneven_batch_graph_matches_eager_bitwise():
    m, n = 24, 12
    inps = [_dirichlet(s, 10 + i, m, n, 0, 1.0) for i, s in enumerate([9, 30, 17])]
    a = UnevenBatch.from_inputs(inps, clamp=False)
    b = UnevenBatch.from_inputs(inps, clamp=False)
    a.step(2)
    b.step(2, use_graph=True)
    for _ in range(3):
        a.synchronize()
        b.synchronize()
        for i in range(len(inps)):
            assert np.array_equal(a.get_U(i), b.get_U(i))
        a.step(1)
        b.step(1, use_graph=True)
    a.delete()
    b.delete()


def test_uneven_batch_log_uniform_sides_vs_oracle():
    """The uneven workload's shape at a test size: 16 2D Dirichlet Laplacians with
    sides drawn log-uniformly in [32, 448] (seed 3: sides 40 to 399, 1.6k to 159k
    rows, several problems spanning hundreds of row blocks), m = 40, 32 shifts, no
    preconditioner, SpMM schedule chosen by the create-time autotuner (whose
    timing steps must leave the state as it found it).  Every problem against the
    oracle's C restatement to 1e-10."""
    from oracle import cfab
    m, n, steps = 40, 32, 2
    sides = problems.uneven_batch_sides(16, 32, 448, seed=3)
    inps = [_dirichlet(int(s), 100 + i, m, n, 0, 1.0) for i, s in enumerate(sides)]
    print(f"case uneven batch 16 problems sides={sorted(int(x) for x in sides)}", flush=True)
    ub = UnevenBatch.from_inputs(inps, clamp=False, autotune=True)
    info = ub.info()
    print(f"autotuned schedule: {info['spmm_tile']}-nonzero row blocks, "
          f"{info['index_bytes']}-byte indices, early loads {info['early_loads']}", flush=True)
    assert info["row_blocks"] > len(inps) and info["spmm_tile"] in (512, 1024, 2048)
    ub.step(steps)
    ub.synchronize()
    for i, inp in enumerate(inps):
        U = ub.get_U(i)
        u, gu, _ = cfab.cpu_steps_c(oracle_setup(inp), inp.U[:, 0], inp.U[:, 1], steps)
        e0, e1 = rel(U[:, 0], u), rel(U[:, 1], inp.dt * gu)
        print(f"problem {i} n={inp.A.shape[0]}: {e0:.2e} {e1:.2e}", flush=True)
        assert e0 < TOL and e1 < TOL
    ub.delete()
