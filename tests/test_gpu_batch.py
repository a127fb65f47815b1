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
