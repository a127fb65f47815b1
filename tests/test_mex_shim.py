"""The MEX gateways (gpumatfun_amd/mex/gm_mex.cpp) driven through the fake-MEX harness.

CPU tests: argument checking with the reference's message identifiers
(lanczos_create.cpp:40-733, get_U.cpp:12-26).  GPU tests: the step sequence of
nicksmethodgpu.m:8-21 through the gateways, against the engine and the oracle.
"""

import os
import sys

import numpy as np
import pytest

from conftest import ROOT, rel

sys.path.insert(0, os.path.join(ROOT, "tests", "fake_mex"))
import harness  # noqa: E402


@pytest.fixture(scope="module")
def mex():
    harness.build()
    return harness.Mex()


def create_args(inp):
    """The 26 arguments of lanczosEngine.m:9, valvec = -1:1."""
    return [inp.U, inp.A.tocsc(), inp.Q, inp.lam, float(inp.m), inp.gamma, inp.gamma_s_real,
            inp.gamma_s_imag, inp.exp_lambda, inp.phi_lambda, inp.dt, inp.ode, inp.aode,
            inp.bode, inp.code, inp.vrest, inp.vamp, inp.vth, inp.vpeak, inp.c1ion, inp.c2ion,
            inp.mass, inp.wts1, inp.wts2, inp.shifts, np.array([-1.0, 0.0, 1.0])]


def small_inputs():
    from cases import dirichlet_case
    return dirichlet_case(side=12, m=10, n=8)


@pytest.mark.parametrize("idx,bad,ident", [
    (24, "cut", "MATLAB:lanczos_create:bad_shifts"),
    (22, "cut", "MATLAB:lanczos_create:bad_wts1"),
    (1, "dense", "MATLAB:lanczos_create:inputNotSparse"),
    (4, "vector", "MATLAB:lanczos_create:bad_N"),
    (11, "cut", "MATLAB:lanczos_create:bad_ode"),
    (21, "cut", "MATLAB:lanczos_create:bad_mass"),
    (6, "transpose", "MATLAB:lanczos_create:bad_gamma_s_real"),
    (3, "extra", "MATLAB:lanczos_create:bad_lambda"),
    (0, "complex", "MATLAB:lanczos_create:inputNotReal"),
])
def test_create_argument_errors(mex, idx, bad, ident):
    args = create_args(small_inputs())
    a = args[idx]
    if bad == "cut":
        args[idx] = np.asarray(a)[:-2] if np.ndim(a) == 1 else np.asarray(a)[:-1]
    elif bad == "dense":
        args[idx] = a.toarray()
    elif bad == "vector":
        args[idx] = np.array([10.0, 11.0])
    elif bad == "transpose":
        args[idx] = np.asarray(a).T
    elif bad == "extra":
        args[idx] = np.array([1.0])
    elif bad == "complex":
        args[idx] = np.asarray(a) + 0j
    with pytest.raises(harness.MexError) as e:
        mex("lanczos_create", *args)
    assert e.value.ident == ident


def test_create_arity(mex):
    args = create_args(small_inputs())
    with pytest.raises(harness.MexError) as e:
        mex("lanczos_create", *args[:25])
    assert e.value.ident == "MATLAB:lanczos_create:invalidNumInputs"
    with pytest.raises(harness.MexError) as e:
        mex("lanczos_create", *args, nlhs=0)
    assert e.value.ident == "MATLAB:lanczos_create:maxlhs"


def test_handle_gateways_check_their_argument(mex):
    with pytest.raises(harness.MexError) as e:
        mex("get_U_mex", np.array([1.0]))
    assert e.value.ident == "MATLAB:get_U:inputNotHandle"
    with pytest.raises(harness.MexError) as e:
        mex("build_subspace_mex")
    assert e.value.ident == "MATLAB:build_subspace:invalidNumInputs"
    with pytest.raises(harness.MexError) as e:   # unknown handle -> library status
        mex("build_subspace_mex", np.uint64(12345))
    assert e.value.ident == "gpumatfun:GM_EBADHANDLE"


@pytest.mark.gpu
def test_mex_step_sequence_matches_engine_and_oracle(mex):
    """nicksmethodgpu.m:8-21 through the gateways (reference semantics: clamp on)."""
    import oracle
    from cases import dirichlet_case
    from gpumatfun_amd.engine import LanczosEngine
    inp = dirichlet_case(side=32, m=24, n=16)
    h = mex("lanczos_create", *create_args(inp))
    eng = LanczosEngine.from_inputs(inp, clamp=True)
    ref = oracle.GpuSemantics(inp.U, inp.A, inp.Q, inp.m, inp.gamma, inp.gamma_s_real,
                              inp.gamma_s_imag, inp.exp_lambda, inp.phi_lambda, inp.dt, inp.mass,
                              inp.wts1, inp.wts2, inp.shifts, clamp=True)
    for _ in range(2):
        mex("build_subspace_mex", h)
        mex("lin_sys_solves_mex", h)
        mex("apply_V_mex", h)
        X = mex("get_xsol_mex", h)
        mex("update_U_mex", h)
        eng.step(1)
        ref.build_subspace()
        ref.lin_sys_solves()
        ref.apply_V()
        assert rel(X, ref.get_xsol()) < 1e-10
        ref.update_U()
    U = mex("get_U_mex", h)
    assert np.array_equal(U, eng.get_U())
    assert rel(U, ref.get_U()) < 1e-10
    ab = mex("get_aabb_mex", h)
    assert ab.shape == (inp.m + 1, 4)
    mex("lanczos_delete", h)
    with pytest.raises(harness.MexError) as e:
        mex("get_U_mex", h)
    assert e.value.ident == "gpumatfun:GM_EBADHANDLE"
    eng.delete()


@pytest.mark.gpu
def test_mex_get_aabb_and_set_y(mex):
    """The two gateways lanczosEngine.m:25-31 calls: get_aabb_mex returns the
    Lanczos coefficients [alpha1 alpha2 beta1 beta2] ((m+1) x 4) of the last
    build -- against the GPU-semantics model and bit for bit against the library
    call -- and set_y_mex replaces the shifted solves' output, so apply_V_mex
    recombines a caller-supplied Y (here the model's solutions scaled by 2 and
    with the phi systems negated: apply_V is linear in Y)."""
    import oracle
    from cases import dirichlet_case
    from gpumatfun_amd.engine import LanczosEngine
    inp = dirichlet_case(side=32, m=24, n=16)
    h = mex("lanczos_create", *create_args(inp))
    eng = LanczosEngine.from_inputs(inp, clamp=True)
    ref = oracle.GpuSemantics(inp.U, inp.A, inp.Q, inp.m, inp.gamma, inp.gamma_s_real,
                              inp.gamma_s_imag, inp.exp_lambda, inp.phi_lambda, inp.dt, inp.mass,
                              inp.wts1, inp.wts2, inp.shifts, clamp=True)
    mex("build_subspace_mex", h)
    eng.build_subspace()
    ref.build_subspace()
    ab = mex("get_aabb_mex", h)
    assert ab.shape == (inp.m + 1, 4)
    assert np.array_equal(ab, eng.get_aabb())
    assert rel(ab, ref.aabb) < 1e-10
    ref.lin_sys_solves()
    n = inp.extra["n"]
    sign = np.concatenate([np.ones(n), -np.ones(n)])
    yr = np.asfortranarray(2.0 * ref.Yr * sign)
    yi = np.asfortranarray(2.0 * ref.Yi * sign)
    with pytest.raises(harness.MexError) as e:
        mex("set_y_mex", h, yr[:, :-1], yi)
    assert e.value.ident == "MATLAB:set_y:badY"
    mex("set_y_mex", h, yr, yi)
    mex("apply_V_mex", h)
    X = mex("get_xsol_mex", h)
    ref.Yr, ref.Yi = yr.copy(), yi.copy()
    ref.apply_V()
    Xr = ref.get_xsol()
    assert rel(X[:, 0], Xr[:, 0]) < 1e-10 and rel(X[:, 1], Xr[:, 1]) < 1e-10
    mex("lanczos_delete", h)
    eng.delete()


def test_eigs_gateway_argument_errors(mex):
    from gpumatfun_amd import problems
    A = problems.laplacian_2d_dirichlet(8)
    cases = [((A.toarray(), 4.0), "MATLAB:eigs_smallest:inputNotSparse"),
             ((A, np.array([1.0, 2.0])), "MATLAB:eigs_smallest:bad_k"),
             ((A, 2.5), "MATLAB:eigs_smallest:bad_k"),
             ((A, 4.0, np.array([16.0, 40.0])), "MATLAB:eigs_smallest:bad_opts"),
             ((A,), "MATLAB:eigs_smallest:invalidNumInputs"),
             ((A, 4.0, np.array([24.0, 40.0, 10.0, 1e-8])), "gpumatfun:GM_EINVAL")]
    for args, ident in cases:
        with pytest.raises(harness.MexError) as e:
            mex("eigs_smallest_mex", *args, nlhs=2)
        assert e.value.ident == ident, (args, e.value)


@pytest.mark.gpu
def test_eigs_gateway_matches_library(mex):
    """[Q, lambda] = eigs_smallest_mex(A, k): the 2D Dirichlet closed form, and
    the same pairs as the Python wrapper (same seed, same arithmetic)."""
    from gpumatfun_amd import problems
    from gpumatfun_amd.eigs import eigs_smallest
    A = problems.laplacian_2d_dirichlet(40)
    Q, lam = mex("eigs_smallest_mex", A.tocsc(), 6.0, nlhs=2)
    exact = np.sort(problems.dirichlet_eigs_2d(40).ravel())[:6]
    assert Q.shape == (1600, 6) and lam.shape == (6, 1)
    assert np.abs(lam[:, 0] - exact).max() <= 1e-10 * exact[0]
    lam2, Q2 = eigs_smallest(A, 6, block_size=16, degree=40, max_passes=200, tol=1e-10)
    assert np.array_equal(lam[:, 0], lam2) and np.array_equal(Q, Q2)


def test_shift_invert_gateways_match_library_and_arpack(mex):
    """[Q, lambda] = eigs_shift_invert_mex(A, k) and lmax = lambda_max_mex(A) --
    host code, no GPU: the same pairs as the Python wrapper (same default shift,
    same arithmetic), eigenvalues as ARPACK's, and lambda_max as ARPACK's."""
    from gpumatfun_amd import problems
    from gpumatfun_amd.eigs import eigs_shift_invert, lambda_max
    A, mass, nodes = problems.fem_disc(2000, 200, 4)
    Q, lam = mex("eigs_shift_invert_mex", A.tocsc(), 7.0, nlhs=2)
    assert Q.shape == (2000, 7) and lam.shape == (7, 1)
    lam2, Q2 = eigs_shift_invert(A, 7)
    assert np.array_equal(lam[:, 0], lam2) and np.array_equal(Q, Q2)
    lam_a, _ = problems.smallest_eigenpairs(A, 7)
    lmax = float(np.asarray(mex("lambda_max_mex", A.tocsc(), nlhs=1)).ravel()[0])
    assert lmax == lambda_max(A)
    assert lmax == pytest.approx(problems.largest_eigenvalue(A), rel=1e-11)
    assert np.abs(lam[:, 0] - lam_a).max() <= 1e-12 * lmax
    # an explicit shift and options reach the library
    Q3, lam3 = mex("eigs_shift_invert_mex", A.tocsc(), 7.0, -0.5, np.array([16.0, 10.0, 1e-13]),
                   nlhs=2)
    lam4, _ = eigs_shift_invert(A, 7, sigma=-0.5, block_size=16, max_blocks=10, tol=1e-13)
    assert np.array_equal(lam3[:, 0], lam4)


def test_shift_invert_gateway_argument_errors(mex):
    from gpumatfun_amd import problems
    A = problems.laplacian_2d_dirichlet(8)
    lam1 = float(problems.dirichlet_eigs_2d(8).min())
    cases = [((A.toarray(), 4.0), "MATLAB:eigs_shift_invert:inputNotSparse"),
             ((A, np.array([1.0, 2.0])), "MATLAB:eigs_shift_invert:bad_k"),
             ((A, 2.5), "MATLAB:eigs_shift_invert:bad_k"),
             ((A, 4.0, np.array([1.0, 2.0])), "MATLAB:eigs_shift_invert:bad_sigma"),
             ((A, 4.0, -1.0, np.array([16.0, 40.0])), "MATLAB:eigs_shift_invert:bad_opts"),
             ((A,), "MATLAB:eigs_shift_invert:invalidNumInputs"),
             ((A, 4.0, 2.0 * lam1), "gpumatfun:GM_EBREAKDOWN")]
    for args, ident in cases:
        with pytest.raises(harness.MexError) as e:
            mex("eigs_shift_invert_mex", *args, nlhs=2)
        assert e.value.ident == ident, (args, e.value)
    with pytest.raises(harness.MexError) as e:
        mex("lambda_max_mex", A.toarray(), nlhs=1)
    assert e.value.ident == "MATLAB:lambda_max:inputNotSparse"
