"""ctypes wrapper of the C restatement ``oracle/c/fab_oracle.c`` (test infrastructure).

``build()`` compiles it with gcc/OpenMP twice: in double precision into
``oracle/_build/libfab_oracle.so`` (the restatement) and in long double into
``oracle/_build/libfab_oracle_ld.so`` (the same evaluation with rounding errors
2^11 times smaller, the reference point for comparing FP64 evaluation orders).
``cpu_steps_c`` runs the step loop of ``problem_startup_CPU_nonuniform.m``
(mass = 1 gives ``problem_startup_CPU.m``) on ``threads`` host threads.
"""

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "c", "fab_oracle.c")
OUT = os.path.join(HERE, "_build", "libfab_oracle.so")
OUT_LD = os.path.join(HERE, "_build", "libfab_oracle_ld.so")

_libs = {}


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    for out, extra in ((OUT, []), (OUT_LD, ["-DFAB_REAL=long double"])):
        tmp = out + ".tmp"
        cmd = (["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=c11"] + extra
               + ["-o", tmp, SRC, "-lm"])
        subprocess.run(cmd, check=True)
        os.replace(tmp, out)
    return OUT


def load(extended=False):
    """The double-precision library, or with ``extended`` the long-double one."""
    key = bool(extended)
    if key not in _libs:
        path = OUT_LD if extended else OUT
        if not os.path.exists(path):
            build()
        lib = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        lib.fab_cpu_steps.restype = C.c_int
        lib.fab_cpu_steps.argtypes = [C.c_long, ip, ip, dp, C.c_int, dp, dp, dp, C.c_double,
                                      C.c_int, C.c_int, dp, C.c_int, dp, dp, dp, dp, C.c_double,
                                      dp, dp, C.c_int, dp, dp]
        lib.fab_set_threads.restype = C.c_int
        lib.fab_set_threads.argtypes = [C.c_int]
        lib.fab_set_variant.restype = C.c_uint
        lib.fab_set_variant.argtypes = [C.c_uint]
        _libs[key] = lib
    return _libs[key]


# Arithmetic variants of the C restatement (fab_oracle.c): each switches one
# rounding choice to the one a fused GPU evaluation makes, or adds full
# reorthogonalisation (the rounding-free Lanczos answer).
V_SCALE, V_FMA, V_TREE, V_DEFL, V_REORTH = 1, 2, 4, 8, 16
# long-double build only: round one stage's results to double
V_RDEFL, V_RLANCZOS, V_RSPMV = 32, 64, 128


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _cplx(z):
    z = np.asarray(z, dtype=np.complex128).ravel()
    out = np.empty(2 * z.size)
    out[0::2] = z.real
    out[1::2] = z.imag
    return out


def cpu_steps_c(setup, u0, gu0, steps, threads=0, variant=0, extended=False):
    """Same contract as ``oracle.cpu_steps_nonuniform`` (mass from setup, default ones).

    ``variant`` ORs the V_* switches (0: the literal restatement); ``extended``
    runs the long-double build (inputs and outputs stay float64)."""
    lib = load(extended)
    used = lib.fab_set_threads(int(threads))
    lib.fab_set_variant(int(variant))
    A = setup["A"].tocsr()
    A.sort_indices()
    n = A.shape[0]
    rp = np.ascontiguousarray(A.indptr, dtype=np.int32)
    ci = np.ascontiguousarray(A.indices, dtype=np.int32)
    va = np.ascontiguousarray(A.data, dtype=np.float64)
    Q = setup.get("Q")
    K = 0 if Q is None else Q.shape[1]
    Qf = np.ascontiguousarray(np.asarray(Q, dtype=np.float64).T.ravel()) if K else np.zeros(1)
    el = np.ascontiguousarray(setup["exp_lambda"], dtype=np.float64) if K else np.zeros(1)
    pl = np.ascontiguousarray(setup["phi_lambda"], dtype=np.float64) if K else np.zeros(1)
    gamma = np.ascontiguousarray(setup["gamma"], dtype=np.float64)
    nd = gamma.size
    gsr = np.ascontiguousarray(np.asarray(setup["gamma_s_real"]).reshape(nd, -1).T.ravel())
    gsi = np.ascontiguousarray(np.asarray(setup["gamma_s_imag"]).reshape(nd, -1).T.ravel())
    ns = int(setup["n"])
    w1 = _cplx(setup["wts1_cpu"][:ns])
    w2 = _cplx(setup["wts2_cpu"][:ns])
    ex = _cplx(setup["etaxi"][:ns])
    mass = setup.get("mass")
    mass = np.ones(n) if mass is None else np.ascontiguousarray(mass, dtype=np.float64)
    u = np.array(u0, dtype=np.float64)
    gu = np.array(gu0, dtype=np.float64)
    bad = lib.fab_cpu_steps(n, rp.ctypes.data_as(C.POINTER(C.c_int)),
                            ci.ctypes.data_as(C.POINTER(C.c_int)), _p(va), K, _p(Qf), _p(el),
                            _p(pl), float(setup["dt"]), int(setup["m"]), nd, _p(gamma), ns,
                            _p(gsr), _p(gsi), _p(w1), _p(w2), float(setup["s"]), _p(ex),
                            _p(mass), int(steps), _p(u), _p(gu))
    lib.fab_set_variant(0)
    if bad:
        raise ZeroDivisionError("singular shifted tridiagonal system in the C oracle")
    return u, gu, used
