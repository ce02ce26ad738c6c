"""ctypes binding of the compiled CPU restatement ``oracle/cpu/pmx_cpu.cpp``
(C++ / OpenMP, FP64): the reference's setup and PCG on the host cores.

TEST INFRASTRUCTURE ONLY: the CPU baseline of ``bench.py`` and a compiled
second checker, pinned against the numpy oracle by
``tests/test_cpu_baseline.py``.  The product path never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
SRC = HERE / "cpu" / "pmx_cpu.cpp"
LIB = HERE / "_build" / "libpmx_cpu.so"
PHASES = ("operator+lv1", "lv2_eigen", "lv2_galerkin", "lv2_factor", "lv3", "pcg", "total")


class _Problem(C.Structure):
    _fields_ = [("M", C.c_int), ("N", C.c_int), ("P", C.c_int), ("dom", C.c_double * 3),
                ("kx", C.c_void_p), ("ky", C.c_void_p), ("kz", C.c_void_p),
                ("sd", C.c_int * 3), ("en_lv1", C.c_int), ("en_lv2", C.c_int),
                ("pin", C.c_int * 3), ("rhs", C.c_void_p), ("rtol", C.c_double),
                ("max_it", C.c_int), ("threads", C.c_int), ("eig_tol", C.c_double)]


class _Result(C.Structure):
    _fields_ = [("x", C.c_void_p), ("its", C.c_int), ("reason", C.c_int),
                ("rnorm", C.c_double), ("rnorm0", C.c_double), ("t", C.c_double * 7),
                ("nc", C.c_int), ("lanczos_steps_max", C.c_int),
                ("lanczos_unconverged", C.c_int), ("threads", C.c_int),
                ("eig_min_lv1", C.c_double), ("eig_max_lv1", C.c_double),
                ("eig_lv2", C.c_double * 16), ("eig_lv1", C.c_void_p)]


def build(verbose=False):
    """g++ -O3 -fopenmp (x86-64-v3) into oracle/_build/libpmx_cpu.so."""
    LIB.parent.mkdir(exist_ok=True)
    cmd = ["g++", "-O3", "-fopenmp", "-march=x86-64-v3", "-std=c++17", "-shared", "-fPIC",
           str(SRC), "-o", str(LIB)]
    subprocess.run(cmd, check=True)
    if verbose:
        print(f"built {LIB}")
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise ImportError(f"{LIB} is missing: run oracle.cpu_baseline.build()")
        _lib = C.CDLL(str(LIB))
        _lib.pmxc_setup_solve.argtypes = [C.POINTER(_Problem), C.POINTER(_Result)]
        _lib.pmxc_setup_solve.restype = C.c_int
    return _lib


def setup_solve(prob, threads=0, eig_tol=1e-8, max_it=10000, want_eig=False):
    """Setup + PCG of ``prob`` (a premixfem_amd.problems.Problem, pinned or
    not); returns dict(x, its, reason, times per phase, stats)."""
    M, N, P = prob.mesh
    n = M * N * P
    k = [np.ascontiguousarray(prob.kappa[d].ravel(), dtype=np.float64) for d in range(3)]
    b = np.ascontiguousarray(prob.rhs.ravel(), dtype=np.float64)
    x = np.zeros(n)
    sd = prob.sd if isinstance(prob.sd, tuple) else (prob.sd,) * 3
    pin = prob.pin_cell if prob.pin_cell is not None else (-1, -1, -1)
    pr = _Problem(M, N, P, (C.c_double * 3)(*prob.dom), k[0].ctypes.data, k[1].ctypes.data,
                  k[2].ctypes.data, (C.c_int * 3)(*sd), prob.en_lv1, prob.en_lv2,
                  (C.c_int * 3)(*pin), b.ctypes.data, prob.rtol, max_it, threads, eig_tol)
    nsub = int(np.prod(sd))
    ev = np.zeros(nsub * prob.en_lv1) if want_eig else None
    rs = _Result()
    rs.x = x.ctypes.data
    rs.eig_lv1 = ev.ctypes.data if want_eig else None
    rc = lib().pmxc_setup_solve(C.byref(pr), C.byref(rs))
    if rc != 0:
        raise RuntimeError(f"pmxc_setup_solve failed ({rc})")
    out = dict(x=x, its=rs.its, reason=rs.reason, rnorm=rs.rnorm, rnorm0=rs.rnorm0,
               t=dict(zip(PHASES, list(rs.t))), nc=rs.nc, threads=rs.threads,
               lanczos_steps_max=rs.lanczos_steps_max,
               lanczos_unconverged=rs.lanczos_unconverged,
               eig_lv2=list(rs.eig_lv2)[:prob.en_lv2])
    if want_eig:
        out["eig_lv1"] = ev.reshape(nsub, prob.en_lv1)
    return out
