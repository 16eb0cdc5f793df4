"""ctypes wrapper of the CPU oracle (libcdx_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg -- never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libcdx_oracle.so")

ORC_OK, ORC_NOT_FACTORIZED, ORC_NONFINITE, ORC_DIVERGED, ORC_BAD_ARG = 0, 1, 2, 3, 4

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Bodies(C.Structure):
    _fields_ = [("nb", C.c_int), ("dof", _ip), ("minv", _dp), ("v", _dp), ("fext", _dp), ("fgyro", _dp)]


class Contacts(C.Structure):
    _fields_ = [("nc", C.c_int), ("body_a", _ip), ("body_b", _ip), ("mu", _dp), ("frame", _dp), ("angA", _dp),
                ("angB", _dp), ("rest", _dp), ("vbias", _dp), ("warm", _dp)]


class Params(C.Structure):
    _fields_ = [("alpha", C.c_double), ("relaxation", C.c_double), ("max_iter", C.c_int), ("tol_abs", C.c_double),
                ("tol_rel", C.c_double), ("warmstart", C.c_int), ("nesterov", C.c_int),
                ("nesterov_beta_threshold", C.c_double), ("nesterov_restart_limit", C.c_int), ("chebyshev", C.c_int),
                ("block_precond", C.c_int), ("anderson", C.c_int), ("anderson_m", C.c_int),
                ("implicit_gyro", C.c_int)]


class Joints(C.Structure):
    _fields_ = [("ns", C.c_int), ("nd", C.c_int), ("body_a", _ip), ("body_b", _ip), ("wa", _dp), ("wb", _dp),
                ("comp", _dp), ("vsrc", _dp), ("gerr", _dp), ("lam_prev", _dp), ("bias_factor", C.c_double)]


class Result(C.Structure):
    _fields_ = [("iters", C.c_int), ("residual", C.c_double), ("nrows", C.c_int), ("ncolors", C.c_int),
                ("rho", C.c_double)]


_lib = None
_variants = {}


def _declare(L):
    for name in ("orc_pj_solve",):
        getattr(L, name).argtypes = [C.POINTER(Bodies), C.POINTER(Contacts), C.c_double, C.c_double,
                                     C.POINTER(Params), C.c_int, _dp, _dp, C.POINTER(Result)]
        getattr(L, name).restype = C.c_int
    for name in ("orc_pgs_solve", "orc_colored_gs_solve"):
        getattr(L, name).argtypes = [C.POINTER(Bodies), C.POINTER(Contacts), C.c_double, C.c_double,
                                     C.POINTER(Params), _dp, _dp, C.POINTER(Result)]
        getattr(L, name).restype = C.c_int
    L.orc_lstsq_colpiv.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]
    L.orc_lstsq_colpiv.restype = C.c_int
    L.orc_pj_solve_joints.argtypes = [C.POINTER(Bodies), C.POINTER(Contacts), C.POINTER(Joints), C.c_double,
                                      C.c_double, C.POINTER(Params), _dp, _dp, _dp, C.POINTER(Result)]
    L.orc_pj_solve_joints.restype = C.c_int
    L.orc_color_blocks.argtypes = [C.c_int, C.c_int, _ip, _ip, _ip]
    L.orc_color_blocks.restype = C.c_int
    L.orc_tangent_frame.argtypes = [_dp, _dp, _dp]
    L.orc_tangent_frame.restype = None
    L.orc_position_update.argtypes = [C.c_int, _ip, _dp, _dp, C.c_double]
    L.orc_position_update.restype = None
    L.orc_gyro_force.argtypes = [C.c_int, _ip, _dp, _dp, _dp]
    L.orc_gyro_force.restype = None
    L.orc_complementarity.argtypes = [C.POINTER(Bodies), C.POINTER(Contacts), _dp, _dp]
    L.orc_complementarity.restype = C.c_double
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-C", HERE, "-s"], check=True)
        _lib = _declare(C.CDLL(LIB))
    return _lib


def mutant(name):
    """The oracle built with one deliberate mistake (macro ORC_MUT_<name>, `make mutants`): used
    only to show that the pinning tests fail on it."""
    if name not in _variants:
        path = os.path.join(HERE, "_mut", f"libcdx_oracle_{name}.so")
        subprocess.run(["make", "-C", HERE, "-s", os.path.join("_mut", f"libcdx_oracle_{name}.so")], check=True)
        _variants[name] = _declare(C.CDLL(path))
    return _variants[name]


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


class _Pack:
    """Keeps contiguous copies alive while the C structs point at them."""

    def __init__(self, prob):
        f = lambda x: np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        g = lambda x: np.ascontiguousarray(x, dtype=np.int32).reshape(-1)
        self.dof = g(prob.dof)
        self.minv, self.v, self.fext, self.fgyro = f(prob.minv), f(prob.v), f(prob.fext), f(prob.fgyro)
        self.a, self.b = g(prob.body_a), g(prob.body_b)
        self.mu, self.frame, self.angA, self.angB = f(prob.mu), f(prob.frame), f(prob.angA), f(prob.angB)
        self.rest, self.vbias, self.warm = f(prob.rest), f(prob.vbias), f(prob.warm)
        self.B = Bodies(len(self.dof), _i(self.dof), _d(self.minv), _d(self.v), _d(self.fext), _d(self.fgyro))
        self.C = Contacts(len(self.a), _i(self.a), _i(self.b), _d(self.mu), _d(self.frame), _d(self.angA),
                          _d(self.angB), _d(self.rest), _d(self.vbias), _d(self.warm))


def lstsq_colpiv(A, b):
    """min ||A g - b|| by the column-pivoted Householder QR of the Anderson step; returns (g, pivots)."""
    A = np.asfortranarray(np.array(A, dtype=np.float64, copy=True))
    rows, cols = A.shape
    b = np.ascontiguousarray(b, dtype=np.float64)
    g = np.zeros(cols)
    nz = lib().orc_lstsq_colpiv(rows, cols, _d(A.reshape(-1, order="F")), _d(b), _d(g))
    return g, nz


def params(alpha=0.3, relaxation=1.0, max_iter=200000, tol_abs=1e-5, tol_rel=1e-5, warmstart=1, nesterov=0,
           nesterov_beta_threshold=1.0, nesterov_restart_limit=10000, chebyshev=0, block_precond=0, anderson=0,
           anderson_m=2, implicit_gyro=0):
    return Params(alpha, relaxation, max_iter, tol_abs, tol_rel, warmstart, nesterov, nesterov_beta_threshold,
                  nesterov_restart_limit, chebyshev, block_precond, anderson, anderson_m, implicit_gyro)


def _solve(name, prob, dt, prm, nthreads=None, L=None):
    L = L or lib()
    pk = _Pack(prob)
    v = np.zeros(int(pk.dof.sum()))
    imp = np.zeros((len(pk.a), 3))
    res = Result()
    if name == "orc_pj_solve":
        st = L.orc_pj_solve(C.byref(pk.B), C.byref(pk.C), dt, 1.0, C.byref(prm), nthreads or 1, _d(v), _d(imp),
                            C.byref(res))
    else:
        st = getattr(L, name)(C.byref(pk.B), C.byref(pk.C), dt, 1.0, C.byref(prm), _d(v), _d(imp), C.byref(res))
    return st, v, imp, res


def pj_solve_joints(prob, rows, dt, prm, theta=1.0, lam_prev=None, bias_factor=0.001):
    """Extended-system PJ (orc_pj_solve_joints).  `rows`: the dict of joints.spherical_rows (ns,
    nd, body_a, body_b, wa, wb, comp, vsrc, gerr).  Returns (status, v, impulses, Lambda, result)."""
    L = lib()
    pk = _Pack(prob)
    f = lambda x: np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    g = lambda x: np.ascontiguousarray(x, dtype=np.int32).reshape(-1)
    n = int(rows["ns"]) + int(rows["nd"])
    keep = [g(rows["body_a"]), g(rows["body_b"]), f(rows["wa"]), f(rows["wb"]), f(rows["comp"]), f(rows["vsrc"]),
            f(rows["gerr"]), f(np.zeros(n) if lam_prev is None else lam_prev)]
    J = Joints(int(rows["ns"]), int(rows["nd"]), _i(keep[0]), _i(keep[1]), *[_d(x) for x in keep[2:]], bias_factor)
    v = np.zeros(int(pk.dof.sum()))
    imp = np.zeros((len(pk.a), 3))
    lam = np.zeros(max(n, 1))
    res = Result()
    st = L.orc_pj_solve_joints(C.byref(pk.B), C.byref(pk.C), C.byref(J), dt, theta, C.byref(prm), _d(v), _d(imp),
                               _d(lam), C.byref(res))
    return st, v, imp, lam[:n], res


def pj_solve(prob, dt, prm, nthreads=1, L=None):
    """L: a variant library (mutant()) in place of the oracle."""
    return _solve("orc_pj_solve", prob, dt, prm, nthreads, L)


def pgs_solve(prob, dt, prm, L=None):
    return _solve("orc_pgs_solve", prob, dt, prm, L=L)


def colored_gs_solve(prob, dt, prm, L=None):
    return _solve("orc_colored_gs_solve", prob, dt, prm, L=L)


def color_blocks(nb, body_a, body_b):
    a = np.ascontiguousarray(body_a, np.int32)
    b = np.ascontiguousarray(body_b, np.int32)
    col = np.empty(len(a), np.int32)
    n = lib().orc_color_blocks(nb, len(a), _i(a), _i(b), _i(col))
    return n, col


def tangent_frame(n):
    n = np.ascontiguousarray(n, dtype=np.float64)
    t1 = np.empty(3)
    t2 = np.empty(3)
    lib().orc_tangent_frame(_d(n), _d(t1), _d(t2))
    return t1, t2


def position_update(dof, q, v, h):
    dof = np.ascontiguousarray(dof, np.int32)
    q = np.ascontiguousarray(q, dtype=np.float64).copy()
    v = np.ascontiguousarray(v, dtype=np.float64)
    lib().orc_position_update(len(dof), _i(dof), _d(q), _d(v), h)
    return q


def gyro_force(dof, minv, v):
    dof = np.ascontiguousarray(dof, np.int32)
    minv = np.ascontiguousarray(minv, np.float64)
    v = np.ascontiguousarray(v, np.float64)
    f = np.empty_like(v)
    lib().orc_gyro_force(len(dof), _i(dof), _d(minv), _d(v), _d(f))
    return f


def complementarity(prob, v, imp):
    pk = _Pack(prob)
    v = np.ascontiguousarray(v, np.float64)
    imp = np.ascontiguousarray(imp, np.float64).reshape(-1)
    return lib().orc_complementarity(C.byref(pk.B), C.byref(pk.C), _d(v), _d(imp))
