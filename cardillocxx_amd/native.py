"""ctypes bindings of the C ABI (include/cdx/cdx.h) and the CPU contact front end.

The device library must be built in-tree (``python -m cardillocxx_amd.build``); loading fails
loudly when it is missing -- there is no CPU fallback for the solve.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libcdx.so")
SCENE_PATH = os.path.join(PKG, "libcdx_scene.so")

CDX_OK = 0
CDX_NOT_FACTORIZED = 1
CDX_NONFINITE = 2
CDX_DIVERGED = 3
CDX_BAD_ARG = 4
CDX_HIP_ERROR = 5
CDX_NO_STATE = 6
STATUS_NAMES = {
    0: "CDX_OK",
    1: "CDX_NOT_FACTORIZED",
    2: "CDX_NONFINITE",
    3: "CDX_DIVERGED",
    4: "CDX_BAD_ARG",
    5: "CDX_HIP_ERROR",
    6: "CDX_NO_STATE",
}
CDX_ACCEL_NONE = 0
CDX_ACCEL_NESTEROV = 1
CDX_ACCEL_CHEBYSHEV = 2
CDX_ACCEL_ANDERSON = 3

# Every symbol include/cdx/cdx.h declares.
EXPORTED = [
    "cdx_default_params",
    "cdx_ctx_create",
    "cdx_ctx_destroy",
    "cdx_last_error",
    "cdx_synchronize",
    "cdx_set_bodies",
    "cdx_set_contacts",
    "cdx_pj_solve",
    "cdx_set_joints",
    "cdx_get_joint_impulses",
    "cdx_gs_colored_solve",
    "cdx_get_velocities",
    "cdx_get_impulses",
    "cdx_get_positions",
    "cdx_nonfinite_bodies",
    "cdx_moreau_predrift",
    "cdx_moreau_update",
    "cdx_set_profiling",
    "cdx_kernel_time",
    "cdx_set_scenes",
    "cdx_batch_pj_solve",
    "cdx_batch_set_launch",
    "cdx_batch_autotune",
    "cdx_set_recording",
    "cdx_pj_adjoint",
    "cdx_pj_adjoint_joints",
    "cdx_pj_solve_noalloc",
    "cdx_set_graphs",
    "cdx_set_range_sweep",
    "cdx_range_sweep_info",
    "cdx_set_contact_points",
    "cdx_set_geometry_rows",
    "cdx_geometry_rows_info",
    "cdx_kernel_stat",
    "cdx_get_stream",
]
CDX_ADJ_IMPLICIT = 0
CDX_ADJ_UNROLLED = 1
CDX_ADJ_AUTO = 2


class SolverParams(C.Structure):
    _fields_ = [
        ("alpha", C.c_double),
        ("relaxation", C.c_double),
        ("max_iter", C.c_int),
        ("tol_abs", C.c_double),
        ("tol_rel", C.c_double),
        ("check_every", C.c_int),
        ("warmstart", C.c_int),
        ("accel", C.c_int),
        ("nesterov_beta_threshold", C.c_double),
        ("nesterov_restart_limit", C.c_int),
        ("block_precond", C.c_int),
        ("anderson_m", C.c_int),
        ("implicit_gyro", C.c_int),
    ]


class AdjointParams(C.Structure):
    _fields_ = [("mode", C.c_int), ("max_iter", C.c_int), ("tol", C.c_double)]


class AdjointInfo(C.Structure):
    _fields_ = [("iters", C.c_int), ("residual", C.c_double), ("mode_used", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("iters", C.c_int), ("residual", C.c_double), ("nrows", C.c_int), ("ncolors", C.c_int)]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class DeviceProblem(C.Structure):
    """cdx_device_problem: device pointers (as integers) in the layouts of cdx_set_bodies /
    cdx_set_contacts."""
    _fields_ = [("nb", C.c_int), ("nc", C.c_int)] + [(n, _vp) for n in (
        "dof", "minv", "v", "fext", "fgyro", "body_a", "body_b", "mu", "frame", "angA", "angB", "rest", "vbias", "warm")]

_lib = None
_scene = None


def _declare(lib):
    sig = {
        "cdx_default_params": (None, [C.POINTER(SolverParams)]),
        "cdx_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
        "cdx_ctx_destroy": (None, [_vp]),
        "cdx_last_error": (C.c_char_p, [_vp]),
        "cdx_synchronize": (C.c_int, [_vp]),
        "cdx_set_bodies": (C.c_int, [_vp, C.c_int, _ip, _dp, _dp, _dp, _dp, _dp]),
        "cdx_set_contacts": (C.c_int, [_vp, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "cdx_pj_solve": (C.c_int, [_vp, C.c_double, C.c_double, C.POINTER(SolverParams), C.POINTER(SolveInfo)]),
        "cdx_set_joints": (C.c_int, [_vp, C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double]),
        "cdx_get_joint_impulses": (C.c_int, [_vp, _dp]),
        "cdx_gs_colored_solve": (C.c_int, [_vp, C.c_double, C.c_double, C.POINTER(SolverParams), C.POINTER(SolveInfo)]),
        "cdx_get_velocities": (C.c_int, [_vp, _dp]),
        "cdx_get_impulses": (C.c_int, [_vp, _dp]),
        "cdx_get_positions": (C.c_int, [_vp, _dp]),
        "cdx_nonfinite_bodies": (C.c_int, [_vp, _ip]),
        "cdx_moreau_predrift": (C.c_int, [_vp, C.c_double, C.c_double]),
        "cdx_moreau_update": (C.c_int, [_vp, C.c_double, C.c_double]),
        "cdx_set_profiling": (C.c_int, [_vp, C.c_int]),
        "cdx_set_graphs": (C.c_int, [_vp, C.c_int]),
        "cdx_set_range_sweep": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
        "cdx_range_sweep_info": (C.c_int, [_vp, _ip, _ip, _ip, _ip]),
        "cdx_set_contact_points": (C.c_int, [_vp, _dp]),
        "cdx_set_geometry_rows": (C.c_int, [_vp, C.c_int]),
        "cdx_geometry_rows_info": (C.c_int, [_vp, _ip, _ip, _ip]),
        "cdx_get_stream": (C.c_int, [_vp, C.POINTER(_vp)]),
        "cdx_kernel_stat": (C.c_int, [_vp, C.c_int, C.c_char_p, C.c_int, _ip, _dp]),
        "cdx_kernel_time": (C.c_int, [_vp, C.c_char_p, _ip, _dp]),
        "cdx_set_scenes": (C.c_int, [_vp, C.c_int, _ip, _ip]),
        "cdx_set_recording": (C.c_int, [_vp, C.c_int, C.c_int]),
        "cdx_pj_adjoint": (C.c_int, [_vp, C.c_double, C.POINTER(SolverParams), C.POINTER(AdjointParams), _dp, _dp, _dp, _dp,
                                     _dp, _dp, C.POINTER(AdjointInfo)]),
        "cdx_pj_adjoint_joints": (C.c_int, [_vp, C.c_double, C.POINTER(SolverParams), C.POINTER(AdjointParams), _dp, _dp,
                                            _dp, _dp, _dp, _dp, C.POINTER(AdjointInfo)]),
        "cdx_batch_pj_solve": (C.c_int, [_vp, C.c_double, C.c_double, C.POINTER(SolverParams), _ip, _dp]),
        "cdx_batch_set_launch": (C.c_int, [_vp, C.c_int, C.c_int]),
        "cdx_batch_autotune": (C.c_int, [_vp, C.c_double, C.c_double, C.POINTER(SolverParams), _ip, _ip, _dp]),
        "cdx_pj_solve_noalloc": (C.c_int, [C.POINTER(DeviceProblem), C.c_double, C.c_double, C.POINTER(SolverParams),
                                           _vp, _vp, C.POINTER(SolveInfo), _vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


def lib():
    """The HIP device library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m cardillocxx_amd.build`")
        _lib = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        _declare(_lib)
    return _lib


class Heightfield(C.Structure):
    """cdx_heightfield (cdx_scene.h): samples h[i + nx*j] at (x0 + i*dx, y0 + j*dy)."""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("x0", C.c_double), ("y0", C.c_double), ("dx", C.c_double),
                ("dy", C.c_double), ("h", _dp), ("mu", C.c_double)]


def scene_lib():
    global _scene
    if _scene is None:
        if not os.path.exists(SCENE_PATH):
            raise RuntimeError(f"{SCENE_PATH} is missing: build it with `python -m cardillocxx_amd.build`")
        _scene = C.CDLL(SCENE_PATH)
        _scene.cdx_tangent_frame.restype = None
        _scene.cdx_tangent_frame.argtypes = [_dp, _dp, _dp]
        _scene.cdx_sphere_contacts.restype = C.c_int
        _scene.cdx_sphere_contacts.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, _dp, _dp, _dp, C.c_double, C.c_int,
                                               _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp]
        _scene.cdx_shape_contacts.restype = C.c_int
        _scene.cdx_shape_contacts.argtypes = [C.c_int, _ip, _dp, _ip, _dp, _dp, C.c_int, _dp, _dp, _dp, C.c_double, C.c_int,
                                              _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _ip]
        _scene.cdx_shape_contacts_hf.restype = C.c_int
        _scene.cdx_shape_contacts_hf.argtypes = [C.c_int, _ip, _dp, _ip, _dp, _dp, C.c_int, _dp, _dp, _dp,
                                                 C.POINTER(Heightfield), C.c_int, _ip, _ip, C.c_double, C.c_int, _ip, _ip,
                                                 _dp, _dp, _dp, _dp, _dp, _dp, _ip]
        _scene.cdx_scene_rsa.restype = C.c_int
        _scene.cdx_scene_rsa.argtypes = [C.c_int, _ip, _dp, _dp, _dp, C.c_uint64, _dp]
        _scene.cdx_match_warmstart.restype = C.c_int
        _scene.cdx_match_warmstart.argtypes = [C.c_int, _ip, _ip, _dp, _dp, C.c_int, _ip, _ip, _dp, C.c_double, _dp]
        _scene.cdx_scene_sphere_grid.restype = C.c_int
        _scene.cdx_scene_sphere_grid.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                                 C.c_uint64, _dp]
        _scene.cdx_scene_sphere_pile.restype = C.c_int
        _scene.cdx_scene_sphere_pile.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                                 C.c_uint64, C.c_int, _dp, _dp]
    return _scene


def dptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "float64 C-contiguous array required"
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"], "int32 C-contiguous array required"
    return a.ctypes.data_as(_ip)


def default_params() -> SolverParams:
    p = SolverParams()
    lib().cdx_default_params(C.byref(p))
    return p


class CdxError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class Context:
    """One device context (one HIP stream) -- a thin owner of ``cdx_ctx*``."""

    def __init__(self, device: int = 0):
        self._l = lib()
        h = _vp()
        st = self._l.cdx_ctx_create(device, C.byref(h))
        if st != CDX_OK:
            raise CdxError(st, f"cdx_ctx_create(device={device}) failed")
        self.h = h
        self.nb = 0
        self.nv = 0
        self.nq = 0
        self.nc = 0

    def close(self):
        if getattr(self, "h", None):
            self._l.cdx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int, what: str):
        if st != CDX_OK:
            raise CdxError(st, f"{what}: {self._l.cdx_last_error(self.h).decode()}")

    def set_bodies(self, dof, minv, v, fext, fgyro=None, q=None):
        dof = np.ascontiguousarray(dof, dtype=np.int32)
        arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (minv, v, fext)]
        fg = None if fgyro is None else np.ascontiguousarray(fgyro, dtype=np.float64)
        qq = None if q is None else np.ascontiguousarray(q, dtype=np.float64)
        self._check(self._l.cdx_set_bodies(self.h, len(dof), iptr(dof), *[dptr(a) for a in arr], dptr(fg), dptr(qq)),
                    "cdx_set_bodies")
        self.nb = len(dof)
        self.nv = int(dof.sum())
        self.nq = int(np.where(dof == 6, 7, 3).sum())

    def set_contacts(self, body_a, body_b, mu, frame, angA, angB, rest, vbias, warm):
        ia = np.ascontiguousarray(body_a, dtype=np.int32)
        ib = np.ascontiguousarray(body_b, dtype=np.int32)
        d = [np.ascontiguousarray(x, dtype=np.float64).reshape(-1) for x in (mu, frame, angA, angB, rest, vbias, warm)]
        self._check(self._l.cdx_set_contacts(self.h, len(ia), iptr(ia), iptr(ib), *[dptr(x) for x in d]),
                    "cdx_set_contacts")
        self.nc = len(ia)

    def set_joints(self, rows, lam_prev=None, bias_factor=0.001):
        """Joint rows in the layout of joints.spherical_rows (ns, nd, body_a, body_b, wa, wb, comp,
        vsrc, gerr); lam_prev: Lambda of the previous step (zeros if None)."""
        ns, nd = int(rows["ns"]), int(rows["nd"])
        n = ns + nd
        f = lambda x: np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        a = np.ascontiguousarray(rows["body_a"], dtype=np.int32)
        b = np.ascontiguousarray(rows["body_b"], dtype=np.int32)
        arr = [f(rows["wa"]), f(rows["wb"]), f(rows["comp"]), f(rows["vsrc"]), f(rows["gerr"]),
               f(np.zeros(n) if lam_prev is None else lam_prev)]
        self._check(self._l.cdx_set_joints(self.h, ns, nd, iptr(a), iptr(b), *[dptr(x) for x in arr], bias_factor),
                    "cdx_set_joints")
        self.njrows = n

    def joint_impulses(self):
        out = np.zeros(max(getattr(self, "njrows", 0), 1))
        self._check(self._l.cdx_get_joint_impulses(self.h, dptr(out)), "cdx_get_joint_impulses")
        return out[:getattr(self, "njrows", 0)]

    def _solve(self, fn, dt, theta, params):
        info = SolveInfo()
        st = fn(self.h, dt, theta, C.byref(params), C.byref(info))
        return st, info

    def pj_solve(self, dt, theta, params):
        st, info = self._solve(self._l.cdx_pj_solve, dt, theta, params)
        self._check(st, "cdx_pj_solve")
        return info

    def gs_colored_solve(self, dt, theta, params, raise_on_divergence=True):
        st, info = self._solve(self._l.cdx_gs_colored_solve, dt, theta, params)
        if st != CDX_OK and (raise_on_divergence or st != CDX_DIVERGED):
            self._check(st, "cdx_gs_colored_solve")
        return info

    def set_scenes(self, body_off, contact_off):
        bo = np.ascontiguousarray(body_off, dtype=np.int32)
        co = np.ascontiguousarray(contact_off, dtype=np.int32)
        self._check(self._l.cdx_set_scenes(self.h, len(bo) - 1, iptr(bo), iptr(co)), "cdx_set_scenes")
        self.nscenes = len(bo) - 1

    def batch_pj_solve(self, dt, theta, params):
        it = np.zeros(self.nscenes, np.int32)
        res = np.zeros(self.nscenes)
        self._check(self._l.cdx_batch_pj_solve(self.h, dt, theta, C.byref(params), iptr(it), dptr(res)),
                    "cdx_batch_pj_solve")
        return it, res

    def batch_set_launch(self, block, groups):
        self._check(self._l.cdx_batch_set_launch(self.h, block, groups), "cdx_batch_set_launch")

    def batch_autotune(self, dt, theta, params):
        b, g, ms = C.c_int(0), C.c_int(0), C.c_double(0)
        self._check(self._l.cdx_batch_autotune(self.h, dt, theta, C.byref(params), C.byref(b), C.byref(g), C.byref(ms)),
                    "cdx_batch_autotune")
        return b.value, g.value, ms.value

    def set_recording(self, on, max_sweeps=0):
        self._check(self._l.cdx_set_recording(self.h, 1 if on else 0, max_sweeps), "cdx_set_recording")

    def pj_adjoint(self, dt, params, g_v, g_p=None, mode=0, max_iter=100000, tol=1e-12):
        """dL/dv_n, dL/dM^-1, dL/df_ext (stacked, nv) and dL/dmu (per contact) of the last PJ solve."""
        a = AdjointParams(mode, max_iter, tol)
        info = AdjointInfo()
        gv = np.ascontiguousarray(g_v, dtype=np.float64)
        gp = None if g_p is None else np.ascontiguousarray(g_p, dtype=np.float64).reshape(-1)
        ov, om, of = np.zeros(self.nv), np.zeros(self.nv), np.zeros(self.nv)
        omu = np.zeros(max(self.nc, 1))
        self._check(self._l.cdx_pj_adjoint(self.h, dt, C.byref(params), C.byref(a), dptr(gv), dptr(gp), dptr(ov), dptr(om),
                                           dptr(of), dptr(omu), C.byref(info)), "cdx_pj_adjoint")
        return dict(v=ov, minv=om, fext=of, mu=omu[:self.nc], info=info)

    def pj_adjoint_joints(self, dt, params, g_v, g_p=None, max_iter=100000, tol=1e-12, mode=CDX_ADJ_IMPLICIT):
        """dL/dv_n, dL/df_ext (stacked, nv), dL/dmu per contact and dL/dcomp per joint row (caller's
        row order) of the last joint solve, by the implicit adjoint (mode CDX_ADJ_AUTO accepts the
        least-squares adjoint where the system is inconsistent; info.residual tells)."""
        a = AdjointParams(mode, max_iter, tol)
        info = AdjointInfo()
        gv = np.ascontiguousarray(g_v, dtype=np.float64)
        gp = None if g_p is None else np.ascontiguousarray(g_p, dtype=np.float64).reshape(-1)
        ov, of = np.zeros(self.nv), np.zeros(self.nv)
        omu = np.zeros(max(self.nc, 1))
        oc = np.zeros(max(self.njrows, 1))
        self._check(self._l.cdx_pj_adjoint_joints(self.h, dt, C.byref(params), C.byref(a), dptr(gv), dptr(gp), dptr(ov),
                                                  dptr(of), dptr(omu), dptr(oc), C.byref(info)), "cdx_pj_adjoint_joints")
        return dict(v=ov, fext=of, mu=omu[:self.nc], comp=oc[:self.njrows], info=info)

    def velocities(self):
        out = np.empty(self.nv, dtype=np.float64)
        self._check(self._l.cdx_get_velocities(self.h, dptr(out)), "cdx_get_velocities")
        return out

    def impulses(self):
        out = np.empty((self.nc, 3), dtype=np.float64)
        self._check(self._l.cdx_get_impulses(self.h, dptr(out)), "cdx_get_impulses")
        return out

    def positions(self):
        out = np.empty(self.nq, dtype=np.float64)
        self._check(self._l.cdx_get_positions(self.h, dptr(out)), "cdx_get_positions")
        return out

    def nonfinite_bodies(self):
        """Bodies whose solved twist or pose holds a NaN or Inf, counted on the device."""
        n = C.c_int(0)
        self._check(self._l.cdx_nonfinite_bodies(self.h, C.byref(n)), "cdx_nonfinite_bodies")
        return n.value

    def moreau_predrift(self, dt, theta):
        self._check(self._l.cdx_moreau_predrift(self.h, dt, theta), "cdx_moreau_predrift")

    def moreau_update(self, dt, theta):
        self._check(self._l.cdx_moreau_update(self.h, dt, theta), "cdx_moreau_update")

    def synchronize(self):
        self._check(self._l.cdx_synchronize(self.h), "cdx_synchronize")

    def set_profiling(self, on: bool):
        self._check(self._l.cdx_set_profiling(self.h, 1 if on else 0), "cdx_set_profiling")

    def kernel_stats(self):
        """{kernel name: (calls, total ms)} of every kernel timed since profiling was enabled."""
        out, buf = {}, C.create_string_buffer(64)
        calls, ms = C.c_int(0), C.c_double(0)
        i = 0
        while self._l.cdx_kernel_stat(self.h, i, buf, 64, C.byref(calls), C.byref(ms)) == CDX_OK:
            out[buf.value.decode()] = (calls.value, ms.value)
            i += 1
        return out

    def stream(self):
        """The context's hipStream_t as an integer."""
        s = _vp()
        self._check(self._l.cdx_get_stream(self.h, C.byref(s)), "cdx_get_stream")
        return s.value or 0

    def set_graphs(self, on: bool):
        self._check(self._l.cdx_set_graphs(self.h, 1 if on else 0), "cdx_set_graphs")

    def set_range_sweep(self, mode: int, contacts_per_range: int = 0, block: int = 0):
        """Range sweep of projected Jacobi: mode 0 off, 1 from 32,768 contacts (default), 2 always."""
        self._check(self._l.cdx_set_range_sweep(self.h, mode, contacts_per_range, block), "cdx_set_range_sweep")

    def set_contact_points(self, point):
        """World contact points [nc, 3] for the geometry form of the rows (cdx_set_contact_points)."""
        pt = np.ascontiguousarray(point, dtype=np.float64).reshape(-1)
        self._check(self._l.cdx_set_contact_points(self.h, dptr(pt)), "cdx_set_contact_points")

    def set_geometry_rows(self, mode: int):
        self._check(self._l.cdx_set_geometry_rows(self.h, mode), "cdx_set_geometry_rows")

    def geometry_rows_info(self):
        """(usable, contacts whose rebuilt rows differ or -1, used by the last solve)."""
        v = [C.c_int(0) for _ in range(3)]
        self._check(self._l.cdx_geometry_rows_info(self.h, *[C.byref(x) for x in v]), "cdx_geometry_rows_info")
        return bool(v[0].value), v[1].value, bool(v[2].value)

    def range_sweep_info(self):
        """(used by the last solve, ranges, partial sums per sweep, contacts per range)."""
        v = [C.c_int(0) for _ in range(4)]
        self._check(self._l.cdx_range_sweep_info(self.h, *[C.byref(x) for x in v]), "cdx_range_sweep_info")
        return bool(v[0].value), v[1].value, v[2].value, v[3].value

    def kernel_time(self, name: str):
        calls = C.c_int(0)
        ms = C.c_double(0)
        self._check(self._l.cdx_kernel_time(self.h, name.encode(), C.byref(calls), C.byref(ms)), "cdx_kernel_time")
        return calls.value, ms.value


def pj_solve_noalloc(nb, nc, ptrs, dt, theta, params, v_out, imp_out, stream=None):
    """cdx_pj_solve_noalloc on device arrays given as integer device addresses: `ptrs` maps the
    cdx_device_problem field names (dof, minv, v, fext, fgyro, body_a, ..., warm) to addresses (0 or
    missing = NULL); v_out / imp_out are output addresses; `stream` a hipStream_t as an integer."""
    L = lib()
    pb = DeviceProblem(nb, nc, *[ptrs.get(n) or None for n, _ in DeviceProblem._fields_[2:]])
    info = SolveInfo()
    st = L.cdx_pj_solve_noalloc(C.byref(pb), dt, theta, C.byref(params), v_out or None, imp_out or None,
                                C.byref(info), stream or None)
    if st != CDX_OK:
        raise CdxError(st, f"cdx_pj_solve_noalloc: {L.cdx_last_error(None).decode()}")
    return info
