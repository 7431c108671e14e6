"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker.  The product path never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# QPSO_ORACLE_LIB selects another build of the oracle (the sanitizer build
# of tests/test_sanitizers.py)
LIB_PATH = os.environ.get("QPSO_ORACLE_LIB") or os.path.join(HERE, "_build", "liboracle.so")

ROWS, COLS = 480, 640
EMPTY = 60000


class Cam(C.Structure):
    _fields_ = [("fx", C.c_double), ("fy", C.c_double), ("cx", C.c_double), ("cy", C.c_double)]


class Params(C.Structure):
    _fields_ = [
        ("n_particles", C.c_int), ("n_iters", C.c_int), ("w", C.c_double),
        ("c1", C.c_float), ("c2", C.c_float),
        ("x_lo", C.c_float * 3), ("x_hi", C.c_float * 3),
        ("rect", C.c_int * 4), ("num_points", C.c_uint32), ("face_cap", C.c_int),
        ("cam", Cam), ("wrap", C.c_int),
    ]


class RenderInfo(C.Structure):
    _fields_ = [("aabb", C.c_int32 * 4), ("degenerate", C.c_int), ("raster_px", C.c_uint64)]


class FitnessInfo(C.Structure):
    _fields_ = [("seg", C.c_uint32), ("tr", C.c_uint32), ("n_rendered", C.c_uint32),
                ("err", C.c_float), ("visit_px", C.c_uint64)]


class SolveOut(C.Structure):
    _fields_ = [
        ("best_pose", C.c_float * 7), ("best_fit", C.c_float),
        ("fit_trace", C.c_void_p), ("gbest_trace", C.c_void_p), ("argmin_trace", C.c_void_p),
        ("pos", C.c_void_p), ("vel", C.c_void_p), ("pbest", C.c_void_p),
        ("pbest_fit", C.c_void_p), ("frames", C.c_void_p),
        ("raster_px", C.c_uint64), ("visit_px", C.c_uint64), ("n_degenerate", C.c_int),
    ]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def load(path):
    """A CDLL of an oracle build at `path`, with its prototypes set."""
    L = C.CDLL(path)
    P = C.c_void_p
    L.or_mt19937_raw.argtypes = [C.c_uint32, P, C.c_size_t]
    L.or_rand_table.argtypes = [C.c_uint32, P, C.c_size_t]
    L.or_rand_stride.argtypes = [C.c_int]
    L.or_rand_stride.restype = C.c_size_t
    for n in ("or_sinf", "or_cosf", "or_acosf", "or_rcp"):
        getattr(L, n).argtypes = [C.c_float]
        getattr(L, n).restype = C.c_float
    L.or_rint_sat.argtypes = [C.c_float]
    L.or_rint_sat.restype = C.c_int32
    L.or_project.argtypes = [P, P, P, P]
    L.or_render.argtypes = [P, C.c_int, C.c_int, P, P, P, P, P]
    L.or_fitness.argtypes = [P, P, C.c_uint32, P, P, P]
    L.or_fitness.restype = C.c_float
    L.or_render_ex.argtypes = [P, C.c_int, C.c_int, P, P, P, C.c_int, P, P]
    L.or_fitness_ex.argtypes = [P, P, C.c_uint32, C.c_int, P, P, P]
    L.or_fitness_ex.restype = C.c_float
    L.or_argmin.argtypes = [P, C.c_int]
    L.or_argmin.restype = C.c_int
    L.or_init.argtypes = [P, P, P, P, P, P, P]
    L.or_update_best.argtypes = [C.c_int, P, P, P, P, P, C.c_int]
    L.or_update_particles.argtypes = [P, P, C.c_int, P, P, P, P]
    L.or_kinematics.argtypes = [P, P]
    L.or_solve.argtypes = [P, C.c_int, P, P, P, C.c_int, P]
    L.or_solve.restype = C.c_int
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load(LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def cam_struct(cam):
    return Cam(*[float(x) for x in cam])


def make_params(n, iters, w, c1, c2, x_lo, x_hi, rect, num_points, cam, face_cap=0, wrap=0):
    p = Params()
    p.wrap = int(wrap)
    p.n_particles, p.n_iters, p.w, p.c1, p.c2 = n, iters, w, c1, c2
    for i in range(3):
        p.x_lo[i], p.x_hi[i] = float(x_lo[i]), float(x_hi[i])
    for i in range(4):
        p.rect[i] = int(rect[i])
    p.num_points, p.face_cap, p.cam = int(num_points), int(face_cap), cam_struct(cam)
    return p


def mt_raw(seed, n):
    out = np.empty(n, np.uint32)
    lib().or_mt19937_raw(seed, _p(out), n)
    return out


def rand_table(seed, n_particles, n_iters):
    s = lib().or_rand_stride(n_iters)
    out = np.empty(s * n_particles, np.float32)
    lib().or_rand_table(seed, _p(out), out.size)
    return out


def project(pose, v, cam):
    pose = np.ascontiguousarray(pose, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    out = np.empty(3, np.float32)
    c = cam_struct(cam)
    lib().or_project(_p(pose), _p(v), C.byref(c), _p(out))
    return out


def render(mesh, pose, cam, rect, frame=None, face_cap=0, wrap=0):
    """Render one particle into `frame` (int32 480x640, min-accumulated)."""
    mesh = np.ascontiguousarray(mesh, np.float32)
    F = mesh.size // 9
    if frame is None:
        frame = np.full((ROWS, COLS), EMPTY, np.int32)
    pose = np.ascontiguousarray(pose, np.float32)
    r = (C.c_int * 4)(*[int(x) for x in rect])
    c = cam_struct(cam)
    info = RenderInfo()
    lib().or_render_ex(_p(mesh), F, face_cap, _p(pose), C.byref(c), r, int(wrap), _p(frame), C.byref(info))
    return frame, info


def fitness(rinfo, rect, num_points, frame, obs, wrap=0):
    r = (C.c_int * 4)(*[int(x) for x in rect])
    obs = np.ascontiguousarray(obs, np.uint16)
    fi = FitnessInfo()
    f = lib().or_fitness_ex(C.byref(rinfo), r, int(num_points), int(wrap), _p(frame), _p(obs), C.byref(fi))
    return np.float32(f), fi


def argmin(fit):
    fit = np.ascontiguousarray(fit, np.float32)
    return lib().or_argmin(_p(fit), fit.size)


def solve(mesh, obs, params, rand, n_threads=1, want_frames=False):
    mesh = np.ascontiguousarray(mesh, np.float32)
    obs = np.ascontiguousarray(obs, np.uint16)
    rand = np.ascontiguousarray(rand, np.float32)
    N, I = params.n_particles, params.n_iters
    res = {
        "fit_trace": np.zeros((I, N), np.float32),
        "gbest_trace": np.zeros((I, 8), np.float32),
        "argmin_trace": np.zeros(I, np.int32),
        "pos": np.zeros((7, N), np.float32),
        "vel": np.zeros((7, N), np.float32),
        "pbest": np.zeros((7, N), np.float32),
        "pbest_fit": np.zeros(N, np.float32),
    }
    if want_frames:
        res["frames"] = np.zeros((N, ROWS, COLS), np.int32)
    o = SolveOut()
    for k, v in res.items():
        setattr(o, k, v.ctypes.data)
    rc = lib().or_solve(_p(mesh), mesh.size // 9, _p(obs), C.byref(params), _p(rand),
                        n_threads, C.byref(o))
    if rc != 0:
        raise RuntimeError(f"or_solve failed: {rc}")
    res["best_pose"] = np.array(o.best_pose[:], np.float32)
    res["best_fit"] = np.float32(o.best_fit)
    res["raster_px"] = int(o.raster_px)
    res["visit_px"] = int(o.visit_px)
    res["n_degenerate"] = int(o.n_degenerate)
    return res
