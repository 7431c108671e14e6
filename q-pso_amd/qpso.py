"""ctypes binding of libqpso.so (include/qpso.h, qpso_scene.h, qpso_io.h).

Host-side mirror of the reference's estimate-pose block for Python callers
(tests, bench): every call goes through the C ABI into the gfx950 kernels.
There is no CPU fallback: if the library or a GPU is missing, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libqpso.so")

ROWS, COLS = 480, 640

OK, E_INVAL, E_HIP, E_NOMEM, E_BUDGET, E_NOTFOUND, E_STALE, E_PENDING = 0, -1, -2, -3, -4, -5, -6, -7
E_CANCELLED, E_TIMEOUT = -8, -9

# Symbols every header in include/ declares (checked by the CPU test suite).
EXPORTS = [
    "qpso_params_default", "qpso_create", "qpso_destroy", "qpso_last_error", "qpso_status_str",
    "qpso_mesh_upload", "qpso_mesh_release", "qpso_estimate_pose", "qpso_estimate_pose_noalloc", "qpso_solve_create",
    "qpso_solve_run", "qpso_solve_get_trace", "qpso_solve_destroy", "qpso_solve_run_timed",
    "qpso_solve_run_reference_launches",
    "qpso_solve_run_batch", "qpso_autotune", "qpso_set_tuning",
    "qpso_solve_eval_poses", "qpso_server_create", "qpso_server_set_slots", "qpso_server_submit", "qpso_server_poll",
    "qpso_server_wait", "qpso_server_cancel", "qpso_server_set_timeout", "qpso_server_stats_get",
    "qpso_server_destroy", "qpso_solve_bytes",
    "qpso_mesh_icosphere", "qpso_mesh_uv_ellipsoid", "qpso_random_quat", "qpso_scene_render",
    "qpso_scene_cluster", "qpso_read_obj", "qpso_read_dpt", "qpso_write_dpt", "qpso_free",
    "qpso_write_bestpose", "qpso_render_cpu", "qpso_render_cpu_csv", "qpso_clusters_from_labels",
]


class Camera(C.Structure):
    _fields_ = [("fx", C.c_double), ("fy", C.c_double), ("cx", C.c_double), ("cy", C.c_double)]


class Params(C.Structure):
    _fields_ = [
        ("n_particles", C.c_int32), ("n_iters", C.c_int32), ("w", C.c_double),
        ("c1", C.c_float), ("c2", C.c_float), ("t_bound", C.c_float), ("t_bound_z", C.c_float),
        ("seed", C.c_uint32), ("cam", Camera), ("face_cap", C.c_int32), ("offframe", C.c_int32),
    ]


class Cluster(C.Structure):
    _fields_ = [("rect", C.c_int32 * 4), ("centroid_mm", C.c_float * 3),
                ("num_points", C.c_uint32), ("depth", C.c_void_p)]


class Result(C.Structure):
    _fields_ = [("pose", C.c_float * 7), ("fit", C.c_float), ("raster_px", C.c_uint64),
                ("visit_px", C.c_uint64), ("n_degenerate", C.c_uint32), ("status", C.c_int32),
                ("latency_ms", C.c_double)]


class Trace(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in
                ("fit_trace", "gbest_trace", "argmin_trace", "pos", "vel", "pbest", "pbest_fit", "frames")]


class TuneEntry(C.Structure):
    _fields_ = [("config", C.c_int32), ("threads", C.c_int32), ("tile_words", C.c_int32),
                ("big_face", C.c_int32), ("ms", C.c_double)]


class Timing(C.Structure):
    _fields_ = [("ms_solve", C.c_double), ("ms_render_fitness", C.c_double), ("ms_pso", C.c_double),
                ("ms_init", C.c_double), ("launches", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double),
                ("solves_per_s", C.c_double), ("steps", C.c_uint64), ("mean_active", C.c_double),
                ("cancelled", C.c_uint64), ("timed_out", C.c_uint64)]


class SceneObject(C.Structure):
    _fields_ = [("mesh_soa", C.c_void_p), ("n_faces", C.c_int), ("pose", C.c_float * 7)]


class ClusterRoi(C.Structure):
    _fields_ = [("rect", C.c_int * 4), ("centroid_mm", C.c_float * 3), ("num_points", C.c_uint32)]


class QpsoError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"qpso status {code}: {msg}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise QpsoError(E_NOTFOUND, f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        P, I32, U32, U64 = C.c_void_p, C.c_int32, C.c_uint32, C.c_uint64
        sig = {
            "qpso_params_default": (None, [P]),
            "qpso_create": (I32, [C.c_int, P]),
            "qpso_destroy": (I32, [P]),
            "qpso_last_error": (C.c_char_p, [P]),
            "qpso_status_str": (C.c_char_p, [C.c_int]),
            "qpso_mesh_upload": (I32, [P, P, U32, P]),
            "qpso_mesh_release": (I32, [P, U32]),
            "qpso_estimate_pose": (I32, [P, U32, P, P, P]),
            "qpso_estimate_pose_noalloc": (I32, [P, U32, P, P, P]),
            "qpso_solve_create": (I32, [P, U32, P, P, C.c_int, P]),
            "qpso_solve_run": (I32, [P, P]),
            "qpso_solve_get_trace": (I32, [P, P]),
            "qpso_solve_destroy": (I32, [P]),
            "qpso_solve_run_timed": (I32, [P, P, P]),
            "qpso_solve_run_reference_launches": (I32, [P, P]),
            "qpso_solve_run_batch": (I32, [P, C.c_int, P, P]),
            "qpso_solve_eval_poses": (I32, [P, C.c_int, P, P, P, P]),
            "qpso_autotune": (I32, [P, C.c_int, P, C.c_int, P, C.c_int, P]),
            "qpso_set_tuning": (I32, [P, C.c_int, C.c_int]),
            "qpso_server_create": (I32, [P, U64, P]),
            "qpso_server_set_slots": (I32, [P, C.c_int64]),
            "qpso_server_submit": (I32, [P, U32, P, P, P]),
            "qpso_server_poll": (I32, [P, U64, P]),
            "qpso_server_wait": (I32, [P, U64, P]),
            "qpso_server_cancel": (I32, [P, U64]),
            "qpso_server_set_timeout": (I32, [P, C.c_double]),
            "qpso_server_stats_get": (I32, [P, P]),
            "qpso_server_destroy": (I32, [P]),
            "qpso_solve_bytes": (U64, [P, P]),
            "qpso_mesh_icosphere": (C.c_int, [C.c_int, C.c_float, C.c_float, C.c_float, P]),
            "qpso_mesh_uv_ellipsoid": (C.c_int, [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                                 C.c_int, C.c_float, C.c_float, U64, P]),
            "qpso_random_quat": (None, [U64, P]),
            "qpso_scene_render": (C.c_int, [P, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_float, C.c_float, C.c_float, U64, P, P]),
            "qpso_scene_cluster": (C.c_int, [P, P, C.c_int, C.c_double, C.c_double, C.c_double,
                                             C.c_double, P, P]),
            "qpso_read_obj": (I32, [C.c_char_p, P, P]),
            "qpso_read_dpt": (I32, [C.c_char_p, P, P, P]),
            "qpso_write_dpt": (I32, [C.c_char_p, P, I32, I32]),
            "qpso_free": (None, [P]),
            "qpso_write_bestpose": (I32, [C.c_char_p, C.c_float, P]),
            "qpso_render_cpu": (I32, [P, P, U32, P, P]),
            "qpso_render_cpu_csv": (I32, [C.c_char_p, P, P, U32, P]),
            "qpso_clusters_from_labels": (I32, [P, P, C.c_int, P, P, P, P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _check(rc, ctx=None):
    if rc != OK:
        msg = lib().qpso_status_str(rc).decode()
        if ctx is not None and getattr(ctx, "handle", None):
            msg += " / " + lib().qpso_last_error(ctx.handle).decode()
        raise QpsoError(rc, msg)


# ---------------------------------------------------------------- parameters
def default_params():
    p = Params()
    lib().qpso_params_default(C.byref(p))
    return p


def make_params(n_particles, n_iters, w=0.3, c1=1.0, c2=1.0, cam=(525.0, 525.0, 319.5, 239.5),
                t_bound=0.3, t_bound_z=0.3, seed=0, face_cap=0, offframe=0):
    p = default_params()
    p.offframe = offframe
    p.n_particles, p.n_iters, p.w, p.c1, p.c2 = n_particles, n_iters, w, c1, c2
    p.t_bound, p.t_bound_z, p.seed, p.face_cap = t_bound, t_bound_z, seed, face_cap
    p.cam = Camera(*[float(x) for x in cam])
    return p


def make_cluster(rect, centroid_mm, num_points, depth):
    depth = np.ascontiguousarray(depth, np.uint16).reshape(ROWS, COLS)
    c = Cluster()
    for i in range(4):
        c.rect[i] = int(rect[i])
    for i in range(3):
        c.centroid_mm[i] = float(centroid_mm[i])
    c.num_points = int(num_points)
    c.depth = depth.ctypes.data
    c._keep = depth  # keep the buffer alive with the struct
    return c


# ---------------------------------------------------------------- context
class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().qpso_create(device, C.byref(h)))
        self.handle = h
        self._meshes = {}

    def close(self):
        if self.handle:
            lib().qpso_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload_mesh(self, soa_m):
        soa = np.ascontiguousarray(soa_m, np.float32)
        mid = C.c_uint32()
        _check(lib().qpso_mesh_upload(self.handle, _p(soa), soa.size // 9, C.byref(mid)), self)
        return mid.value

    def release_mesh(self, mid):
        _check(lib().qpso_mesh_release(self.handle, mid), self)

    def estimate_pose(self, mesh_id, cluster, params):
        r = Result()
        _check(lib().qpso_estimate_pose(self.handle, mesh_id, C.byref(cluster), C.byref(params), C.byref(r)), self)
        return r

    def solve(self, mesh_id, cluster, params, trace=False):
        return Solve(self, mesh_id, cluster, params, trace)

    def run_batch(self, solves):
        arr = (C.c_void_p * len(solves))(*[s.handle for s in solves])
        res = (Result * len(solves))()
        _check(lib().qpso_solve_run_batch(self.handle, len(solves), arr, res), self)
        return list(res)

    def autotune(self, solves, iters=3):
        """Time every configuration on these solves (run together) and keep
        the fastest in each and as the context's default; returns the
        measured table, fastest first."""
        arr = (C.c_void_p * len(solves))(*[s.handle for s in solves])
        table = (TuneEntry * 64)()
        n = C.c_int()
        _check(lib().qpso_autotune(self.handle, len(solves), arr, int(iters), table, 64, C.byref(n)), self)
        rows = [dict(config=t.config, threads=t.threads, tile_words=t.tile_words, big_face=t.big_face, ms=t.ms)
                for t in table[:min(n.value, 64)]]
        return sorted(rows, key=lambda r: r["ms"])

    def set_tuning(self, config, big_face):
        """Default configuration and big-face threshold of solves and servers created from now on."""
        _check(lib().qpso_set_tuning(self.handle, int(config), int(big_face)), self)


class Solve:
    def __init__(self, ctx, mesh_id, cluster, params, trace=False):
        self.ctx = ctx
        self.N, self.I = params.n_particles, params.n_iters
        h = C.c_void_p()
        _check(lib().qpso_solve_create(ctx.handle, mesh_id, C.byref(cluster), C.byref(params),
                                       1 if trace else 0, C.byref(h)), ctx)
        self.handle = h

    def close(self):
        if self.handle:
            lib().qpso_solve_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def run(self):
        r = Result()
        _check(lib().qpso_solve_run(self.handle, C.byref(r)), self.ctx)
        return r

    def run_reference_launches(self):
        """The same solve launched the reference's way (one launch per particle, 16 streams, a host read per iteration)."""
        r = Result()
        _check(lib().qpso_solve_run_reference_launches(self.handle, C.byref(r)), self.ctx)
        return r

    def run_timed(self):
        r, t = Result(), Timing()
        _check(lib().qpso_solve_run_timed(self.handle, C.byref(r), C.byref(t)), self.ctx)
        return r, t

    def trace(self, frames=False):
        N, I = self.N, self.I
        out = {
            "fit_trace": np.zeros((I, N), np.float32), "gbest_trace": np.zeros((I, 8), np.float32),
            "argmin_trace": np.zeros(I, np.int32), "pos": np.zeros((7, N), np.float32),
            "vel": np.zeros((7, N), np.float32), "pbest": np.zeros((7, N), np.float32),
            "pbest_fit": np.zeros(N, np.float32),
        }
        if frames:
            out["frames"] = np.zeros((N, ROWS, COLS), np.int32)
        t = Trace()
        for k, v in out.items():
            setattr(t, k, v.ctypes.data)
        _check(lib().qpso_solve_get_trace(self.handle, C.byref(t)), self.ctx)
        return out

    def eval_poses(self, poses_7xn, frames=False):
        poses = np.ascontiguousarray(poses_7xn, np.float32)
        n = poses.shape[1]
        fit = np.zeros(n, np.float32)
        aabb = np.zeros((n, 4), np.int32)
        fr = np.zeros((n, ROWS, COLS), np.int32) if frames else None
        _check(lib().qpso_solve_eval_poses(self.handle, n, _p(poses), _p(fit), _p(aabb),
                                           _p(fr) if frames else None), self.ctx)
        return fit, aabb, fr


class Server:
    def __init__(self, ctx, budget_bytes):
        h = C.c_void_p()
        _check(lib().qpso_server_create(ctx.handle, int(budget_bytes), C.byref(h)), ctx)
        self.handle, self.ctx = h, ctx

    def set_slots(self, max_resident_particles):
        _check(lib().qpso_server_set_slots(self.handle, int(max_resident_particles)))

    def submit(self, mesh_id, cluster, params):
        t = C.c_uint64()
        _check(lib().qpso_server_submit(self.handle, mesh_id, C.byref(cluster), C.byref(params), C.byref(t)))
        return t.value

    def wait(self, ticket):
        r = Result()
        _check(lib().qpso_server_wait(self.handle, ticket, C.byref(r)))
        return r

    def poll(self, ticket):
        r = Result()
        rc = lib().qpso_server_poll(self.handle, ticket, C.byref(r))
        if rc == E_PENDING:
            return None
        _check(rc)
        return r

    def wait_status(self, ticket):
        """(status, result) without raising on a per-request error."""
        r = Result()
        rc = lib().qpso_server_wait(self.handle, ticket, C.byref(r))
        return rc, r

    def cancel(self, ticket):
        _check(lib().qpso_server_cancel(self.handle, ticket))

    def set_timeout(self, timeout_ms):
        _check(lib().qpso_server_set_timeout(self.handle, float(timeout_ms)))

    def stats(self):
        s = ServerStats()
        _check(lib().qpso_server_stats_get(self.handle, C.byref(s)))
        return s

    def close(self):
        if self.handle:
            lib().qpso_server_destroy(self.handle)
            self.handle = None


# ---------------------------------------------------------------- scenes
def mesh_icosphere(subdiv, ax, ay, az):
    F = lib().qpso_mesh_icosphere(subdiv, ax, ay, az, None)
    out = np.zeros(9 * F, np.float32)
    lib().qpso_mesh_icosphere(subdiv, ax, ay, az, _p(out))
    return out


def mesh_uv_ellipsoid(slices, stacks, ax, ay, az, n_boxes=0, box_min=10.0, box_max=20.0, seed=0):
    F = lib().qpso_mesh_uv_ellipsoid(slices, stacks, ax, ay, az, n_boxes, box_min, box_max, seed, None)
    out = np.zeros(9 * F, np.float32)
    lib().qpso_mesh_uv_ellipsoid(slices, stacks, ax, ay, az, n_boxes, box_min, box_max, seed, _p(out))
    return out


def random_quat(seed):
    q = np.zeros(4, np.float32)
    lib().qpso_random_quat(seed, _p(q))
    return q


def scene_render(objects, cam, plane_z=1100.0, sigma=2.0, hole_frac=0.05, seed=0):
    """objects: list of (mesh_soa_mm, pose7 with translation in mm). Returns depth, labels."""
    arr = (SceneObject * len(objects))()
    keep = []
    for i, (mesh, pose) in enumerate(objects):
        m = np.ascontiguousarray(mesh, np.float32)
        keep.append(m)
        arr[i].mesh_soa = m.ctypes.data
        arr[i].n_faces = m.size // 9
        for d in range(7):
            arr[i].pose[d] = float(pose[d])
    depth = np.zeros((ROWS, COLS), np.uint16)
    labels = np.zeros((ROWS, COLS), np.int32)
    rc = lib().qpso_scene_render(arr, len(objects), *[float(x) for x in cam], plane_z, sigma,
                                 hole_frac, seed, _p(depth), _p(labels))
    _check(rc)
    return depth, labels


def scene_cluster(depth, labels, obj, cam):
    roi = ClusterRoi()
    cd = np.zeros((ROWS, COLS), np.uint16)
    rc = lib().qpso_scene_cluster(_p(np.ascontiguousarray(depth, np.uint16)),
                                  _p(np.ascontiguousarray(labels, np.int32)), obj,
                                  *[float(x) for x in cam], C.byref(roi), _p(cd))
    if rc != 0:
        return None
    return {"rect": list(roi.rect), "centroid_mm": list(roi.centroid_mm),
            "num_points": int(roi.num_points), "depth": cd}
