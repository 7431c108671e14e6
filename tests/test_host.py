"""CPU tests of the host side: the C ABI library loads and exports every
symbol the headers declare, the reference defaults, the synthetic scene and
cluster-ROI provider, and the file formats (OBJ, .dpt, bestpose.csv)."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import qpso
import workloads as W

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(qpso_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    L = C.CDLL(qpso.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert names == set(qpso.EXPORTS), names ^ set(qpso.EXPORTS)


def test_params_default_are_reference_constants():
    p = qpso.default_params()
    assert (p.n_particles, p.n_iters, p.w, p.c1, p.c2) == (1024, 20, 0.3, 1.0, 1.0)
    assert abs(p.t_bound - 0.3) < 1e-7 and abs(p.t_bound_z - 0.3) < 1e-7 and p.seed == 0
    assert (p.cam.fx, p.cam.fy, p.cam.cx, p.cam.cy) == (572.41140, 573.57043, 325.26110, 242.04899)


def test_status_strings():
    L = qpso.lib()
    for code in (0, -1, -2, -3, -4, -5, -6, -7, -8, -9):
        assert L.qpso_status_str(code) != b"unknown status"
    assert L.qpso_status_str(-99) == b"unknown status"


def test_create_without_gpu_fails_loudly():
    import subprocess
    probe = subprocess.run(["python3", "-c", "import torch; print(torch.cuda.device_count())"],
                           capture_output=True, text=True)
    if probe.returncode == 0 and probe.stdout.strip() not in ("", "0"):
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert qpso.lib().qpso_create(0, C.byref(h)) == qpso.E_HIP and not h.value


def test_solve_bytes_grows_with_swarm():
    p = qpso.make_params(128, 60)
    a = qpso.lib().qpso_solve_bytes(C.byref(p), (C.c_int32 * 4)(0, 0, 100, 100))
    p.n_particles = 256
    b = qpso.lib().qpso_solve_bytes(C.byref(p), (C.c_int32 * 4)(0, 0, 100, 100))
    assert 0 < a < b
    # the reference-faithful off-frame mode adds a 640x480 int32 buffer per particle
    p.offframe = 1
    c = qpso.lib().qpso_solve_bytes(C.byref(p), (C.c_int32 * 4)(0, 0, 100, 100))
    assert b + 256 * 640 * 480 * 4 <= c < b + 256 * 640 * 480 * 4 + 4096


def soa_faces(soa):
    F = soa.size // 9
    return soa.reshape(3, 3, F)  # [xyz][slot][face]


@pytest.mark.parametrize("subdiv", [0, 1, 3])
def test_icosphere_closed_and_outward(subdiv):
    m = qpso.mesh_icosphere(subdiv, 50.0, 35.0, 25.0)
    F = m.size // 9
    assert F == 20 * 4 ** subdiv
    V = soa_faces(m).astype(np.float64)
    v0, v1, v2 = V[:, 0].T, V[:, 1].T, V[:, 2].T
    n = np.cross(v1 - v0, v2 - v0)
    c = (v0 + v1 + v2) / 3
    assert np.all(np.einsum("ij,ij->i", n, c) > 0)  # outward winding = front-facing rule
    # closed: every directed edge has its reverse
    e = {}
    for f in range(F):
        for a, b in ((0, 1), (1, 2), (2, 0)):
            e[(tuple(V[:, a, f]), tuple(V[:, b, f]))] = 1
    assert all((b, a) in e for (a, b) in e)


def test_uv_ellipsoid_face_count_and_boxes():
    assert qpso.lib().qpso_mesh_uv_ellipsoid(80, 65, 1, 1, 1, 0, 10, 20, 0, None) == 10240
    m = qpso.mesh_uv_ellipsoid(80, 65, 160.0, 115.0, 85.0, 8, 10.0, 20.0, 7)
    assert m.size // 9 == 10240 + 96
    m2 = qpso.mesh_uv_ellipsoid(80, 65, 160.0, 115.0, 85.0, 8, 10.0, 20.0, 7)
    assert np.array_equal(m, m2)


def test_scene_plane_only_and_noise():
    cam = W.CAM
    depth, labels = qpso.scene_render([], cam, 1100.0, 0.0, 0.0, 1)
    assert np.all(depth == 1100) and np.all(labels == -1)
    depth, _ = qpso.scene_render([], cam, 1100.0, 2.0, 0.05, 1)
    holes = (depth == 0).mean()
    assert 0.04 < holes < 0.06
    d = depth[depth > 0].astype(np.float64)
    assert abs(d.mean() - 1100) < 0.05 and 1.8 < d.std() < 2.2


def test_scene_sphere_depth_and_cluster_roi():
    cam = W.CAM
    mesh = qpso.mesh_icosphere(4, 50.0, 50.0, 50.0)
    pose = np.array([0, 0, 800, 1, 0, 0, 0], np.float32)
    depth, labels = qpso.scene_render([(mesh, pose)], cam, 1100.0, 0.0, 0.0, 0)
    # pixel-centre ray through the image centre hits the sphere front at ~750 mm
    assert abs(int(depth[240, 320]) - 750) <= 1
    c = qpso.scene_cluster(depth, labels, 0, cam)
    r = c["rect"]
    # silhouette radius ~ f * R / sqrt(d^2 - R^2)
    rad = 525 * 50 / np.sqrt(800 ** 2 - 50 ** 2)
    assert abs((r[2] - r[0]) / 2 - rad) < 2 and abs((r[3] - r[1]) / 2 - rad) < 2
    assert c["num_points"] == (labels == 0).sum()
    assert np.all(c["depth"][labels != 0] == 0) and np.all(c["depth"][labels == 0] == depth[labels == 0])
    assert abs(c["centroid_mm"][2] - 765) < 10 and abs(c["centroid_mm"][0]) < 2


def test_workload_configs_build():
    s0, p0 = W.config0(0)
    assert s0.objects[0][0].size // 9 == 1280 and p0.n_particles == 32 and p0.n_iters == 30
    assert s0.clusters[0] is not None
    s1, p1 = W.config1(0)
    assert s1.objects[0][0].size // 9 == 10336 and p1.n_particles == 256 and p1.n_iters == 100


def test_obj_reader_layout(tmp_path):
    path = tmp_path / "quad.obj"
    path.write_text("o first\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\nf 1/1/1 3//2 4\n"
                    "o second\nv 5 5 5\nf 5 5 5\n")
    soa = C.c_void_p()
    nf = C.c_uint32()
    assert qpso.lib().qpso_read_obj(str(path).encode(), C.byref(soa), C.byref(nf)) == 0
    F = nf.value
    a = np.ctypeslib.as_array(C.cast(soa, C.POINTER(C.c_float)), shape=(9 * F,)).copy()
    qpso.lib().qpso_free(soa)
    assert F == 3  # quad fan-split into 2, plus 1 triangle; second object ignored
    V = a.reshape(3, 3, F)
    assert list(V[:, :, 0].T.reshape(-1)) == [0, 0, 0, 1, 0, 0, 1, 1, 0]
    assert list(V[:, :, 1].T.reshape(-1)) == [0, 0, 0, 1, 1, 0, 0, 1, 0]


def test_dpt_round_trip(tmp_path):
    path = str(tmp_path / "d.dpt").encode()
    d = (np.arange(480 * 640) % 65536).astype(np.uint16).reshape(480, 640)
    assert qpso.lib().qpso_write_dpt(path, d.ctypes.data, 480, 640) == 0
    raw = open(path, "rb").read()
    assert np.frombuffer(raw[:8], np.int32).tolist() == [480, 640] and len(raw) == 8 + 2 * 480 * 640
    out, r, c = C.c_void_p(), C.c_int32(), C.c_int32()
    assert qpso.lib().qpso_read_dpt(path, C.byref(out), C.byref(r), C.byref(c)) == 0
    got = np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_uint16)), shape=(480 * 640,)).copy()
    qpso.lib().qpso_free(out)
    assert (r.value, c.value) == (480, 640) and np.array_equal(got.reshape(480, 640), d)


def test_bestpose_csv_format(tmp_path):
    path = str(tmp_path / "x-0-bestpose.csv")
    pose = np.array([0.1, -0.2, 0.8, 1, 0, 0, 0], np.float32)
    assert qpso.lib().qpso_write_bestpose(path.encode(), C.c_float(0.5), pose.ctypes.data) == 0
    assert open(path).read() == "0.500000,0.100000,-0.200000,0.800000,1.000000,0.000000,0.000000,0.000000"


def test_clusters_from_label_map():
    labels = np.zeros((480, 640), np.uint16)
    depth = np.full((480, 640), 1000, np.uint16)
    labels[100:120, 200:230] = 2  # cluster index 1
    depth[110, 210] = 0
    cam = qpso.Camera(525.0, 525.0, 319.5, 239.5)
    rect = (C.c_int32 * 4)()
    cen = (C.c_float * 3)()
    npts = C.c_uint32()
    out = np.zeros((480, 640), np.uint16)
    L = qpso.lib()
    assert L.qpso_clusters_from_labels(labels.ctypes.data, depth.ctypes.data, 1, C.byref(cam), rect, cen,
                                       C.byref(npts), out.ctypes.data) == 0
    assert list(rect) == [200, 100, 230, 120] and npts.value == 600
    assert out.sum() == 599 * 1000 and abs(cen[2] - 1000) < 1e-3
    assert L.qpso_clusters_from_labels(labels.ctypes.data, depth.ctypes.data, 0, C.byref(cam), rect, cen,
                                       C.byref(npts), out.ctypes.data) == qpso.E_NOTFOUND


# ---------------------------------------------------------------- cluster centroid
def centroid_restated(labels_mask, depth, cam):
    """compute3DCentroid (GraphCannySeg.h:2569-2597) with projectPixel2CameraRF
    (GraphCannySeg.h:2510-2517): pixels with depth z != 0 give the float
    vector (u z, v z, z) times K^-1, whose entries are OpenCV's 3x3 cofactor
    inverse of the float camera matrix (double, rounded once per entry) and
    whose rows are summed in double and rounded once; the points are summed
    in float in row-major order and divided by their number in float."""
    fx, fy, cx, cy = (np.float64(np.float32(c)) for c in cam)
    det = fx * fy
    k00, k02 = np.float32(fy / det), np.float32(-(cx * fy) / det)
    k11, k12 = np.float32(fx / det), np.float32(-(fx * cy) / det)
    v, u = np.nonzero(labels_mask & (depth != 0))  # row-major order
    z = depth[v, u].astype(np.int64)
    p0, p1, p2 = (u * z).astype(np.float32), (v * z).astype(np.float32), z.astype(np.float32)
    X = (np.float64(k00) * p0 + np.float64(k02) * p2).astype(np.float32)
    Y = (np.float64(k11) * p1 + np.float64(k12) * p2).astype(np.float32)
    n = np.float32(len(z))
    return [np.cumsum(a, dtype=np.float32)[-1] / n for a in (X, Y, p2)]


def test_centroid_matches_graphcannyseg_restatement():
    rng = np.random.default_rng(4)
    cam = (572.41140, 573.57043, 325.26110, 242.04899)
    L = qpso.lib()
    for trial in range(4):
        labels = np.zeros((480, 640), np.uint16)
        y0, x0 = rng.integers(0, 300), rng.integers(0, 400)
        h, w = rng.integers(20, 170), rng.integers(20, 230)
        labels[y0:y0 + h, x0:x0 + w] = 1
        labels[rng.random((480, 640)) < 0.02] = 0  # ragged mask
        depth = rng.integers(300, 4000, (480, 640)).astype(np.uint16)
        depth[rng.random((480, 640)) < 0.1] = 0  # holes: counted in num_points, not in the centroid
        rect, cen, npts = (C.c_int32 * 4)(), (C.c_float * 3)(), C.c_uint32()
        camc = qpso.Camera(*cam)
        assert L.qpso_clusters_from_labels(labels.ctypes.data, depth.ctypes.data, 0, C.byref(camc), rect, cen,
                                           C.byref(npts), None) == 0
        want = centroid_restated(labels == 1, depth, cam)
        assert np.array_equal(np.array(cen[:], np.float32).view(np.uint32), np.array(want, np.float32).view(np.uint32))
        assert npts.value == int((labels == 1).sum())
        # the synthetic scenes' provider is the same computation
        c = qpso.scene_cluster(depth, np.where(labels == 1, 0, -1).astype(np.int32), 0, cam)
        assert np.array_equal(np.array(c["centroid_mm"], np.float32).view(np.uint32),
                              np.array(want, np.float32).view(np.uint32))
        # and close to the exact mean back-projection
        v, u = np.nonzero((labels == 1) & (depth != 0))
        zz = depth[v, u].astype(np.float64)
        exact = [np.mean((u - cam[2]) * zz / cam[0]), np.mean((v - cam[3]) * zz / cam[1]), zz.mean()]
        assert np.allclose(want, exact, rtol=1e-4, atol=0.05)


# ---------------------------------------------------------------- CPU visualisation render
def render_cpu_restated(pose, soa, cam):
    """RenderCPU + CPUfixedPointRender (main.h:1836-1991, 1993-2010): host
    float arithmetic without contraction, iround(v) = (int)(v + 0.5f), depth
    Z1 + cx2 z2 + cx3 z3 in float, kept where strictly below the buffer
    (60000.0f).  Every edge value is evaluated directly at its pixel,
    C + DX (16 y) - DY (16 x) mod 2^32, instead of the reference's
    incremental walk; off-frame pixels are clipped."""
    f32 = np.float32
    F = soa.size // 9
    buf = np.full((480, 640), f32(60000.0), np.float32)
    t = [f32(x) for x in pose]
    q0, q1, q2, q3 = t[3:]
    for i in range(F):
        U, V, Z = [], [], []
        for j in range(3):
            vx, vy, vz = f32(soa[i + j * F]), f32(soa[3 * F + i + j * F]), f32(soa[6 * F + i + j * F])
            c = [q2 * vz - q3 * vy, -(q1 * vz - q3 * vx), q1 * vy - q2 * vx]
            tt = [f32(2.0) * c[0], f32(2.0) * c[1], f32(2.0) * c[2]]
            k = [q2 * tt[2] - q3 * tt[1], -(q1 * tt[2] - q3 * tt[0]), q1 * tt[1] - q2 * tt[0]]
            px = vx + q0 * tt[0] + k[0] + t[0]
            py = vy + q0 * tt[1] + k[1] + t[1]
            pz = vz + q0 * tt[2] + k[2] + t[2]
            r = f32(1.0) / pz
            U.append(f32(cam[2] + cam[0] * float(px * r)))
            V.append(f32(cam[3] + cam[1] * float(py * r)))
            Z.append(f32(1000.0) * pz)
        X = [int(f32(f32(16.0) * u) + f32(0.5)) for u in U]  # C truncation toward zero
        Y = [int(f32(f32(16.0) * v) + f32(0.5)) for v in V]
        area = (X[0] - X[1]) * (Y[2] - Y[0]) - (X[2] - X[0]) * (Y[0] - Y[1])
        if area <= 0:
            continue
        inv = f32(1.0) / f32(area)
        minx, maxx = (min(X) + 15) >> 4, (max(X) + 15) >> 4
        miny, maxy = (min(Y) + 15) >> 4, (max(Y) + 15) >> 4
        if maxx - minx <= 0 or maxy - miny <= 0:
            continue
        yy, xx = np.mgrid[miny:maxy, minx:maxx].astype(np.int64)

        def edge(a, b):
            dx, dy = X[a] - X[b], Y[a] - Y[b]
            c0 = dy * X[a] - dx * Y[a] + (1 if (dy < 0 or (dy == 0 and dx > 0)) else 0)
            e = (c0 + dx * 16 * yy - dy * 16 * xx) & 0xFFFFFFFF
            return np.where(e >= 2 ** 31, e - 2 ** 32, e)

        cx2 = ((edge(1, 2) + 15) >> 8).astype(np.float32)
        cx3 = ((edge(2, 0) + 15) >> 8).astype(np.float32)
        z2, z3 = (Z[1] - Z[0]) * inv, (Z[2] - Z[0]) * inv
        d = (Z[0] + cx2 * z2) + cx3 * z3
        ok = (xx >= 0) & (xx < 640) & (yy >= 0) & (yy < 480)
        for y, x, dv in zip(yy[ok], xx[ok], d[ok]):
            if dv < buf[y, x]:
                buf[y, x] = dv
    return buf


@pytest.mark.parametrize("case", range(4))
def test_render_cpu_matches_restatement(tmp_path, case):
    rng = np.random.default_rng(50 + case)
    soa = (qpso.mesh_icosphere(2, 60.0, 45.0, 35.0) * np.float32(1e-3)).astype(np.float32)
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    q = q if q[0] >= 0 else -q
    # the last case straddles the frame's left edge
    z = rng.uniform(0.5, 0.9)
    tx = -325.26 / 572.41 * z if case == 3 else rng.uniform(-0.1, 0.1)
    pose = np.array([tx, rng.uniform(-0.1, 0.1), z, *q], np.float32)
    cam = qpso.Camera(572.41140, 573.57043, 325.26110, 242.04899)
    buf = np.full((480, 640), np.float32(60000.0), np.float32)
    L = qpso.lib()
    assert L.qpso_render_cpu(pose.ctypes.data, soa.ctypes.data, soa.size // 9, C.byref(cam), buf.ctypes.data) == 0
    want = render_cpu_restated(pose, soa, (cam.fx, cam.fy, cam.cx, cam.cy))
    assert (want < 60000).sum() > 200
    assert np.array_equal(buf.view(np.uint32), want.view(np.uint32))
    # saveBestParticleRender2MatlabCSV (main.h:1993-2010): ostream default
    # float format, ", " between columns, one row per line
    path = tmp_path / "best.csv"
    assert L.qpso_render_cpu_csv(str(path).encode(), pose.ctypes.data, soa.ctypes.data, soa.size // 9,
                                 C.byref(cam)) == 0
    lines = path.read_text().split("\n")
    assert len(lines) == 481 and lines[-1] == ""
    for r in range(0, 480, 7):
        assert lines[r] == ", ".join(f"{float(v):g}" for v in want[r])
