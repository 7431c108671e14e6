"""Seeded synthetic workloads for the pose-search hot path.

Scenes follow the benchmark definitions: an object mesh (mm) at a random
ground-truth pose, the depth frame from the independent ray caster over a
background plane with Gaussian noise and zero holes, and the cluster ROI of
each object (rect, centroid, pixel count, masked depth).  Meshes are handed
to the solver in metres, like the reference's loadOBJ3 output.

  config0  icosphere (1,280 faces, 100x70x50 mm), N 32, I 30, one cluster
  config1  UV ellipsoid 80x65 (10,240 faces) + 8 boxes, ROI 220x220, N 256, I 100
  config2  16 objects in one frame, faces log-uniform [500, 20000],
           N in {64,128,256,512}, I 60, ROI 40^2..250^2
  config3  serving stream: 8 resident meshes (1k..50k faces), frames of 4..16
           clusters, N 128, I 60, Poisson arrivals
  config4  100k-face mesh, ROI up to 320^2, N 2048, I 200, sigma 4, 15% holes;
           config4_mix: 64 of those solves mixed with 64 1k-face/N 64 solves
  reference_config  the reference's own defaults (N 1024, I 20, ACCV camera,
           3072-face cap) on the config1 object
"""
from __future__ import annotations

import math

import numpy as np

import qpso

CAM = (525.0, 525.0, 319.5, 239.5)


def uv_dims(faces):
    """slices, stacks of a UV ellipsoid with about `faces` faces (2*sl*(st-1))."""
    sl = max(6, int(round(math.sqrt(faces))))
    st = max(3, int(round(faces / (2 * sl))) + 1)
    return sl, st


def gt_pose(rng, xy=100.0, z=(600.0, 1000.0), seed=0):
    q = qpso.random_quat(seed)
    t = [rng.uniform(-xy, xy), rng.uniform(-xy, xy), rng.uniform(*z)]
    return np.array(t + list(q), np.float32)


class Scene:
    """One frame: objects (mesh mm, pose mm), depth, labels and cluster ROIs."""

    def __init__(self, objects, depth, labels, cam=CAM):
        self.objects, self.depth, self.labels, self.cam = objects, depth, labels, cam
        self.clusters = [qpso.scene_cluster(depth, labels, i, cam) for i in range(len(objects))]

    def cluster(self, i=0):
        c = self.clusters[i]
        return qpso.make_cluster(c["rect"], c["centroid_mm"], c["num_points"], c["depth"])

    def mesh_m(self, i=0):
        return (self.objects[i][0] * np.float32(1e-3)).astype(np.float32)


def config0(seed=0):
    rng = np.random.default_rng(seed)
    mesh = qpso.mesh_icosphere(3, 50.0, 35.0, 25.0)
    pose = gt_pose(rng, seed=seed)
    depth, labels = qpso.scene_render([(mesh, pose)], CAM, 1100.0, 2.0, 0.05, seed + 1)
    scene = Scene([(mesh, pose)], depth, labels)
    params = qpso.make_params(32, 30, w=0.7, c1=1.5, c2=1.5, cam=CAM, seed=0)
    return scene, params


def config1(seed=0, n_particles=256, n_iters=100):
    rng = np.random.default_rng(1000 + seed)
    # semi-axes chosen so that at this pose the cluster rect is 220x220 px
    mesh = qpso.mesh_uv_ellipsoid(80, 65, 164.0, 168.0, 162.0, 8, 10.0, 20.0, seed + 7)
    pose = gt_pose(rng, xy=60.0, z=(780.0, 820.0), seed=1000 + seed)
    depth, labels = qpso.scene_render([(mesh, pose)], CAM, 1100.0, 2.0, 0.05, seed + 2)
    scene = Scene([(mesh, pose)], depth, labels)
    params = qpso.make_params(n_particles, n_iters, w=0.7, c1=1.5, c2=1.5, cam=CAM, seed=0)
    return scene, params


def _place(rng, n, radii_px, cam, zr):
    """Non-overlapping image-plane placement; returns translations (mm)."""
    placed, out = [], []
    for i in range(n):
        r = radii_px[i]
        for _ in range(2000):
            u, v = rng.uniform(r + 2, 640 - r - 2), rng.uniform(r + 2, 480 - r - 2)
            if all((u - pu) ** 2 + (v - pv) ** 2 > (r + pr + 3) ** 2 for pu, pv, pr in placed):
                break
        else:
            return None
        placed.append((u, v, r))
        z = rng.uniform(*zr)
        out.append(((u - cam[2]) * z / cam[0], (v - cam[3]) * z / cam[1], z))
    return out


def config2(seed=0, n_objects=16, n_iters=60):
    rng = np.random.default_rng(2000 + seed)
    for attempt in range(50):
        faces = np.exp(rng.uniform(math.log(500), math.log(20000), n_objects))
        roi = np.exp(rng.uniform(math.log(40), math.log(250), n_objects))
        roi = np.sort(roi)[::-1] * (0.97 ** attempt)
        zs = rng.uniform(600, 1200, n_objects)
        trans = _place(rng, n_objects, roi / 2.0 + 4, CAM, (600, 1200))
        if trans is not None:
            break
    objects = []
    for i in range(n_objects):
        z = trans[i][2]
        a = roi[i] * z / CAM[0] / 2.0  # semi-axis (mm) for the wanted extent
        sl, st = uv_dims(max(12, faces[i] - 24))
        mesh = qpso.mesh_uv_ellipsoid(sl, st, a, a * 0.7, a * 0.5, 2, 0.1 * a, 0.2 * a, seed * 100 + i)
        q = qpso.random_quat(2000 + seed * 100 + i)
        objects.append((mesh, np.array(list(trans[i]) + list(q), np.float32)))
    depth, labels = qpso.scene_render(objects, CAM, 1500.0, 2.0, 0.05, seed + 3)
    scene = Scene(objects, depth, labels)
    sizes = rng.choice([64, 128, 256, 512], n_objects)
    params = [qpso.make_params(int(sizes[i]), n_iters, w=0.7, c1=1.5, c2=1.5, cam=CAM, seed=0)
              for i in range(n_objects)]
    return scene, params


def config3_meshes(seed=0):
    """8 resident meshes, 1k..50k faces (mm) and their semi-axes."""
    rng = np.random.default_rng(3000 + seed)
    faces = np.exp(np.linspace(math.log(1000), math.log(50000), 8))
    out = []
    for i in range(8):
        a = rng.uniform(35, 90)
        sl, st = uv_dims(faces[i] - 48)
        out.append((qpso.mesh_uv_ellipsoid(sl, st, a, a * 0.7, a * 0.5, 4, 8.0, 16.0, seed * 10 + i), a))
    return out


def config3_frame(meshes, rng, frame_seed):
    """One 30 fps frame: 4..16 clusters drawn from the resident meshes."""
    k = int(rng.integers(4, 17))
    ids = rng.integers(0, len(meshes), k)
    radii = [meshes[i][1] * 525.0 / 900.0 + 4 for i in ids]
    trans = _place(rng, k, radii, CAM, (600, 1200))
    while trans is None:
        k -= 1
        ids, radii = ids[:k], radii[:k]
        trans = _place(rng, k, radii, CAM, (600, 1200))
    objects = [(meshes[i][0], np.array(list(trans[j]) + list(qpso.random_quat(frame_seed * 31 + j)), np.float32))
               for j, i in enumerate(ids)]
    depth, labels = qpso.scene_render(objects, CAM, 1500.0, 2.0, 0.05, frame_seed)
    return Scene(objects, depth, labels), list(ids)


def config4(seed=0, n_particles=2048, n_iters=200):
    rng = np.random.default_rng(4000 + seed)
    mesh = qpso.mesh_uv_ellipsoid(250, 201, 150.0, 110.0, 80.0, 0, 10.0, 20.0, seed)
    pose = gt_pose(rng, xy=40.0, z=(620.0, 680.0), seed=4000 + seed)
    depth, labels = qpso.scene_render([(mesh, pose)], CAM, 1100.0, 4.0, 0.15, seed + 4)
    scene = Scene([(mesh, pose)], depth, labels)
    params = qpso.make_params(n_particles, n_iters, w=0.7, c1=1.5, c2=1.5, cam=CAM, seed=0)
    return scene, params


def config4_mix(seed=0, n_stress=64, n_small=64, stress_iters=200):
    """The stress load: `n_stress` solves of the 100k-face scene (2048
    particles, `stress_iters` iterations, each with its own random-table seed)
    mixed with `n_small` solves of a 1k-face object (64 particles, 20..200
    iterations).  Returns (stress scene, small scene, [(scene index, params)])
    in submission order: the kinds interleaved."""
    big, _ = config4(seed)
    rng = np.random.default_rng(4400 + seed)
    sl, st = uv_dims(1000 - 24)
    mesh = qpso.mesh_uv_ellipsoid(sl, st, 40.0, 28.0, 20.0, 2, 4.0, 8.0, seed + 44)
    pose = gt_pose(rng, xy=80.0, z=(600.0, 900.0), seed=4401 + seed)
    depth, labels = qpso.scene_render([(mesh, pose)], CAM, 1100.0, 4.0, 0.15, seed + 45)
    small = Scene([(mesh, pose)], depth, labels)
    jobs = []
    for k in range(max(n_stress, n_small)):
        if k < n_stress:
            jobs.append((0, qpso.make_params(2048, stress_iters, w=0.7, c1=1.5, c2=1.5, cam=CAM, seed=1 + k)))
        if k < n_small:
            iters = int(rng.integers(20, 201))
            jobs.append((1, qpso.make_params(64, iters, w=0.7, c1=1.5, c2=1.5, cam=CAM, seed=1000 + k)))
    return big, small, jobs


def edge_config(seed=0, n_particles=64, n_iters=12):
    """The 1,280-face object cut by the left frame edge (its centre about 25 px
    from it), for the off-frame modes: many particles' boxes cross x = 0."""
    rng = np.random.default_rng(5000 + seed)
    mesh = qpso.mesh_icosphere(3, 50.0, 35.0, 25.0)
    z = rng.uniform(650.0, 750.0)
    pose = np.array([(25.0 - CAM[2]) * z / CAM[0], rng.uniform(-40.0, 40.0), z, *qpso.random_quat(5000 + seed)],
                    np.float32)
    depth, labels = qpso.scene_render([(mesh, pose)], CAM, 1100.0, 2.0, 0.05, seed + 5)
    scene = Scene([(mesh, pose)], depth, labels)
    params = qpso.make_params(n_particles, n_iters, w=0.7, c1=1.5, c2=1.5, cam=CAM, seed=0, offframe=1)
    return scene, params


# FX, FY, CX, CY of the reference (main.h:30-33)
ACCV_CAM = (572.41140, 573.57043, 325.26110, 242.04899)


def reference_config(seed=0):
    """The reference's own run: qpso_params_default (N 1024, I 20, w 0.3,
    c1 = c2 = 1, the ACCV intrinsics) with its 6 x 512-thread render grid,
    i.e. faces 0..3071 only (face_cap 3072), on the config1 object seen
    through the ACCV camera."""
    rng = np.random.default_rng(1000 + seed)
    mesh = qpso.mesh_uv_ellipsoid(80, 65, 164.0, 168.0, 162.0, 8, 10.0, 20.0, seed + 7)
    pose = gt_pose(rng, xy=60.0, z=(780.0, 820.0), seed=1000 + seed)
    depth, labels = qpso.scene_render([(mesh, pose)], ACCV_CAM, 1100.0, 2.0, 0.05, seed + 2)
    scene = Scene([(mesh, pose)], depth, labels, cam=ACCV_CAM)
    params = qpso.default_params()
    params.face_cap = 3072
    return scene, params


def oracle_params(oracle_mod, scene, params, i=0):
    """The oracle's parameter struct for the same solve (bounds as main.cu:1303-1331)."""
    c = scene.clusters[i]
    cen = [np.float32(x) * np.float32(0.001) for x in c["centroid_mm"]]
    tb, tbz = np.float32(params.t_bound), np.float32(params.t_bound_z)
    lo = [cen[0] - tb, cen[1] - tb, cen[2] - tbz]
    hi = [cen[0] + tb, cen[1] + tb, cen[2] + tbz]
    cam = (params.cam.fx, params.cam.fy, params.cam.cx, params.cam.cy)
    return oracle_mod.make_params(params.n_particles, params.n_iters, params.w, params.c1, params.c2,
                                  lo, hi, c["rect"], c["num_points"], cam, params.face_cap, params.offframe)
