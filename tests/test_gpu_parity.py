"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Stage parity (rendered depth buffers, per-particle AABB and fitness for the
same poses) and end-to-end parity of whole solves (every iteration's fitness,
argmin, gbest, the final swarm state and the carried depth pixels), all
bit-exact.
"""
import os

import numpy as np
import pytest

import oracle as O
import qpso
import workloads as W
from helpers import assert_same_f32, assert_same_i32

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = qpso.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def scene0():
    return W.config0(0)


@pytest.fixture(scope="module")
def scene1():
    return W.config1(0, n_particles=48, n_iters=4)


def sample_poses(scene, n, seed, spread_mm=25.0, wild=0.3):
    """Poses (7 x n, metres) around the ground truth plus a share drawn like InitQPSO."""
    rng = np.random.default_rng(seed)
    gt = scene.objects[0][1]
    out = np.zeros((7, n), np.float32)
    c = scene.clusters[0]["centroid_mm"]
    for i in range(n):
        if rng.random() < wild:
            t = np.array(c) * 1e-3 + rng.uniform(-0.3, 0.3, 3)
            q = rng.normal(size=4)
        else:
            t = gt[:3] * 1e-3 + rng.normal(0, spread_mm * 1e-3, 3)
            q = gt[3:] + rng.normal(0, 0.15, 4)
        q = q / np.linalg.norm(q)
        if q[0] < 0:
            q = -q
        out[:3, i], out[3:, i] = t, q
    return out


def oracle_eval(scene, params, poses):
    op = W.oracle_params(O, scene, params)
    cam = (params.cam.fx, params.cam.fy, params.cam.cx, params.cam.cy)
    c = scene.clusters[0]
    fits, aabbs, frames = [], [], []
    for i in range(poses.shape[1]):
        fr, info = O.render(scene.mesh_m(), poses[:, i], cam, c["rect"], face_cap=params.face_cap)
        frames.append(fr.copy())
        aabbs.append(list(info.aabb))
        f, _ = O.fitness(info, c["rect"], c["num_points"], fr, c["depth"])
        fits.append(f)
    return np.array(fits, np.float32), np.array(aabbs, np.int32), np.array(frames)


def check_eval(ctx, scene, params, poses, frames=True):
    """Frames requested: the kernel keeps exact depth over its whole tile and
    dumps it.  Without: exact depth only inside the cluster rect and the two
    unvisited lines, a rendered-mask elsewhere (the solve path)."""
    mid = ctx.upload_mesh(scene.mesh_m())
    try:
        with ctx.solve(mid, scene.cluster(), params) as s:
            fit, aabb, fr = s.eval_poses(poses, frames=frames)
    finally:
        ctx.release_mesh(mid)
    ofit, oaabb, oframes = oracle_eval(scene, params, poses)
    assert_same_i32(aabb, oaabb, "aabb")
    if frames:
        assert_same_i32(fr, oframes, "rendered depth buffers")
    assert_same_f32(fit, ofit, "fitness")
    return fit


def test_eval_poses_config0(ctx, scene0):
    scene, params = scene0
    poses = sample_poses(scene, params.n_particles, seed=1)
    fit = check_eval(ctx, scene, params, poses)
    assert np.isfinite(fit).sum() > params.n_particles // 2


def test_eval_poses_config1_bands(ctx, scene1):
    # 10k faces, ROI ~220^2: the union box exceeds one tile, so rows are banded
    scene, params = scene1
    poses = sample_poses(scene, params.n_particles, seed=2)
    check_eval(ctx, scene, params, poses)


@pytest.mark.parametrize("wild", [0.0, 1.0])
def test_eval_poses_mask_path(ctx, scene0, scene1, wild):
    # far-off poses render mostly outside the cluster rect: mask, lines, bands
    for k, (scene, params) in enumerate((scene0, scene1)):
        poses = sample_poses(scene, params.n_particles, seed=11 + k, wild=wild)
        check_eval(ctx, scene, params, poses, frames=False)


def test_rect_strips_quirk_width(ctx, scene0):
    # a 121-px cluster rect (a width whose fp32 row index skips and doubles
    # pixels of the union box's first column) over the scene depth inside it
    # (a cluster's depth is zero outside its rect), so nearly every rect pixel
    # is observed: rect rows above and below the rendered tile are counted
    # per row and the x = minX column by its visits
    scene, params = scene0
    c = scene.clusters[0]
    cx, cy = (c["rect"][0] + c["rect"][2]) // 2, (c["rect"][1] + c["rect"][3]) // 2
    rect = [cx - 60, cy - 55, cx + 61, cy + 55]
    depth = np.zeros_like(np.asarray(scene.depth, np.uint16))
    depth[rect[1]:rect[3], rect[0]:rect[2]] = np.asarray(scene.depth, np.uint16)[rect[1]:rect[3], rect[0]:rect[2]]
    num_points = (rect[2] - rect[0]) * (rect[3] - rect[1])
    poses = sample_poses(scene, params.n_particles, seed=41, spread_mm=10.0, wild=0.0)
    mid = ctx.upload_mesh(scene.mesh_m())
    try:
        with ctx.solve(mid, qpso.make_cluster(rect, c["centroid_mm"], num_points, depth), params) as s:
            fit, aabb, _ = s.eval_poses(poses, frames=False)
    finally:
        ctx.release_mesh(mid)
    cam = (params.cam.fx, params.cam.fy, params.cam.cx, params.cam.cy)
    ofit = []
    for i in range(poses.shape[1]):
        fr, info = O.render(scene.mesh_m(), poses[:, i], cam, rect, face_cap=params.face_cap)
        ofit.append(O.fitness(info, rect, num_points, fr, depth)[0])
        a = list(info.aabb)
        assert a[0] >= rect[0] and a[1] <= rect[2] and (a[2] > rect[1] or a[3] < rect[3] - 1)
    assert_same_f32(fit, np.array(ofit, np.float32), "fitness")
    assert np.isfinite(fit).all()


def run_solve_pair(ctx, scene, params, threads=8):
    mid = ctx.upload_mesh(scene.mesh_m())
    try:
        with ctx.solve(mid, scene.cluster(), params, trace=True) as s:
            res = s.run()
            tr = s.trace(frames=True)
    finally:
        ctx.release_mesh(mid)
    op = W.oracle_params(O, scene, params)
    rand = O.rand_table(params.seed, params.n_particles, params.n_iters)
    ores = O.solve(scene.mesh_m(), scene.clusters[0]["depth"], op, rand, n_threads=threads, want_frames=True)
    return res, tr, ores


def check_solve(res, tr, ores):
    assert res.status == 0
    assert_same_f32(tr["fit_trace"], ores["fit_trace"], "fitness trace")
    assert_same_i32(tr["argmin_trace"], ores["argmin_trace"], "argmin trace")
    assert_same_f32(tr["gbest_trace"], ores["gbest_trace"], "gbest trace")
    assert_same_f32(tr["pos"], ores["pos"], "final positions")
    assert_same_f32(tr["vel"], ores["vel"], "final velocities")
    assert_same_f32(tr["pbest"], ores["pbest"], "final personal bests")
    assert_same_f32(tr["pbest_fit"], ores["pbest_fit"], "final personal best fitness")
    assert_same_i32(tr["frames"], ores["frames"], "carried depth pixels")
    assert_same_f32(np.array(res.pose[:]), ores["best_pose"], "best pose")
    assert_same_f32(np.float32(res.fit), ores["best_fit"], "best fit")
    assert res.raster_px == ores["raster_px"]
    assert res.visit_px == ores["visit_px"]


def test_solve_config0_end_to_end(ctx, scene0):
    scene, params = scene0
    check_solve(*run_solve_pair(ctx, scene, params))


def test_solve_config1_short(ctx, scene1):
    scene, params = scene1
    check_solve(*run_solve_pair(ctx, scene, params))


def test_solve_config1_full_size(ctx):
    # the benchmark's solve itself (10,336 faces, 256 particles, 100
    # iterations): every iteration's fitness, argmin and gbest, the final
    # swarm and the best pose, bit for bit
    scene, params = W.config1(0)
    mid = ctx.upload_mesh(scene.mesh_m())
    try:
        with ctx.solve(mid, scene.cluster(), params, trace=True) as s:
            res = s.run()
            tr = s.trace()
    finally:
        ctx.release_mesh(mid)
    op = W.oracle_params(O, scene, params)
    rand = O.rand_table(params.seed, params.n_particles, params.n_iters)
    ores = O.solve(scene.mesh_m(), scene.clusters[0]["depth"], op, rand, n_threads=16)
    assert res.status == 0
    assert_same_f32(tr["fit_trace"], ores["fit_trace"], "fitness trace")
    assert_same_i32(tr["argmin_trace"], ores["argmin_trace"], "argmin trace")
    assert_same_f32(tr["gbest_trace"], ores["gbest_trace"], "gbest trace")
    for k in ("pos", "vel", "pbest", "pbest_fit"):
        assert_same_f32(tr[k], ores[k], k)
    assert_same_f32(np.array(res.pose[:]), ores["best_pose"], "best pose")
    assert res.raster_px == ores["raster_px"] and res.visit_px == ores["visit_px"]


def test_solve_config1_face_cap(ctx, scene1):
    # the reference's 3,072-face cap on the packed vertex cache: only the
    # vertices of the first 3,072 faces are cached and drawn
    scene, params = scene1
    p = qpso.make_params(params.n_particles, params.n_iters, w=0.7, c1=1.5, c2=1.5, cam=W.CAM, seed=3,
                         face_cap=3072)
    check_solve(*run_solve_pair(ctx, scene, p))


@pytest.mark.parametrize("face_cap", [0, 12345])
def test_big_mesh_windows(ctx, face_cap):
    # 16k faces, 8.2k unique vertices: more than the LDS vertex cache holds,
    # so the faces are drawn run by run from projected vertex windows; with a
    # face cap a window can reach past the capped faces' vertices
    mesh = qpso.mesh_uv_ellipsoid(103, 81, 70.0, 50.0, 35.0, 4, 8.0, 16.0, 5)
    rng = np.random.default_rng(77)
    pose = W.gt_pose(rng, xy=60.0, z=(700.0, 900.0), seed=77)
    depth, labels = qpso.scene_render([(mesh, pose)], W.CAM, 1100.0, 2.0, 0.05, 78)
    scene = W.Scene([(mesh, pose)], depth, labels)
    params = qpso.make_params(24, 3, w=0.7, c1=1.5, c2=1.5, cam=W.CAM, seed=0, face_cap=face_cap)
    poses = sample_poses(scene, params.n_particles, seed=21, wild=0.3)
    check_eval(ctx, scene, params, poses, frames=False)
    check_eval(ctx, scene, params, poses)
    check_solve(*run_solve_pair(ctx, scene, params))


def test_big_mesh_windows_far_vertices(ctx):
    # windows of the 16k-face mesh use the packed (16-bit X, Y) vertex cache;
    # poses close to the camera put vertices beyond 2047 pixels while part of
    # the mesh stays in the frame (finite fitness, rendered pixels): those
    # particles project their windows in three words instead
    mesh = qpso.mesh_uv_ellipsoid(103, 81, 70.0, 50.0, 35.0, 4, 8.0, 16.0, 5)
    rng = np.random.default_rng(77)
    pose = W.gt_pose(rng, xy=60.0, z=(700.0, 900.0), seed=77)
    depth, labels = qpso.scene_render([(mesh, pose)], W.CAM, 1100.0, 2.0, 0.05, 78)
    scene = W.Scene([(mesh, pose)], depth, labels)
    params = qpso.make_params(16, 2, w=0.7, c1=1.5, c2=1.5, cam=W.CAM, seed=0)
    poses = sample_poses(scene, params.n_particles, seed=23, wild=0.3)
    gt = scene.objects[0][1]
    r = np.random.default_rng(31)
    for i in range(8):
        q = gt[3:] + r.normal(0, 0.05, 4)
        q = q / np.linalg.norm(q)
        poses[:, 2 * i] = [0.03 + 0.005 * i, r.uniform(-0.02, 0.02), 0.06 + 0.005 * i, *q]
    check_eval(ctx, scene, params, poses)
    fit = check_eval(ctx, scene, params, poses, frames=False)
    near = poses[:, 0::2]
    fo, ao, fr = oracle_eval(scene, params, near)
    far = ao[:, 1] > 2046
    assert (far & np.isfinite(fo) & ((fr < 60000).sum(axis=(1, 2)) > 0)).any()
    assert np.isfinite(fit).sum() > params.n_particles // 2


@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 4])
def test_big_mesh_windows_configs(cfg):
    # the 16k-face mesh (8.2k vertices) in every (threads, tile) variant: one
    # packed window on the 40,448-word tiles (projected once, in pass A), two
    # to five windows on the smaller ones, and one to several bands
    mesh = qpso.mesh_uv_ellipsoid(103, 81, 70.0, 50.0, 35.0, 4, 8.0, 16.0, 5)
    rng = np.random.default_rng(77)
    pose = W.gt_pose(rng, xy=60.0, z=(700.0, 900.0), seed=77)
    depth, labels = qpso.scene_render([(mesh, pose)], W.CAM, 1100.0, 2.0, 0.05, 78)
    scene = W.Scene([(mesh, pose)], depth, labels)
    params = qpso.make_params(16, 2, w=0.7, c1=1.5, c2=1.5, cam=W.CAM, seed=0)
    old = os.environ.get("QPSO_RF_CFG")
    os.environ["QPSO_RF_CFG"] = str(cfg)
    try:
        with qpso.Context(0) as c:
            poses = sample_poses(scene, params.n_particles, seed=40 + cfg, wild=0.4)
            check_eval(c, scene, params, poses, frames=False)
            check_eval(c, scene, params, poses)
    finally:
        if old is None:
            del os.environ["QPSO_RF_CFG"]
        else:
            os.environ["QPSO_RF_CFG"] = old


@pytest.mark.parametrize("mesh_kind", ["cached", "windowed"])
def test_queue_overflow(scene1, mesh_kind):
    # a big-face threshold of 1 pixel sends nearly every face to the wave
    # queue: it overflows into the small-face queue, which overflows into
    # the drawing lane (cached mesh: the 10k-face config1 mesh; windowed: the
    # 16k-face mesh, whose queues are drawn at the end of each band)
    if mesh_kind == "cached":
        scene, params = scene1
        cfg = 0
    else:
        mesh = qpso.mesh_uv_ellipsoid(103, 81, 70.0, 50.0, 35.0, 4, 8.0, 16.0, 5)
        rng = np.random.default_rng(77)
        pose = W.gt_pose(rng, xy=60.0, z=(700.0, 900.0), seed=77)
        depth, labels = qpso.scene_render([(mesh, pose)], W.CAM, 1100.0, 2.0, 0.05, 78)
        scene = W.Scene([(mesh, pose)], depth, labels)
        params = qpso.make_params(16, 2, w=0.7, c1=1.5, c2=1.5, cam=W.CAM, seed=0)
        cfg = 2
    with qpso.Context(0) as c:
        c.set_tuning(cfg, 1)
        poses = sample_poses(scene, params.n_particles, seed=50, spread_mm=10.0, wild=0.2)
        check_eval(c, scene, params, poses, frames=False)
        check_eval(c, scene, params, poses)


def test_big_mesh_without_runs(ctx):
    # the same 16k-face mesh with its faces shuffled: corners of neighbouring
    # faces are far apart in vertex order, so there are no vertex windows and
    # every face projects its own corners
    mesh = qpso.mesh_uv_ellipsoid(103, 81, 70.0, 50.0, 35.0, 4, 8.0, 16.0, 5)
    F = mesh.size // 9
    perm = np.random.default_rng(5).permutation(F)
    mesh = np.ascontiguousarray(mesh.reshape(3, 3, F)[:, :, perm].reshape(-1))
    rng = np.random.default_rng(77)
    pose = W.gt_pose(rng, xy=60.0, z=(700.0, 900.0), seed=77)
    depth, labels = qpso.scene_render([(mesh, pose)], W.CAM, 1100.0, 2.0, 0.05, 78)
    scene = W.Scene([(mesh, pose)], depth, labels)
    params = qpso.make_params(16, 2, w=0.7, c1=1.5, c2=1.5, cam=W.CAM, seed=0)
    poses = sample_poses(scene, params.n_particles, seed=22, wild=0.3)
    check_eval(ctx, scene, params, poses, frames=False)
    check_solve(*run_solve_pair(ctx, scene, params))


def test_packed_cache_far_vertices(ctx, scene1):
    # the 220x220 cluster rect takes the packed vertex cache (16-bit X, Y);
    # a pose close to the camera and off-axis puts vertices beyond 2047
    # pixels (with the union box still under the guard's limit), so that
    # particle projects its faces where needed instead
    scene, params = scene1
    poses = sample_poses(scene, params.n_particles, seed=31)
    rng = np.random.default_rng(31)
    gt = scene.objects[0][1]
    for i in range(8):
        q = gt[3:] + rng.normal(0, 0.05, 4)
        q = q / np.linalg.norm(q)
        poses[:, 5 * i] = [0.2 + 0.01 * i, rng.uniform(-0.02, 0.02), 0.2 + 0.005 * i, *q]
    poses[:, 3] = [0.3, 0.0, 0.2, *gt[3:]]
    fit = check_eval(ctx, scene, params, poses, frames=False)
    far = [i for i in range(0, 40, 5)]
    fo, ao, _ = oracle_eval(scene, params, poses[:, far])
    assert (ao[:, 1] > 2046).any() and np.isfinite(fo[ao[:, 1] > 2046]).any()
    assert np.isfinite(fit).sum() > params.n_particles // 2


@pytest.mark.parametrize("cfg", [1, 2, 3, 4])
def test_tile_configs_agree(scene1, cfg):
    # every (threads, tile) variant of the render+fitness kernel is bit-exact
    scene, params = scene1
    old = os.environ.get("QPSO_RF_CFG")
    os.environ["QPSO_RF_CFG"] = str(cfg)
    try:
        with qpso.Context(0) as c:
            poses = sample_poses(scene, params.n_particles, seed=3 + cfg, wild=0.5)
            check_eval(c, scene, params, poses)
            check_eval(c, scene, params, poses, frames=False)
    finally:
        if old is None:
            del os.environ["QPSO_RF_CFG"]
        else:
            os.environ["QPSO_RF_CFG"] = old


def test_mesh_release_waits_for_its_solves(ctx, scene0):
    """Releasing a mesh that a solve still uses defers the free: the solve
    runs as before, new solves on the mesh are refused, and the mesh goes
    with the last solve."""
    scene, params = scene0
    mid = ctx.upload_mesh(scene.mesh_m())
    s = ctx.solve(mid, scene.cluster(), params)
    try:
        before = s.run()
        ctx.release_mesh(mid)
        after = s.run()
        assert np.float32(before.fit).view(np.uint32) == np.float32(after.fit).view(np.uint32)
        assert np.array_equal(np.array(before.pose[:], np.float32), np.array(after.pose[:], np.float32))
        with pytest.raises(qpso.QpsoError) as e:
            ctx.solve(mid, scene.cluster(), params)
        assert e.value.code == qpso.E_NOTFOUND
    finally:
        s.close()
    with pytest.raises(qpso.QpsoError) as e:
        ctx.release_mesh(mid)
    assert e.value.code == qpso.E_NOTFOUND


@pytest.mark.parametrize("case", ["rect_outside", "rect_empty", "no_points", "no_particles",
                                  "too_many_particles", "no_iters", "too_many_iters", "bad_mesh"])
def test_invalid_inputs_are_refused(ctx, scene0, case):
    """Bad clusters, swarm sizes and mesh ids return a status, with nothing launched."""
    scene, params = scene0
    c = scene.clusters[0]
    rect, npts = list(c["rect"]), c["num_points"]
    p = qpso.make_params(params.n_particles, params.n_iters, cam=W.CAM)
    mid = ctx.upload_mesh(scene.mesh_m())
    try:
        if case == "rect_outside":
            rect = [rect[0], rect[1], 700, rect[3]]
        elif case == "rect_empty":
            rect = [rect[0], rect[1], rect[0], rect[3]]
        elif case == "no_points":
            npts = 0
        elif case == "no_particles":
            p.n_particles = 0
        elif case == "too_many_particles":
            p.n_particles = 65536
        elif case == "no_iters":
            p.n_iters = 0
        elif case == "too_many_iters":
            p.n_iters = (1 << 20) + 1
        cl = qpso.make_cluster(rect, c["centroid_mm"], npts, c["depth"])
        with pytest.raises(qpso.QpsoError) as e:
            ctx.solve(mid + 1000 if case == "bad_mesh" else mid, cl, p)
        assert e.value.code == (qpso.E_NOTFOUND if case == "bad_mesh" else qpso.E_INVAL)
    finally:
        ctx.release_mesh(mid)
