"""The reference-faithful off-frame mode (qpso_params.offframe = 1) against
the oracle in the same mode: depth-buffer accesses at idx = y*640 + x
(main.h:1442, main.cu:674, 688), so pixels past the left or right frame edge
land in the neighbouring row.  The scene's object is cut by the left edge, so
many particles' boxes cross it.  Stage parity (the rendered buffers, AABBs
and fitness of given poses) and whole solves (every trace, the final swarm
and the particles' buffers), all bit-exact."""
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


def edge_poses(scene, n, seed):
    """Poses around the ground truth, half of them shifted across the left or right edge."""
    rng = np.random.default_rng(seed)
    gt = scene.objects[0][1]
    out = np.zeros((7, n), np.float32)
    for i in range(n):
        t = gt[:3] * 1e-3 + rng.normal(0, 0.02, 3)
        if i % 4 == 1:
            t[0] += rng.uniform(-0.05, 0.0)  # further out on the left
        elif i % 4 == 3:
            t[0] = (620.0 - W.CAM[2]) * t[2] / W.CAM[0]  # across the right edge
        q = gt[3:] + rng.normal(0, 0.2, 4)
        q /= np.linalg.norm(q)
        if q[0] < 0:
            q = -q
        out[:3, i], out[3:, i] = t, q
    return out


def test_eval_poses_wrap_matches_oracle(ctx):
    scene, params = W.edge_config(0)
    c = scene.clusters[0]
    assert c["rect"][0] == 0  # the cluster touches the left edge
    poses = edge_poses(scene, params.n_particles, 3)
    mid = ctx.upload_mesh(scene.mesh_m())
    try:
        with ctx.solve(mid, scene.cluster(), params) as s:
            fit, aabb, frames = s.eval_poses(poses, frames=True)
    finally:
        ctx.release_mesh(mid)
    cam = (params.cam.fx, params.cam.fy, params.cam.cx, params.cam.cy)
    crossing = 0
    for i in range(poses.shape[1]):
        fr, info = O.render(scene.mesh_m(), poses[:, i], cam, c["rect"], wrap=1)
        assert_same_i32(frames[i], fr, f"pose {i} buffer")
        assert list(aabb[i]) == list(info.aabb), f"pose {i} aabb"
        f, _ = O.fitness(info, c["rect"], c["num_points"], fr.copy(), c["depth"], wrap=1)
        assert_same_f32(fit[i], f, f"pose {i} fitness")
        crossing += info.aabb[0] < 0 or info.aabb[1] > 640
    assert crossing >= poses.shape[1] // 3


def test_solve_wrap_matches_oracle(ctx):
    scene, params = W.edge_config(1)
    mid = ctx.upload_mesh(scene.mesh_m())
    try:
        with ctx.solve(mid, scene.cluster(), params, trace=True) as s:
            res = s.run()
            tr = s.trace(frames=True)
            again = s.run()  # a second run starts from empty buffers again
        clip = qpso.make_params(params.n_particles, params.n_iters, w=params.w, c1=params.c1, c2=params.c2,
                                cam=W.CAM, seed=params.seed, offframe=0)
        with ctx.solve(mid, scene.cluster(), clip) as s:
            res_clip = s.run()
    finally:
        ctx.release_mesh(mid)
    op = W.oracle_params(O, scene, params)
    assert op.wrap == 1
    rand = O.rand_table(params.seed, params.n_particles, params.n_iters)
    o = O.solve(scene.mesh_m(), scene.clusters[0]["depth"], op, rand, n_threads=16, want_frames=True)
    assert res.status == 0
    assert_same_f32(tr["fit_trace"], o["fit_trace"], "fitness trace")
    assert_same_i32(tr["argmin_trace"], o["argmin_trace"], "argmin trace")
    assert_same_f32(tr["gbest_trace"], o["gbest_trace"], "gbest trace")
    for k in ("pos", "vel", "pbest", "pbest_fit"):
        assert_same_f32(tr[k], o[k], k)
    assert_same_i32(tr["frames"], o["frames"], "final buffers")
    assert res.raster_px == o["raster_px"] and res.visit_px == o["visit_px"]
    assert np.float32(again.fit).view(np.uint32) == np.float32(res.fit).view(np.uint32)
    # the mode matters on this scene
    assert res_clip.raster_px != res.raster_px or res_clip.fit != res.fit


def test_wrap_refused_where_unsupported(ctx):
    scene, params = W.edge_config(0)
    mid = ctx.upload_mesh(scene.mesh_m())
    try:
        with ctx.solve(mid, scene.cluster(), params) as s:
            for call in (lambda: ctx.run_batch([s]), lambda: ctx.autotune([s], iters=1)):
                with pytest.raises(qpso.QpsoError) as e:
                    call()
                assert e.value.code == qpso.E_INVAL
        sv = qpso.Server(ctx, 1 << 30)
        try:
            with pytest.raises(qpso.QpsoError) as e:
                sv.submit(mid, scene.cluster(), params)
            assert e.value.code == qpso.E_INVAL
        finally:
            sv.close()
        import ctypes as C
        r = qpso.Result()
        rc = qpso.lib().qpso_estimate_pose_noalloc(ctx.handle, mid, C.byref(scene.cluster()), C.byref(params), C.byref(r))
        assert rc == qpso.E_INVAL
        # the allocating per-cluster call runs it
        assert ctx.estimate_pose(mid, scene.cluster(), params).status == 0
    finally:
        ctx.release_mesh(mid)
