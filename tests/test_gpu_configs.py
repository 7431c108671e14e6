"""GPU against the oracle on the other workloads at their sizes: the batched
16-object frame (every solve, all 60 iterations), the 2,048-particle swarm
on the 100k-face mesh (its first iterations) and a frame of the served mix."""
import numpy as np
import pytest

import oracle as O
import qpso
import workloads as W
from helpers import assert_same_f32, assert_same_i32

pytestmark = pytest.mark.gpu


def oracle_solve(scene, params, i=0):
    op = W.oracle_params(O, scene, params, i)
    rand = O.rand_table(params.seed, params.n_particles, params.n_iters)
    return O.solve(scene.mesh_m(i), scene.clusters[i]["depth"], op, rand, n_threads=16)


def test_batched_frame_matches_oracle():
    # 16 objects, 500..20k faces, 64..512 particles, ROIs 40..250 px, 60
    # iterations, one batched graph: each solve's result is the oracle's
    scene, plist = W.config2(0)
    idx = [i for i in range(len(plist)) if scene.clusters[i] is not None]
    with qpso.Context(0) as ctx:
        meshes = [ctx.upload_mesh(scene.mesh_m(i)) for i in range(len(scene.objects))]
        solves = [ctx.solve(meshes[i], scene.cluster(i), plist[i]) for i in idx]
        try:
            res = ctx.run_batch(solves)
        finally:
            for s in solves:
                s.close()
            for m in meshes:
                ctx.release_mesh(m)
    for k, i in enumerate(idx):
        o = oracle_solve(scene, plist[i], i)
        r = res[k]
        assert r.status == 0, f"solve {i}"
        assert_same_f32(np.array(r.pose[:]), o["best_pose"], f"solve {i} best pose")
        assert_same_f32(np.float32(r.fit), o["best_fit"], f"solve {i} best fit")
        assert r.raster_px == o["raster_px"] and r.visit_px == o["visit_px"], f"solve {i} work counts"


def test_large_swarm_matches_oracle():
    # 2,048 particles on a 100k-face mesh (more vertices than the LDS cache
    # holds: drawn window by window), the first 3 iterations in full
    scene, params = W.config4(0, n_iters=3)
    with qpso.Context(0) as ctx:
        mid = ctx.upload_mesh(scene.mesh_m())
        try:
            with ctx.solve(mid, scene.cluster(), params, trace=True) as s:
                res = s.run()
                tr = s.trace()
        finally:
            ctx.release_mesh(mid)
    o = oracle_solve(scene, params)
    assert res.status == 0
    assert_same_f32(tr["fit_trace"], o["fit_trace"], "fitness trace")
    assert_same_i32(tr["argmin_trace"], o["argmin_trace"], "argmin trace")
    assert_same_f32(tr["gbest_trace"], o["gbest_trace"], "gbest trace")
    for k in ("pos", "vel", "pbest", "pbest_fit"):
        assert_same_f32(tr[k], o[k], k)
    assert res.raster_px == o["raster_px"] and res.visit_px == o["visit_px"]


def test_served_frame_matches_oracle():
    # one frame of the served mix (4..16 clusters over 8 resident meshes of
    # 1k..50k faces, the larger ones drawn window by window), submitted to the
    # continuous-batching server together: each result is the oracle's
    meshes = W.config3_meshes(0)
    metres = [(m * np.float32(1e-3)).astype(np.float32) for m, _ in meshes]
    scene, ids = W.config3_frame(meshes, np.random.default_rng(0), 100)
    params = qpso.make_params(128, 10, w=0.7, c1=1.5, c2=1.5, cam=W.CAM, seed=0)
    jobs = [j for j in range(len(ids)) if scene.clusters[j] is not None]
    with qpso.Context(0) as ctx:
        mids = [ctx.upload_mesh(m) for m in metres]
        sv = qpso.Server(ctx, 8 << 30)
        try:
            tickets = [sv.submit(mids[ids[j]], scene.cluster(j), params) for j in jobs]
            res = [sv.wait(t) for t in tickets]
        finally:
            sv.close()
            for m in mids:
                ctx.release_mesh(m)
    for j, r in zip(jobs, res):
        op = W.oracle_params(O, scene, params, j)
        rand = O.rand_table(params.seed, params.n_particles, params.n_iters)
        o = O.solve(metres[ids[j]], scene.clusters[j]["depth"], op, rand, n_threads=16)
        assert r.status == 0, f"cluster {j}"
        assert_same_f32(np.array(r.pose[:]), o["best_pose"], f"cluster {j} best pose")
        assert_same_f32(np.float32(r.fit), o["best_fit"], f"cluster {j} best fit")
        assert r.raster_px == o["raster_px"] and r.visit_px == o["visit_px"], f"cluster {j} work counts"
