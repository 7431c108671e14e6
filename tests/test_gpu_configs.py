"""GPU against the oracle on the other workloads at their sizes: the
reference's own configuration (every trace), the batched 16-object frame
(every solve, all 60 iterations), the 2,048-particle swarm on the 100k-face
mesh (its first 10 iterations in full), a frame of the served mix, and the
stress load (64 such swarms mixed with 64 small solves) through the server."""
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


def test_reference_configuration_matches_oracle():
    # qpso_params_default (N 1024, I 20, w 0.3, c1 = c2 = 1, ACCV intrinsics,
    # main.h:30-33, 84, 99, 103-105) with the 6 x 512 render grid's cap of
    # 3072 faces (main.h:162-163) on a 10,336-face mesh: every trace
    scene, params = W.reference_config(0)
    assert scene.objects[0][0].size // 9 > 3072
    print("reference configuration: 1024 particles, 20 iterations, cap 3072", flush=True)
    with qpso.Context(0) as ctx:
        mid = ctx.upload_mesh(scene.mesh_m())
        try:
            with ctx.solve(mid, scene.cluster(), params, trace=True) as s:
                res = s.run()
                tr = s.trace(frames=True)
        finally:
            ctx.release_mesh(mid)
    op = W.oracle_params(O, scene, params)
    rand = O.rand_table(params.seed, params.n_particles, params.n_iters)
    o = O.solve(scene.mesh_m(), scene.clusters[0]["depth"], op, rand, n_threads=16, want_frames=True)
    assert res.status == 0
    assert_same_f32(tr["fit_trace"], o["fit_trace"], "fitness trace")
    assert_same_i32(tr["argmin_trace"], o["argmin_trace"], "argmin trace")
    assert_same_f32(tr["gbest_trace"], o["gbest_trace"], "gbest trace")
    for k in ("pos", "vel", "pbest", "pbest_fit"):
        assert_same_f32(tr[k], o[k], k)
    assert_same_i32(tr["frames"], o["frames"], "carried depth buffers")
    assert_same_f32(np.array(res.pose[:]), o["best_pose"], "best pose")
    assert res.raster_px == o["raster_px"] and res.visit_px == o["visit_px"]


def test_large_swarm_matches_oracle():
    # 2,048 particles on a 100k-face mesh (more vertices than the LDS cache
    # holds: drawn window by window), the first 10 iterations in full
    scene, params = W.config4(0, n_iters=10)
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


def test_stress_mix_served_equals_plain():
    # the stress load: 64 solves of the 100k-face scene (2,048 particles, 200
    # iterations, each its own random table) and 64 small solves (1k faces,
    # 64 particles, 20..200 iterations), all submitted at once to a server
    # with a 32 GiB budget and slots for every particle: all are admitted
    # together, and each result is its plain solve's, bit for bit
    big, small, jobs = W.config4_mix(0)
    scenes = [big, small]
    print(f"stress mix: {len(jobs)} requests", flush=True)
    with qpso.Context(0) as ctx:
        mids = [ctx.upload_mesh(sc.mesh_m()) for sc in scenes]
        try:
            sv = qpso.Server(ctx, 32 << 30)
            try:
                sv.set_slots(sum(p.n_particles for _, p in jobs))
                tickets = [sv.submit(mids[k], scenes[k].cluster(), p) for k, p in jobs]
                served = [sv.wait(t) for t in tickets]
                st = sv.stats()
            finally:
                sv.close()
            assert st.completed == len(jobs) and st.rejected == 0
            assert st.mean_active > 64  # the whole load ran concurrently for most steps
            for (k, p), r in zip(jobs, served):
                with ctx.solve(mids[k], scenes[k].cluster(), p) as s:
                    plain = s.run()
                assert r.status == 0 and plain.status == 0
                assert_same_f32(np.array(r.pose[:]), np.array(plain.pose[:]), f"seed {p.seed} pose")
                assert_same_f32(np.float32(r.fit), np.float32(plain.fit), f"seed {p.seed} fit")
                assert r.raster_px == plain.raster_px and r.visit_px == plain.visit_px
        finally:
            for m in mids:
                ctx.release_mesh(m)
