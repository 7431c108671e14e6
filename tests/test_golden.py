"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the oracle): the oracle must keep reproducing them on the CPU, and the
HIP path must reproduce them bit-exactly on the GPU."""
import os

import numpy as np
import pytest

import oracle as O
import qpso
import workloads as W
from helpers import assert_same_f32, assert_same_i32

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = {"config0": lambda: W.config0(0), "config1": lambda: W.config1(0, 8, 3)}


def load(name):
    with np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(CASES))
def test_scene_is_stable(name):
    g = load(name)
    scene, _ = CASES[name]()
    c = scene.clusters[0]
    assert list(g["rect"]) == list(c["rect"]) and int(g["num_points"]) == c["num_points"]
    assert np.array_equal(g["centroid_mm"], np.array(c["centroid_mm"], np.float32))


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_fixture(name):
    g = load(name)
    scene, params = CASES[name]()
    c = scene.clusters[0]
    cam = (params.cam.fx, params.cam.fy, params.cam.cx, params.cam.cy)
    for i in range(g["poses"].shape[1]):
        fr, info = O.render(scene.mesh_m(), g["poses"][:, i], cam, c["rect"])
        f, _ = O.fitness(info, c["rect"], c["num_points"], fr, c["depth"])
        assert_same_f32(f, g["fit"][i], f"fit {i}")
        assert list(info.aabb) == list(g["aabb"][i])
    params.n_particles, params.n_iters = int(g["n_particles"]), int(g["n_iters"])
    op = W.oracle_params(O, scene, params)
    rand = O.rand_table(params.seed, params.n_particles, params.n_iters)
    res = O.solve(scene.mesh_m(), c["depth"], op, rand, n_threads=2)
    assert_same_f32(res["fit_trace"], g["fit_trace"], "fit trace")
    assert_same_f32(res["gbest_trace"], g["gbest_trace"], "gbest trace")
    assert_same_i32(res["argmin_trace"], g["argmin_trace"], "argmin trace")
    assert res["raster_px"] == int(g["raster_px"]) and res["visit_px"] == int(g["visit_px"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_reproduces_fixture(name):
    g = load(name)
    scene, params = CASES[name]()
    params.n_particles, params.n_iters = int(g["n_particles"]), int(g["n_iters"])
    with qpso.Context(0) as ctx:
        mid = ctx.upload_mesh(scene.mesh_m())
        with ctx.solve(mid, scene.cluster(), params, trace=True) as s:
            res = s.run()
            tr = s.trace()
            n = g["poses"].shape[1]
            fit, aabb, _ = s.eval_poses(g["poses"][:, :min(n, params.n_particles)])
        ctx.release_mesh(mid)
    k = fit.size
    assert_same_f32(fit, g["fit"][:k], "stage fitness")
    assert_same_i32(aabb, g["aabb"][:k], "stage aabb")
    assert_same_f32(tr["fit_trace"], g["fit_trace"], "fit trace")
    assert_same_f32(tr["gbest_trace"], g["gbest_trace"], "gbest trace")
    assert_same_i32(tr["argmin_trace"], g["argmin_trace"], "argmin trace")
    assert_same_f32(np.array(res.pose[:]), g["best_pose"], "best pose")
    assert res.raster_px == int(g["raster_px"]) and res.visit_px == int(g["visit_px"])
