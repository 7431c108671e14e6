"""Writes tests/golden/*.npz: oracle outputs on seeded synthetic scenes.

Run from the repository root after building the oracle and libqpso (the
scene generator lives in libqpso; it is host code and needs no GPU):
    python3 tests/golden/make_golden.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(REPO, "q-pso_amd"), os.path.join(REPO, "oracle")]

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
import workloads as W  # noqa: E402


def poses_for(scene, n, seed):
    rng = np.random.default_rng(seed)
    gt = scene.objects[0][1]
    out = np.zeros((7, n), np.float32)
    for i in range(n):
        q = gt[3:] + rng.normal(0, 0.1, 4)
        q /= np.linalg.norm(q)
        if q[0] < 0:
            q = -q
        out[:3, i] = gt[:3] * 1e-3 + rng.normal(0, 0.02, 3)
        out[3:, i] = q
    return out


def main():
    for name, (scene, params), n_eval, solve_np in [("config0", W.config0(0), 12, (16, 6)),
                                                     ("config1", W.config1(0, 8, 3), 6, (8, 3))]:
        c = scene.clusters[0]
        cam = (params.cam.fx, params.cam.fy, params.cam.cx, params.cam.cy)
        poses = poses_for(scene, n_eval, 7)
        fits, aabbs = [], []
        for i in range(n_eval):
            fr, info = O.render(scene.mesh_m(), poses[:, i], cam, c["rect"])
            f, _ = O.fitness(info, c["rect"], c["num_points"], fr, c["depth"])
            fits.append(f)
            aabbs.append(list(info.aabb))
        params.n_particles, params.n_iters = solve_np
        op = W.oracle_params(O, scene, params)
        rand = O.rand_table(params.seed, params.n_particles, params.n_iters)
        res = O.solve(scene.mesh_m(), c["depth"], op, rand, n_threads=4)
        np.savez_compressed(
            os.path.join(HERE, f"{name}.npz"),
            rect=np.array(c["rect"], np.int32), centroid_mm=np.array(c["centroid_mm"], np.float32),
            num_points=np.int64(c["num_points"]), poses=poses, fit=np.array(fits, np.float32),
            aabb=np.array(aabbs, np.int32), n_particles=np.int64(params.n_particles),
            n_iters=np.int64(params.n_iters), fit_trace=res["fit_trace"], gbest_trace=res["gbest_trace"],
            argmin_trace=res["argmin_trace"], best_pose=res["best_pose"], best_fit=res["best_fit"],
            raster_px=np.int64(res["raster_px"]), visit_px=np.int64(res["visit_px"]))
        print(name, "written")


if __name__ == "__main__":
    main()
