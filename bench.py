#!/usr/bin/env python3
"""Benchmark of the gfx950 Q-PSO pose-search hot path on one MI355X.

Headline workload: one plain solve of a 10,336-face ellipsoid with 8 box
protrusions (ROI about 200x130 px), 256 particles, 100 iterations, inputs
device-resident before timing.  One step = one whole solve (InitQPSO + 100 x
(render+fitness, argmin/pbest/gbest/update)).  value = solves/s.

Also reported on the same JSON line:
  roofline      the render+fitness kernel against HBM: unique bytes per launch
                (mesh, observed ROI, swarm state) / measured average launch
                time; traffic = HBM bytes per launch from rocprof
  roofline_lds  the same kernel against LDS bandwidth: algorithmic LDS bytes
                per launch (4 B per bbox-fill atomicMin + 10 B per fitness
                visit, counted exactly) / launch time
  roofline_valu the same kernel against VALU issue, the resource that binds
                it: VALU wave instructions per launch (rocprof SQ_INSTS_VALU,
                committed summary) / launch time
  cpu_baseline  the CPU oracle (C, OpenMP over particles) on a bounded sample
  batched       16 uneven solves of one frame: batched graph form vs plain,
                and vs the reference's launch structure (one launch per particle)
  served        Poisson arrivals into the continuous-batching server: p50/p99
"""
import argparse
import json
import os
import sys
import threading
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(REPO, "q-pso_amd"), os.path.join(REPO, "oracle")]

import numpy as np  # noqa: E402

import qpso  # noqa: E402
import workloads as W  # noqa: E402

LDS_PEAK_GBS = 256 * 128 * 2.4  # 256 CUs x 128 B/clk x 2.4 GHz = 78,643 GB/s
HBM_PEAK_GBS = 8000.0
# 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 VALU instruction per SIMD every 4
# clocks (16 lanes wide, packed FP32 off)
VALU_PEAK_TIPS = 256 * 4 * 2.4e9 / 4 / 1e12


def pct(v, q):
    v = sorted(v)
    return v[min(len(v) - 1, int(round(q * (len(v) - 1))))]


def from_profiles(key):
    """A per-launch figure of the render+fitness kernel measured by rocprofv3
    --pmc on the headline workload (committed summary), or None."""
    path = os.path.join(REPO, "profiles", "k_render_fitness_10k_faces_256_particles_pmc.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return float(json.load(f)[key])
    except (OSError, ValueError, KeyError):
        return None


def cpu_baseline(scene, params, iters_sample):
    import oracle as O
    O.lib()
    threads = min(16, os.cpu_count() or 1)
    op = W.oracle_params(O, scene, params)
    op.n_iters = iters_sample
    rand = O.rand_table(params.seed, params.n_particles, params.n_iters)
    t = time.perf_counter()
    O.solve(scene.mesh_m(), scene.clusters[0]["depth"], op, rand, n_threads=threads)
    dt = time.perf_counter() - t
    return {"value": (iters_sample / params.n_iters) / dt, "unit": "solves/s", "cores": threads,
            "kind": "port",
            "sample": f"first {iters_sample} of {params.n_iters} iterations of the same solve "
                      f"({params.n_particles} particles), {dt:.1f} s, scaled to whole solves"}


def batched_section(ctx, reps=3):
    scene, plist = W.config2(0)
    meshes = [ctx.upload_mesh(scene.mesh_m(i)) for i in range(len(scene.objects))]
    solves = []
    for i, p in enumerate(plist):
        if scene.clusters[i] is None:
            continue
        solves.append(ctx.solve(meshes[i], scene.cluster(i), p))
    n = len(solves)
    # plain form: every solve tuned on its own, run one after another
    for s in solves:
        ctx.autotune([s], iters=3)
        s.run()
    t = time.perf_counter()
    for _ in range(reps):
        res_p = [s.run() for s in solves]
    t_p = (time.perf_counter() - t) / reps
    # the reference's launch structure with these kernels: one launch per
    # particle on 16 streams, a host read every iteration (one pass)
    t = time.perf_counter()
    res_r = [s.run_reference_launches() for s in solves]
    t_r = time.perf_counter() - t
    # batched form: tuned for the whole frame, one graph launch
    tune = ctx.autotune(solves, iters=3)
    ctx.run_batch(solves)  # capture + warm-up
    t = time.perf_counter()
    for _ in range(reps):
        res_b = ctx.run_batch(solves)
    t_b = (time.perf_counter() - t) / reps
    same = all(np.array_equal(np.array(a.pose[:]), np.array(b.pose[:])) and a.fit == b.fit
               for a, b in zip(res_b + res_r, res_p + res_p))
    for s in solves:
        s.close()
    for m in meshes:
        ctx.release_mesh(m)
    return {"workload": f"one frame, {n} solves, faces {min(len(o[0]) // 9 for o in scene.objects)}.."
                        f"{max(len(o[0]) // 9 for o in scene.objects)}, particles 64..512, 60 iterations",
            "solves": n, "batched_solves_per_s": n / t_b, "plain_solves_per_s": n / t_p,
            "speedup": t_p / t_b, "reference_launch_solves_per_s": n / t_r,
            "speedup_vs_reference_launches": t_r / t_b, "identical_results": bool(same), "autotune": tune[0]}


def served_section(ctx, lams=(200.0, 1000.0, 4000.0), seconds=1.5, seed=0, slots=8192):
    """Poisson arrivals of (frame, cluster) requests into the continuous-batching
    server, one run per arrival rate; latency is submit -> result."""
    meshes = W.config3_meshes(seed)
    mids = [ctx.upload_mesh((m * np.float32(1e-3)).astype(np.float32)) for m, _ in meshes]
    rng = np.random.default_rng(seed)
    pool = []
    for f in range(6):
        scene, ids = W.config3_frame(meshes, rng, 100 + f)
        for j, mi in enumerate(ids):
            if scene.clusters[j] is not None:
                pool.append((mids[mi], scene.cluster(j)))
    params = qpso.make_params(128, 60, w=0.7, c1=1.5, c2=1.5, cam=W.CAM, seed=0)
    # the kernel configuration for this mix: autotune on a step-sized group
    # of its requests (the server launches all resident solves at once), over
    # whole solves, since early iterations (scattered particles) favour other
    # settings than the mix of stages a busy server runs
    sample = [ctx.solve(mid, cl, params) for mid, cl in pool[:min(len(pool), slots // 128)]]
    try:
        tune = ctx.autotune(sample, iters=params.n_iters)
    finally:
        for s in sample:
            s.close()
    runs = []
    for lam in lams:
        sv = qpso.Server(ctx, 32 << 30)
        sv.set_slots(slots)
        n_req = int(lam * seconds)
        gaps = rng.exponential(1.0 / lam, n_req)
        tickets = []
        t0 = time.perf_counter()
        t_next = t0
        for k in range(n_req):
            t_next += gaps[k]
            while True:
                now = time.perf_counter()
                if now >= t_next:
                    break
                time.sleep(min(0.0005, t_next - now))
            mid, cl = pool[k % len(pool)]
            tickets.append(sv.submit(mid, cl, params))
        for tk in tickets:
            sv.wait(tk)
        st = sv.stats()
        sv.close()
        runs.append({"lambda_per_s": lam, "requests": n_req, "completed": int(st.completed),
                     "solves_per_s": st.solves_per_s, "p50_ms": st.p50_ms, "p99_ms": st.p99_ms,
                     "mean_resident_solves": st.mean_active})
    for m in mids:
        ctx.release_mesh(m)
    return {"workload": f"Poisson arrivals for {seconds:.1f} s per rate, 8 resident meshes 1k..50k faces, "
                        f"128 particles, 60 iterations, 32 GiB budget, {slots} particle slots",
            "runs": runs, "autotune": tune[0]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--timed-solves", type=int, default=5, help="solves timed per kernel with HIP events")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=10)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--lam", type=str, default="200,1000,4000", help="arrival rates of the served section")
    ap.add_argument("--slots", type=int, default=8192, help="server particle slots (particles rendered per step)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(independent solves shard across GPUs with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    scene, params = W.config1(0)
    c = scene.clusters[0]
    F = scene.objects[0][0].size // 9
    rw, rh = c["rect"][2] - c["rect"][0], c["rect"][3] - c["rect"][1]
    ctx = qpso.Context(0)
    mid = ctx.upload_mesh(scene.mesh_m())
    solve = ctx.solve(mid, scene.cluster(), params)
    tune = ctx.autotune([solve], iters=3)
    for _ in range(args.warmup):
        solve.run()
    # timed region: K whole solves, each one graph launch (init + I x 2 kernels)
    lat, results = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        results.append(solve.run())
        lat.append((time.perf_counter() - t) * 1e3)
    wall = time.perf_counter() - t0
    value = args.steps / wall
    # per-kernel durations: HIP events between every two kernels on the
    # launch stream, over the same solve right after the timed region
    k1_ms = k2_ms = 0.0
    for _ in range(args.timed_solves):
        r, tm = solve.run_timed()
        k1_ms += tm.ms_render_fitness
        k2_ms += tm.ms_pso
        results.append(r)
    r0 = results[0]
    assert all(r.fit == r0.fit for r in results), "solves are not deterministic"
    launches = args.timed_solves * params.n_iters
    k1_avg_s = k1_ms / launches * 1e-3
    lds_bytes = (4.0 * r0.raster_px + 10.0 * r0.visit_px) / params.n_iters
    hbm_bytes = 36.0 * F + 2.0 * rw * rh + 7 * 4 * 5 * params.n_particles
    lds_gbs = lds_bytes / k1_avg_s / 1e9
    hbm_gbs = hbm_bytes / k1_avg_s / 1e9
    out = {
        "metric": "pose solves/s + p50/p99 latency, batched-served vs plain, 1 MI355X",
        "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "i32+f32", "data": "synthetic",
        "config": {"workload": f"plain single solve: {F}-face ellipsoid + 8 boxes, ROI {rw}x{rh}, "
                               f"{params.n_particles} particles, {params.n_iters} iterations",
                   "faces": F, "roi": [rw, rh], "particles": params.n_particles,
                   "iterations": params.n_iters, "parallelism": "1 GPU"},
        "p50_ms": pct(lat, 0.5), "p99_ms": pct(lat, 0.99),
        "pixels_per_s": {"raster": r0.raster_px * value, "fitness": r0.visit_px * value},
        "kernel_ms": {"render_fitness_avg": k1_avg_s * 1e3, "pso_step_avg": k2_ms / launches},
        "roofline": {"bound": "hbm", "achieved": hbm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": hbm_gbs / HBM_PEAK_GBS, "traffic": from_profiles("hbm_bytes_per_launch"),
                     "algorithmic_bytes_per_launch": hbm_bytes,
                     "kernel": "k_render_fitness"},
        # the binding resource is on-chip: LDS depth-tile traffic (and VALU issue)
        "roofline_lds": {"bound": "lds", "achieved": lds_gbs, "peak": LDS_PEAK_GBS, "unit": "GB/s",
                         "frac": lds_gbs / LDS_PEAK_GBS, "algorithmic_bytes_per_launch": lds_bytes},
        "autotune": {"chosen": tune[0], "candidates": len(tune), "slowest_ms": tune[-1]["ms"]},
    }
    valu = from_profiles("valu_instr_per_launch")
    if valu is not None:
        tips = valu / k1_avg_s / 1e12
        out["roofline_valu"] = {"bound": "valu", "achieved": tips, "peak": VALU_PEAK_TIPS,
                                "unit": "T wave-instr/s", "frac": tips / VALU_PEAK_TIPS,
                                "instr_per_launch": valu}
    solve.close()
    ctx.release_mesh(mid)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, params, args.cpu_iters)
    if not args.no_batched:
        out["batched"] = batched_section(ctx)
    if not args.no_served:
        out["served"] = served_section(ctx, lams=[float(x) for x in args.lam.split(",")], slots=args.slots)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
