#!/usr/bin/env python3
"""Benchmark of the gfx950 Q-PSO pose-search hot path on one MI355X.

Headline workload: one plain solve of a 10,336-face ellipsoid with 8 box
protrusions (cluster rect 220x220 px), 256 particles, 100 iterations, inputs
device-resident before timing.  One step = one whole solve (InitQPSO + 100 x
(render+fitness, argmin/pbest/gbest/update)), one graph launch.  value =
solves/s.

Also reported on the same JSON line:
  kernel_ms     per-kernel durations over further solves in the same process,
                from start/stop events stamped by each kernel's own dispatch
  roofline      the render+fitness kernel against HBM: unique bytes per launch
                (mesh, observed ROI, swarm state) / measured launch time;
                traffic = HBM bytes per launch from rocprof (committed summary)
  roofline_lds  the same kernel against LDS bandwidth: algorithmic LDS bytes
                per launch (4 B per bbox-fill atomicMin + 10 B per fitness
                visit, counted exactly) / launch time
  roofline_valu the same kernel against VALU issue, the resource that binds
                it: VALU wave instructions per launch (rocprof SQ_INSTS_VALU,
                committed summary) / launch time
  cpu_baseline  the CPU oracle (C, OpenMP over particles) on the headline
                solve; cpu_baselines: the same for the 1,280-face solve and
                a sample of the 16-object frame
  batched       16 uneven solves of one frame: batched graph form vs plain,
                and vs the reference's launch structure (one launch per particle)
  served        Poisson arrivals into the continuous-batching server:
                saturation capacity, p50/p99 over each run's steady window,
                and the plain one-request-per-call rate of the same mix
  stress        64 solves of 2,048 particles x 200 iterations on a 100k-face
                mesh mixed with 64 small solves, through the server
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(REPO, "q-pso_amd"), os.path.join(REPO, "oracle")]

import numpy as np  # noqa: E402

import qpso  # noqa: E402
import workloads as W  # noqa: E402

# 256 CUs x 256 B/clk x 2.4 GHz = 157,286 GB/s (gfx950 LDS; 128-bit reads
# measure 146 TB/s on the device, qpso_peaks)
LDS_PEAK_GBS = 256 * 256 * 2.4
HBM_PEAK_GBS = 8000.0
# 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 VALU instruction per SIMD every 4
# clocks (16 lanes wide, packed FP32 off)
VALU_PEAK_TIPS = 256 * 4 * 2.4e9 / 4 / 1e12
PROFILE = os.path.join(REPO, "profiles", "k_render_fitness_10k_faces_220px_roi_256_particles_pmc.json")
PEAKS_TOOL = os.path.join(REPO, "q-pso_amd", "lib", "qpso_peaks")


def measured_peaks():
    """The device's own HBM copy, LDS atomic and LDS read ceilings
    (q-pso_amd/csrc/qpso_peaks.hip, built by `make -C q-pso_amd cal`), or
    None without the tool."""
    if not os.path.exists(PEAKS_TOOL):
        return None
    import subprocess
    try:
        p = subprocess.run([PEAKS_TOOL], capture_output=True, text=True, timeout=120)
        return json.loads(p.stdout.strip().splitlines()[-1]) if p.returncode == 0 else None
    except (subprocess.SubprocessError, ValueError, IndexError):
        return None


def pct(v, q):
    v = sorted(v)
    return v[min(len(v) - 1, int(round(q * (len(v) - 1))))] if v else None


def from_profiles(key):
    """A per-launch figure of the render+fitness kernel measured by rocprofv3
    --pmc on the headline workload (committed summary), or None."""
    if not os.path.exists(PROFILE):
        return None
    try:
        with open(PROFILE) as f:
            return float(json.load(f)[key])
    except (OSError, ValueError, KeyError):
        return None


def host_cpus():
    """(threads the oracle runs on, nproc, CPU model).  A one-GPU job gets a
    share of the host: OMP_NUM_THREADS where the launcher sets it, else the
    CPUs this process may run on."""
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS") or allowed)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, os.cpu_count(), allowed, model


def cpu_solve_rate(O, scene, params, i=0, iters=None, threads=1):
    """Oracle solves/s on `iters` of the solve's iterations, scaled to whole
    solves; also the seconds taken and the bbox-fill and fitness pixels of
    the run."""
    op = W.oracle_params(O, scene, params, i)
    if iters is not None:
        op.n_iters = iters
    rand = O.rand_table(params.seed, params.n_particles, params.n_iters)
    t = time.perf_counter()
    r = O.solve(scene.mesh_m(i), scene.clusters[i]["depth"], op, rand, n_threads=threads)
    dt = time.perf_counter() - t
    return (op.n_iters / params.n_iters) / dt, dt, r["raster_px"], r["visit_px"]


def cpu_median(runs):
    """Median solves/s of timed runs and the pixels/s of the median run."""
    runs = sorted(runs, key=lambda x: x[0])
    rate, dt, pr, pf = runs[len(runs) // 2]
    return rate, dt, {"raster": pr / dt, "fitness": pf / dt}


def cpu_baselines(scene1, params1):
    import oracle as O
    O.lib()
    threads, nproc, allowed, model = host_cpus()
    cpu_solve_rate(O, scene1, params1, iters=5, threads=threads)  # warm-up
    rate, dt, px = cpu_median([cpu_solve_rate(O, scene1, params1, threads=threads) for _ in range(5)])
    head = {"value": rate, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"median of 5 whole solves of the headline workload ({params1.n_particles} particles, "
                      f"{params1.n_iters} iterations), {dt:.1f} s each",
            "pixels_per_s": px, "nproc": nproc, "cpus_allowed": allowed, "cpu_model": model}
    s0, p0 = W.config0(0)
    cpu_solve_rate(O, s0, p0, threads=threads)  # warm-up
    r0, _, px0 = cpu_median([cpu_solve_rate(O, s0, p0, threads=threads) for _ in range(5)])
    s2, p2 = W.config2(0)
    idx = [i for i in range(len(p2)) if s2.clusters[i] is not None]
    t2, pr2, pf2, dt_run = 0.0, 0, 0, 0.0
    for i in idx:
        _, dt2, pr, pf = cpu_solve_rate(O, s2, p2[i], i, iters=10, threads=threads)
        t2 += dt2 * p2[i].n_iters / 10
        dt_run += dt2
        pr2 += pr
        pf2 += pf
    extra = {
        "config0": {"value": r0, "unit": "solves/s", "cores": threads, "pixels_per_s": px0,
                    "sample": "median of 5 whole solves (1,280 faces, 32 particles, 30 iterations)"},
        "config2_frame": {"value": len(idx) / t2, "unit": "solves/s", "cores": threads,
                          "pixels_per_s": {"raster": pr2 / dt_run, "fitness": pf2 / dt_run},
                          "sample": f"first 10 of 60 iterations of each of the frame's {len(idx)} solves, "
                                    "scaled to whole solves"},
    }
    return head, extra


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
            "batched_vs_plain": t_p / t_b, "reference_launch_solves_per_s": n / t_r,
            "batched_vs_reference_launches": t_r / t_b, "identical_results": bool(same), "autotune": tune[0]}


def run_server(ctx, jobs, arrivals, slots, budget=32 << 30, warm_frac=0.0):
    """Submit jobs[k] at time arrivals[k] (s from start); returns stats and
    the p50/p99 of the requests submitted after warm_frac of the run."""
    sv = qpso.Server(ctx, budget)
    sv.set_slots(slots)
    tickets = []
    t0 = time.perf_counter()
    for k, (mid, cl, p) in enumerate(jobs):
        t_next = t0 + arrivals[k]
        while True:
            now = time.perf_counter()
            if now >= t_next:
                break
            time.sleep(min(0.0005, t_next - now))
        tickets.append(sv.submit(mid, cl, p))
    res = [sv.wait(t) for t in tickets]
    st = sv.stats()
    sv.close()
    t_end = arrivals[-1] if len(arrivals) else 0.0
    lat = [r.latency_ms for r, a in zip(res, arrivals) if a >= warm_frac * t_end]
    return st, res, pct(lat, 0.5), pct(lat, 0.99)


def served_section(ctx, lams=(200.0, 1000.0, 4000.0), seconds=2.5, seed=0, slots=8192, n_sat=1500):
    """(frame, cluster) requests into the continuous-batching server: the
    saturation capacity (every request submitted at once), then Poisson
    arrivals per rate; p50/p99 over requests submitted after the first 30 %
    of each run (the steady window).  A rate above the capacity is an
    overload: its queue grows for the whole run."""
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
    # plain form: one qpso_estimate_pose call per request, one after another
    n_plain = min(len(pool), 48)
    t = time.perf_counter()
    for mid, cl in pool[:n_plain]:
        ctx.estimate_pose(mid, cl, params)
    plain_rate = n_plain / (time.perf_counter() - t)
    # saturation: every request at once
    jobs = [(pool[k % len(pool)][0], pool[k % len(pool)][1], params) for k in range(n_sat)]
    st, _, _, _ = run_server(ctx, jobs, np.zeros(n_sat), slots)
    capacity = st.solves_per_s
    runs = []
    for lam in lams:
        n_req = int(lam * seconds)
        arrivals = np.cumsum(rng.exponential(1.0 / lam, n_req))
        jobs = [(pool[k % len(pool)][0], pool[k % len(pool)][1], params) for k in range(n_req)]
        st, _, p50, p99 = run_server(ctx, jobs, arrivals, slots, warm_frac=0.3)
        runs.append({"lambda_per_s": lam, "requests": n_req, "completed": int(st.completed),
                     "solves_per_s": st.solves_per_s, "p50_ms": p50, "p99_ms": p99,
                     "mean_resident_solves": st.mean_active, "overload": bool(lam > capacity)})
    for m in mids:
        ctx.release_mesh(m)
    return {"workload": f"8 resident meshes 1k..50k faces, 4..16 clusters per frame, 128 particles, 60 iterations, "
                        f"32 GiB budget, {slots} particle slots; Poisson arrivals for {seconds:.1f} s per rate",
            "capacity_solves_per_s": capacity, "plain_solves_per_s": plain_rate,
            "served_vs_plain": capacity / plain_rate, "runs": runs, "autotune": tune[0]}


def stress_section(ctx, n_stress=64, n_small=64, stress_iters=200):
    """The stress load (by default 64 solves of 2,048 particles x 200
    iterations on the 100k-face mesh, mixed with 64 small 1k-face /
    64-particle solves of 20..200 iterations) submitted at once under a 32 GiB
    budget; and the plain rate of the same stress solve (one
    qpso_estimate_pose call)."""
    big, small, jobs = W.config4_mix(0, n_stress, n_small, stress_iters=stress_iters)
    scenes = [big, small]
    mids = [ctx.upload_mesh(sc.mesh_m()) for sc in scenes]
    js = [(mids[k], scenes[k].cluster(), p) for k, p in jobs]
    slots = sum(p.n_particles for _, p in jobs)
    st, res, p50, p99 = run_server(ctx, js, np.zeros(len(js)), slots)
    ok = all(r.status == 0 for r in res)
    lat_big = [r.latency_ms for (k, _), r in zip(jobs, res) if k == 0]
    lat_small = [r.latency_ms for (k, _), r in zip(jobs, res) if k == 1]
    t = time.perf_counter()
    for _, p in [j for j in jobs if j[0] == 0][:2]:
        ctx.estimate_pose(mids[0], big.cluster(), p)
    plain_big = 2 / (time.perf_counter() - t)
    for m in mids:
        ctx.release_mesh(m)
    rect = big.clusters[0]["rect"]
    return {"workload": f"{n_stress} solves of a {big.objects[0][0].size // 9}-face mesh (ROI "
                        f"{rect[2] - rect[0]}x{rect[3] - rect[1]}, 2048 particles, {stress_iters} iterations, sigma 4 mm, "
                        f"15% holes) + {n_small} solves of a {small.objects[0][0].size // 9}-face mesh "
                        f"(64 particles, 20..200 iterations), submitted at once, 32 GiB budget",
            "solves": len(jobs), "all_ok": ok, "solves_per_s": st.solves_per_s,
            "stress_solves_per_s": n_stress / (max(lat_big) * 1e-3) if lat_big else None,
            "p50_ms": p50, "p99_ms": p99, "stress_p50_ms": pct(lat_big, 0.5), "stress_p99_ms": pct(lat_big, 0.99),
            "small_p50_ms": pct(lat_small, 0.5), "small_p99_ms": pct(lat_small, 0.99),
            "mean_resident_solves": st.mean_active, "plain_stress_solves_per_s": plain_big}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--timed-solves", type=int, default=5, help="solves timed per kernel")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-stress", action="store_true")
    ap.add_argument("--lam", type=str, default="200,1000,4000", help="arrival rates of the served section")
    ap.add_argument("--slots", type=int, default=8192, help="server particle slots (particles rendered per step)")
    ap.add_argument("--served-seconds", type=float, default=2.5, help="Poisson arrivals per rate (s)")
    ap.add_argument("--sat-requests", type=int, default=1500, help="requests of the saturation run")
    ap.add_argument("--stress-solves", type=int, default=64, help="solves of each kind in the stress mix")
    ap.add_argument("--stress-iters", type=int, default=200, help="iterations of the stress mix's large solves")
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
    # per-kernel durations over further solves of the same workload: every
    # launch stamps a start and a stop event with its own dispatch
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
    peaks = measured_peaks()
    if peaks:
        out["measured_peaks"] = peaks
        out["roofline"]["peak_measured"] = peaks["hbm_copy_gbs"]
        out["roofline"]["frac_measured"] = hbm_gbs / peaks["hbm_copy_gbs"]
        out["roofline_lds"]["peak_measured"] = peaks["lds_read_gbs"]
        out["roofline_lds"]["frac_measured"] = lds_gbs / peaks["lds_read_gbs"]
        # the raster's pixel writes are LDS atomicMin operations, one per P_r
        atomics = r0.raster_px / params.n_iters / k1_avg_s / 1e9
        out["roofline_lds_atomic"] = {"bound": "lds-atomic", "achieved": atomics, "peak": peaks["lds_atomic_gops"],
                                      "unit": "G atomics/s", "frac": atomics / peaks["lds_atomic_gops"],
                                      "peak_source": "measured on this device (qpso_peaks)"}
    valu = from_profiles("valu_instr_per_launch")
    if valu is not None:
        tips = valu / k1_avg_s / 1e12
        out["roofline_valu"] = {"bound": "valu", "achieved": tips, "peak": VALU_PEAK_TIPS,
                                "unit": "T wave-instr/s", "frac": tips / VALU_PEAK_TIPS,
                                "instr_per_launch": valu}
    solve.close()
    ctx.release_mesh(mid)
    if not args.no_cpu:
        out["cpu_baseline"], out["cpu_baselines"] = cpu_baselines(scene, params)
    if not args.no_batched:
        out["batched"] = batched_section(ctx)
    if not args.no_served:
        out["served"] = served_section(ctx, lams=[float(x) for x in args.lam.split(",")], seconds=args.served_seconds,
                                       slots=args.slots, n_sat=args.sat_requests)
    if not args.no_stress:
        out["stress"] = stress_section(ctx, args.stress_solves, args.stress_solves, args.stress_iters)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
