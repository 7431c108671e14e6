/*
 * qpso.h -- C ABI of the MI355X (gfx950) Q-PSO pose-search hot path.
 *
 * The reference (morpheus1820/Q-PSO) has no library API: its hot path is the
 * per-cluster "estimate pose" block inlined in main() (main.cu:1269-1733),
 * reached from the CLI `main object_model rgb_image depth_image N`
 * (main.cu:1198-1204).  This header is the function boundary that block
 * becomes; each entry point names the reference code it replaces.  Plain C
 * types only; every call returns a status code (the reference exits the
 * process on any CUDA error, handle_error.h:14-23).
 *
 * Units and layouts follow the reference:
 *   mesh      fp32 struct-of-arrays, metres, loadOBJ3 order (main.h:620-679):
 *             x(face f, vertex j) = A[f + j F], y = A[3F + f + j F],
 *             z = A[6F + f + j F]; 9F floats
 *   pose      [tx, ty, tz (m), q0, q1, q2, q3] (main.h:86-94), q0 >= 0
 *   swarm     struct-of-arrays [7][N]: element (d, p) at d*N + p
 *             (idx = tIdx + d*Nparticle_, main.cu:164-170)
 *   depth     480x640 row-major uint16 millimetres (the cluster's masked,
 *             inpainted depth, GraphCannySeg.h:1805-1809)
 */
#ifndef QPSO_H_
#define QPSO_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QPSO_ROWS 480
#define QPSO_COLS 640

enum {
    QPSO_OK = 0,
    QPSO_E_INVAL = -1,     /* bad argument */
    QPSO_E_HIP = -2,       /* HIP runtime error (see qpso_last_error) */
    QPSO_E_NOMEM = -3,     /* device or host allocation failed */
    QPSO_E_BUDGET = -4,    /* request does not fit the server's HBM budget */
    QPSO_E_NOTFOUND = -5,  /* unknown mesh id / ticket */
    QPSO_E_STALE = -6,     /* a particle's stale-pixel list overflowed */
    QPSO_E_PENDING = -7,   /* ticket not finished yet */
    QPSO_E_CANCELLED = -8, /* request cancelled (qpso_server_cancel, qpso_server_destroy) */
    QPSO_E_TIMEOUT = -9    /* request not finished within the server's time-out */
};

typedef struct qpso_ctx qpso_ctx;
typedef struct qpso_solve qpso_solve;
typedef struct qpso_server qpso_server;

/* Camera intrinsics: the reference hard-codes them as double macros
 * FX, FY, CX, CY (main.h:30-33). */
typedef struct {
    double fx, fy, cx, cy;
} qpso_camera;

/* Per-solve PSO parameters.  The reference fixes these at compile time
 * (main.h:83-105, QPSOParams_t main.h:182-204, filled at main.cu:1303-1331). */
typedef struct {
    int32_t n_particles;   /* Nparticle_ (1024); at most 65535 */
    int32_t n_iters;       /* max_steps_ (20); at most 2^20 */
    double w;              /* psoW_ (0.3): inertia, applied in binary64 */
    float c1, c2;          /* c1_, c2_ (1.0f) */
    float t_bound;         /* x/y translation half-range around the centroid, m (0.3) */
    float t_bound_z;       /* z half-range, m (0.3) */
    uint32_t seed;         /* mt19937 seed; 0 means GSL's default 4357 */
    qpso_camera cam;
    int32_t face_cap;      /* 0 = every face; 3072 = the reference's 6x512 grid */
    int32_t offframe;      /* 0: depth-buffer accesses outside the 640x480 frame
                              are dropped (writes) or read as empty; 1: the
                              reference's own indexing, idx = y*640 + x in 32-bit
                              arithmetic (main.h:1442, main.cu:674, 688), so a
                              pixel past the left or right edge lands in the
                              neighbouring row of the particle's buffer; indices
                              outside that buffer (another particle's in the
                              reference) are still dropped.  Mode 1 keeps a
                              640x480 int32 buffer per particle in device memory
                              and runs in plain solves only (not batched, served,
                              autotuned or allocation-free) */
} qpso_params;

/* One segmented cluster (GraphCannySeg::SegResults, GraphCannySeg.h:594-618). */
typedef struct {
    int32_t rect[4];       /* rect_aabb_: tl.x, tl.y, br.x, br.y (br exclusive) */
    float centroid_mm[3];  /* centroid3D (mm) */
    uint32_t num_points;   /* num_points */
    const uint16_t* depth; /* clusterDepth, 480x640, zero outside the cluster;
                              only the rect is read (pixels outside it count
                              as unobserved, as they are in a cluster depth) */
} qpso_cluster;

typedef struct {
    float pose[7];         /* gbest pose (d_solution_best_pose) */
    float fit;             /* gbest fitness (d_solution_best_fit) */
    uint64_t raster_px;    /* bbox-fill pixel writes over the solve (P_r) */
    uint64_t visit_px;     /* fitness pixel visits over the solve (P_f) */
    uint32_t n_degenerate; /* particle evaluations scored NaN by the guard */
    int32_t status;        /* QPSO_OK or a per-solve error */
    double latency_ms;     /* served requests: submit to result; 0 otherwise */
} qpso_result;

/* Optional per-iteration traces of a solve (all host buffers, nullable). */
typedef struct {
    float* fit_trace;      /* [I][N]: d_pso_fit_error after each fitness pass */
    float* gbest_trace;    /* [I][8]: gbest pose + fit after each pbest/gbest update */
    int32_t* argmin_trace; /* [I]: best_particle_Idx of each iteration */
    float* pos;            /* [7][N] final d_pso_pos */
    float* vel;            /* [7][N] final d_pso_vel */
    float* pbest;          /* [7][N] final d_pso_pos_b */
    float* pbest_fit;      /* [N] final d_pso_personal_best_fit */
    int32_t* frames;       /* [N][480][640] final depth buffers (stale pixels) */
} qpso_trace;

/* Reference defaults: N 1024, I 20, w 0.3, c1 = c2 = 1, bounds 0.3 m, seed 0,
 * the ACCV intrinsics of main.h:30-33, no face cap. */
void qpso_params_default(qpso_params* p);

/* Context: device selection, streams, caches.  Replaces the process-level set
 * up of main.cu:1377-1574 (allocations, 16 streams, constant upload). */
int qpso_create(int device, qpso_ctx** out);
int qpso_destroy(qpso_ctx* ctx);
const char* qpso_last_error(qpso_ctx* ctx);
const char* qpso_status_str(int status);

/* Mesh upload (loadOBJ3 + cudaMalloc/cudaMemcpy of d_obj_model,
 * main.cu:1366-1373).  The mesh stays resident until released.  A release
 * while solves created on the mesh exist takes effect when the last of them
 * is destroyed; from the release on, new solves on it fail with
 * QPSO_E_NOTFOUND. */
int qpso_mesh_upload(qpso_ctx* ctx, const float* soa, uint32_t n_faces, uint32_t* mesh_id);
int qpso_mesh_release(qpso_ctx* ctx, uint32_t mesh_id);

/* The per-cluster estimate-pose block, blocking (main.cu:1269-1733):
 * params, random table, uploads, InitQPSO, I iterations of
 * render -> AABB -> fitness -> argmin -> pbest/gbest -> update, readback. */
int qpso_estimate_pose(qpso_ctx* ctx, uint32_t mesh_id, const qpso_cluster* cl,
                       const qpso_params* p, qpso_result* out);

/* qpso_estimate_pose without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  The solve's arena and random
 * table live in a file-scope device array of QPSO_NOALLOC_BYTES that every
 * GPU the library runs on holds until the process ends; a solve that needs
 * more (about N <= 1024 at I <= 200 with a 640x480 rect fits) returns
 * QPSO_E_NOMEM before touching the device.  It calls no hipMalloc, hipFree
 * or graph routine; the mesh must already be uploaded.  Calls on one GPU
 * share the array and run one at a time; it keeps the last call's
 * contents, and each call reinitialises what it reads.  Results equal
 * qpso_estimate_pose's. */
#define QPSO_NOALLOC_BYTES (40u << 20)
int qpso_estimate_pose_noalloc(qpso_ctx* ctx, uint32_t mesh_id, const qpso_cluster* cl,
                               const qpso_params* p, qpso_result* out);

/* The same block split in two so that the inputs are device-resident before
 * the solve runs: create uploads the cluster depth and the random table
 * (main.cu:1377-1557); run executes InitQPSO and the iteration loop
 * (main.cu:1600-1741) and may be called repeatedly. */
int qpso_solve_create(qpso_ctx* ctx, uint32_t mesh_id, const qpso_cluster* cl,
                      const qpso_params* p, int with_trace, qpso_solve** out);
int qpso_solve_run(qpso_solve* s, qpso_result* out);
int qpso_solve_get_trace(qpso_solve* s, qpso_trace* tr);
int qpso_solve_destroy(qpso_solve* s);

/* The same solve with the reference's launch structure, for comparison
 * only (main.cu:1617-1733): per iteration one render+fitness launch per
 * particle, particle p on stream p % 16 (NumStreams, main.h:130;
 * main.cu:1633-1644), then the PSO
 * step and a host read of the best fitness (the reference copies the
 * argmin to the host every iteration, main.cu:1691-1698).  The kernels are
 * this library's, so results equal qpso_solve_run's. */
int qpso_solve_run_reference_launches(qpso_solve* s, qpso_result* out);

/* qpso_solve_run with every kernel timed (the reference's myTimer brackets,
 * main.cu:1617-1733, per kernel): each launch stamps a start and a stop event
 * with its own dispatch, so the figures are kernel durations without event
 * records between the kernels. */
typedef struct {
    double ms_solve;          /* start of InitQPSO to the end of the last kernel */
    double ms_render_fitness; /* sum over the I render+fitness launches */
    double ms_pso;            /* sum over the I argmin/pbest/gbest/update launches */
    double ms_init;           /* InitQPSO */
    int32_t launches;         /* render+fitness launches (= I) */
} qpso_timing;
int qpso_solve_run_timed(qpso_solve* s, qpso_result* out, qpso_timing* tm);

/* Batched form: every listed solve runs concurrently, one fused
 * render+fitness launch per iteration over all their particles plus one PSO
 * launch, the iteration loop captured in a HIP graph.  Results are identical
 * to running each solve alone. */
int qpso_solve_run_batch(qpso_ctx* ctx, int n, qpso_solve* const* solves, qpso_result* out);

/* One measured candidate of qpso_autotune. */
typedef struct {
    int32_t config;      /* render+fitness configuration index */
    int32_t threads;     /* workgroup size (one workgroup per particle) */
    int32_t tile_words;  /* 32-bit LDS words per workgroup: vertex cache, face queues, depth tile */
    int32_t big_face;    /* flat faces over this many clipped pixels (others over half) are drawn by a whole wave */
    double ms;           /* InitQPSO + `iters` iterations of the listed solves */
} qpso_tune_entry;

/* Autotuner: times InitQPSO + `iters` iterations of the listed solves, run
 * together as in a batch, under every workgroup/tile configuration and
 * big-face threshold, and stores the fastest in each solve for its later
 * plain, timed, stage and batched runs; the choice also becomes the
 * context's default (qpso_set_tuning).  Results do not depend on the
 * choice.  `table` (nullable when cap is 0) receives up to `cap` candidates,
 * `*n_out` (nullable) their number.  The swarm state is left undefined; every
 * run starts with InitQPSO.  No counterpart in the reference, whose launch
 * shapes are compile-time constants (main.cu:1607, main.h:84-99). */
int qpso_autotune(qpso_ctx* ctx, int n, qpso_solve* const* solves, int iters,
                  qpso_tune_entry* table, int cap, int* n_out);

/* The context's default tuning: the configuration index (0 .. the number
 * of qpso_autotune candidates per threshold - 1) and big-face threshold
 * (>= 1) of solves and servers created from now on; a server keeps the
 * tuning it was created with.  Defaults: configuration 0 (or the
 * QPSO_RF_CFG environment variable), threshold 128.  No counterpart in the
 * reference. */
int qpso_set_tuning(qpso_ctx* ctx, int config, int big_face);

/* Stage entry point (RenderFixedPointKernelAABB + ComputeAABBFinalReduction +
 * ComputeFitnessKernel, main.cu:1633-1681) on caller-given poses [7][n]
 * (n <= N), starting from empty depth buffers.  Outputs (nullable):
 * fit[n], aabb[n][4] (minX, maxX, minY, maxY: d_finalAABB[p*32+0..3]),
 * frames[n][480][640] = the rendered buffers before the fitness pass. */
int qpso_solve_eval_poses(qpso_solve* s, int n, const float* poses,
                          float* fit, int32_t* aabb, int32_t* frames);

/* Serving: a host thread admits solves as they arrive, under an HBM budget
 * and a number of particle slots, and steps every admitted solve one
 * iteration per device step (continuous batching): a request that arrives
 * while others run joins them at the next step.  Inputs are copied at submit. */
typedef struct {
    uint64_t submitted, completed, rejected;
    double p50_ms, p99_ms, mean_ms;  /* submit -> result latency of completed requests */
    double solves_per_s;             /* completed / (last completion - first submit) */
    uint64_t steps;
    double mean_active;              /* mean solves resident per step */
    uint64_t cancelled, timed_out;
} qpso_server_stats;

/* The budget covers each resident solve's device bytes (qpso_solve_bytes)
 * plus each random table in use (4 S N bytes, S = max(512, 12 (I + 1)),
 * shared by solves of equal seed, N and I) once while resident solves use
 * it.  A request that could never fit is refused at submit (QPSO_E_BUDGET).
 * Requests are admitted in arrival order. */
int qpso_server_create(qpso_ctx* ctx, uint64_t hbm_budget_bytes, qpso_server** out);
/* Slots: the most particles resident at once (default 8192).  A step runs
 * one iteration of every resident solve, so fewer slots mean shorter steps
 * and lower latency, more slots fuller launches and higher throughput.  A
 * step with at most one particle per CU runs the widest configuration (0),
 * fuller steps the server's tuning (qpso_set_tuning).  A
 * solve with more particles than the slots runs alone.  No counterpart in the
 * reference (it solves one cluster at a time). */
int qpso_server_set_slots(qpso_server* sv, int64_t max_resident_particles);
int qpso_server_submit(qpso_server* sv, uint32_t mesh_id, const qpso_cluster* cl,
                       const qpso_params* p, uint64_t* ticket);
/* A ticket's result is read once, by poll (QPSO_E_PENDING while it runs) or
 * wait (blocks); both return the result's status, or QPSO_E_NOTFOUND for an
 * unknown or already read ticket. */
int qpso_server_poll(qpso_server* sv, uint64_t ticket, qpso_result* out);
int qpso_server_wait(qpso_server* sv, uint64_t ticket, qpso_result* out);
/* Cancel a request: a queued one ends at once, a resident one at the next
 * step boundary, with status QPSO_E_CANCELLED and its memory released.  A
 * finished one keeps its result.  QPSO_E_NOTFOUND for unknown or read
 * tickets.  No counterpart in the reference. */
int qpso_server_cancel(qpso_server* sv, uint64_t ticket);
/* Requests submitted from now on that have not finished `timeout_ms` after
 * their submission end with QPSO_E_TIMEOUT (checked at step boundaries);
 * 0 (the default) disables it.  No counterpart in the reference. */
int qpso_server_set_timeout(qpso_server* sv, double timeout_ms);
int qpso_server_stats_get(qpso_server* sv, qpso_server_stats* st);
/* Ends the requests that have no result yet with QPSO_E_CANCELLED, releases
 * threads blocked in qpso_server_wait, and frees the server.  No other call
 * on the server may start once destroy has begun. */
int qpso_server_destroy(qpso_server* sv);

/* Device bytes of one solve's own arena (observed crop, swarm state, carried
 * pixels), the admission unit; its random table is shared and counted
 * separately (see qpso_server_create). */
uint64_t qpso_solve_bytes(const qpso_params* p, const int32_t rect[4]);

#ifdef __cplusplus
}
#endif
#endif
