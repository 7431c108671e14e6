/*
 * qpso_oracle.h -- CPU restatement of the Q-PSO pose-search hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP product
 * path.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it; the product library never links or calls it.
 *
 * It restates, in plain C, the semantics of the reference CUDA code
 * (morpheus1820/Q-PSO):
 *   main.cu:72-198    InitQPSO
 *   main.cu:305-419   RenderFixedPointKernelAABB  (+ main.h:514-545 projection,
 *                     main.h:1306-1491 renderFixedPointAABB,
 *                     main.h:1190-1304 AABBreduction)
 *   main.cu:423-546   ComputeAABBFinalReductionKernel
 *   main.cu:550-764   ComputeFitnessKernel (+ main.h:1522-1573 reduction)
 *   main.cu:1691-1698 thrust::min_element argmin
 *   main.cu:768-855   UpdatePersonalAndGlobalBest
 *   main.cu:858-1060  UpdateParticlesEquationKernel (+ main.h:208-413)
 *   main.cu:1337-1352 GSL mt19937 random table
 *   main.cu:1617-1733 host iteration loop
 *
 * Arithmetic convention (the reference was built with nvcc -use_fast_math,
 * whose multiply-add contraction cannot be observed here): every a*b+c site
 * is a single-rounding fmaf/fma, fusing the FIRST product of a sum of two
 * products (a*b + c*d -> fma(a,b,c*d); a*b - c*d -> fma(a,b,-(c*d));
 * x - a*b -> fma(-a,b,x)).  Divisions written __frcp_rn in the reference are
 * correctly rounded 1/x.  sqrtf is IEEE.  sin/cos/acos are this project's own
 * implementations (Cody-Waite reduction + Taylor series; acos via the
 * half-angle asin identity), not CUDA's.  Signed 32-bit integer arithmetic
 * wraps (as the reference's hardware does).
 *
 * Deviations from the reference, fixed as the project's spec (the reference
 * reads/writes out of bounds or races in these cases):
 *   Q4  depth-buffer accesses outside the 640x480 frame are dropped (writes)
 *       or read as empty (reads) instead of wrapping into other rows or
 *       other particles' buffers (main.h:1442, main.cu:674,688).  With
 *       `wrap` set (the reference-faithful mode) the linear index
 *       y*640 + x, in wrapping 32-bit arithmetic, addresses the particle's
 *       own buffer as the reference's does wherever it falls inside it;
 *       only indices outside it (other particles' buffers or outside the
 *       allocation in the reference) are dropped or read as empty.
 *   Q5  all faces are rendered (the reference ignores faces >= 3072,
 *       main.h:162-163); face_cap > 0 restores the cap.
 *   Q6  the random table has S = max(512, 12*(I+1)) entries per particle,
 *       layout T[p + s*N] (reference: 512 per particle, N = 1024).
 *   Q8  argmin treats NaN as +inf, first index on ties.
 *   Q9  the per-particle AABB is the true min/max (main.cu:533-543 races).
 *   Q3  a pixel visited twice by the fitness scan reads the rendered value on
 *       both visits (the reference races, main.cu:675-681).
 *   G   a particle whose projected 28.4 fixed-point vertex coordinates exceed
 *       2^30 in magnitude, or whose union box exceeds 2^23 pixels, is
 *       "degenerate": its fitness is NaN and its depth buffer is reset to
 *       empty (the reference writes far outside its buffers there).
 */
#ifndef QPSO_ORACLE_H_
#define QPSO_ORACLE_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_ROWS 480
#define OR_COLS 640
#define OR_FRAME (OR_ROWS * OR_COLS)
#define OR_EMPTY 60000           /* maxVal_depthBufferInt, main.h:148 */
#define OR_AABB_MIN_SENT 50000   /* max_AABB_, main.h:150 */
#define OR_AABB_MAX_SENT (-1)    /* min_AABB_, main.h:151 */
#define OR_BEST_FIT_INIT 1.0e10f /* max_best_fit_, main.h:100 */
#define OR_NDIM 7                /* tx,ty,tz,q0,q1,q2,q3: main.h:86-94 */

typedef struct {
    double fx, fy, cx, cy;       /* FX,FY,CX,CY (double macros, main.h:30-33) */
} or_cam;

typedef struct {
    int n_particles;             /* N (Nparticle_, main.h:84) */
    int n_iters;                 /* I (max_steps_, main.h:99) */
    double w;                    /* inertia (psoW_, a double, main.h:105) */
    float c1, c2;                /* c1_, c2_ (main.h:103-104) */
    float x_lo[3], x_hi[3];      /* translation bounds, metres (main.cu:1324-1331) */
    int rect[4];                 /* cluster rect: minx, miny, maxx, maxy; max exclusive */
    uint32_t num_points;         /* SegResults::num_points (main.cu:1318) */
    int face_cap;                /* 0 = all faces; 3072 = reference grid cap */
    or_cam cam;
    int wrap;                    /* 0: off-frame accesses clipped (Q4); 1: linear-index wrap */
} or_params;

typedef struct {
    int32_t aabb[4];             /* minX, maxX, minY, maxY of the rendered faces */
    int degenerate;              /* guard G */
    uint64_t raster_px;          /* in-frame bbox-fill pixel writes (P_r) */
} or_render_info;

typedef struct {
    uint32_t seg, tr, n_rendered;
    float err;
    uint64_t visit_px;           /* in-frame fitness visits (P_f) */
} or_fitness_info;

/* ---- random table (main.cu:1337-1352): GSL mt19937, seed 0 -> 4357 ---- */
void or_mt19937_raw(uint32_t seed, uint32_t* out, size_t n);
void or_rand_table(uint32_t seed, float* out, size_t n);
size_t or_rand_stride(int n_iters);  /* S = max(512, 12*(I+1)) */

/* ---- scalar math shared by the PSO update ---- */
float or_sinf(float x);
float or_cosf(float x);
float or_acosf(float x);
float or_rcp(float x);
int32_t or_rint_sat(float x);

/* ---- stages ---- */
void or_project(const float pose[7], const float v[3], const or_cam* cam, float uvz[3]);
void or_render(const float* mesh, int F, int face_cap, const float pose[7],
               const or_cam* cam, const int rect[4], int32_t* frame, or_render_info* info);
float or_fitness(const or_render_info* rinfo, const int rect[4], uint32_t num_points,
                 int32_t* frame, const uint16_t* obs, or_fitness_info* finfo);
/* the same with the off-frame mode chosen (wrap: see Q4) */
void or_render_ex(const float* mesh, int F, int face_cap, const float pose[7], const or_cam* cam,
                  const int rect[4], int wrap, int32_t* frame, or_render_info* info);
float or_fitness_ex(const or_render_info* rinfo, const int rect[4], uint32_t num_points, int wrap,
                    int32_t* frame, const uint16_t* obs, or_fitness_info* finfo);
int or_argmin(const float* fit, int n);
void or_init(const or_params* p, const float* rand, float* pos, float* vel,
             float* pbest, float* pbest_fit, float* gbest);
void or_update_best(int n, const float* fit, float* pbest_fit, const float* pos,
                    float* pbest, float* gbest, int best);
void or_update_particles(const or_params* p, const float* rand, int rand_idx,
                         float* pos, float* vel, const float* pbest, const float* gbest);
void or_kinematics(float q[4], const float w[3]);

/* ---- whole solve (main.cu:1600-1733) ---- */
typedef struct {
    float best_pose[7];
    float best_fit;
    float* fit_trace;            /* optional [I*N]: fitness of every evaluation */
    float* gbest_trace;          /* optional [I*8]: gbest pose + fit after each iteration */
    int* argmin_trace;           /* optional [I] */
    float* pos, *vel, *pbest;    /* optional [7*N]: final swarm state */
    float* pbest_fit;            /* optional [N] */
    int32_t* frames;             /* optional [N*640*480]: final per-particle buffers */
    uint64_t raster_px, visit_px;
    int n_degenerate;
} or_solve_out;

int or_solve(const float* mesh, int F, const uint16_t* obs, const or_params* p,
             const float* rand, int n_threads, or_solve_out* out);

#ifdef __cplusplus
}
#endif
#endif
