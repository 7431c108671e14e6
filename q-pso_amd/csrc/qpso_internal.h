// qpso_internal.h -- device-visible solve descriptor shared by the host code
// (qpso_api.cpp) and the kernels (qpso_kernels.hip).  Plain C layout.
#ifndef QPSO_INTERNAL_H_
#define QPSO_INTERNAL_H_

#include <stdint.h>

#include "qpso.h"

#define QPSO_EMPTY 60000          // maxVal_depthBufferInt (main.h:148)
#define QPSO_AABB_MIN_SENT 50000  // max_AABB_ (main.h:150)
#define QPSO_AABB_MAX_SENT (-1)   // min_AABB_ (main.h:151)
#define QPSO_BEST_FIT_INIT 1.0e10f
#define QPSO_FXP_LIMIT (1 << 30)  // degenerate guard on 28.4 coordinates
#define QPSO_AREA_LIMIT (1 << 23) // degenerate guard on the union box area

// flags[] bits
#define QPSO_FLAG_STALE_OVERFLOW 1
// a workgroup was launched with an instantiation that cannot hold its mesh
// or dump its frame (the host never does this; the workgroup writes NaN and
// stops)
#define QPSO_FLAG_LAUNCH 2

// carried-over pixels kept per particle and iteration; more set
// QPSO_FLAG_STALE_OVERFLOW
#define QPSO_STALE_CAP 1024

// device scratch of the allocation-free entry point (qpso_estimate_pose_noalloc)
#define QPSO_NOALLOC_SCRATCH_BYTES QPSO_NOALLOC_BYTES

struct QpsoStale {  // one carried-over depth-buffer pixel
    int32_t pos;    // (y << 16) | x
    int32_t val;
};

struct QpsoItem {  // one render+fitness workgroup: (solve index, particle index)
    int32_t solve, particle;
};

// One entry of a mesh's raster table: the three unique-vertex indices of a
// face and its index in the loadOBJ3 order.
struct QpsoFaceRef {
    int32_t v0, v1, v2, face;
};

struct QpsoSolveDev {
    // mesh (metres)
    int32_t F;       // faces
    int32_t F_eff;   // faces rendered (face cap)
    // deduplicated mesh: unique vertices [3][V] (x, y, z) and the raster order
    const float* verts;
    int32_t V;       // unique vertices (row stride of verts)
    int32_t V_eff;   // unique vertices referenced by faces < F_eff (a prefix)
    const QpsoFaceRef* faces;
    const float* corners;  // [F][3][4]: face f's corner j at corners[12 f + 4 j], (x, y, z, vertex index bits)
    int32_t n_faces; // entries in `faces` (= F)
    // runs of consecutive faces whose corners lie in a short vertex range:
    // (first face, first vertex, end vertex) per run, for meshes whose
    // vertices do not fit the projected-vertex cache at once
    const int32_t* runs;
    int32_t n_runs;
    // camera
    double fx, fy, cx, cy;
    // PSO
    int32_t N, I, S;  // particles, iterations, random stride
    double w;
    float c1, c2;
    float x_lo[3], x_hi[3];
    // cluster
    int32_t rect[4];        // minx, miny, maxx, maxy (max exclusive)
    uint32_t num_points;
    const uint16_t* obs;    // rect crop of the cluster depth, pitch rect[2]-rect[0]
    const uint16_t* obs_rows;  // [rect rows][pitch + 1]: observed pixels of the crop row left of each column
    // state
    const float* rand;      // [S][N]
    float* pos;             // [7][N]
    float* vel;
    float* pbest;
    float* pbest_fit;       // [N]
    float* fit;             // [N]
    float* gbest;           // [8]: pose + fit
    int32_t* iter;          // [1]
    QpsoStale* stale[2];    // [N][stale_cap] each
    int32_t* stale_cnt[2];  // [N] each
    int32_t stale_cap;
    int32_t big_face;       // flat faces of more clipped pixels (others: half) are rasterized by a whole wave
    int32_t* flags;         // [1]
    unsigned long long* counters;  // [3]: raster px, visit px, degenerate evals
    // traces (nullable)
    float* fit_trace;       // [I][N]
    float* gbest_trace;     // [I][8]
    int32_t* argmin_trace;  // [I]
    // reference-faithful off-frame mode (qpso_params.offframe = 1): the
    // reference's full per-particle depth buffers, [N][480][640], else null
    int32_t wrap;
    int32_t* frames;
    // debug outputs of the stage entry point (nullable)
    int32_t* dbg_frames;    // [N][480][640]
    int32_t* dbg_aabb;      // [N][4]
};

#endif
