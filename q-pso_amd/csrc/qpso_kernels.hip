// qpso_kernels.hip -- gfx950 kernels of the Q-PSO pose-search hot path.
//
//   k_init            InitQPSO (main.cu:72-198) for every particle of every
//                     listed solve; resets the solve's gbest, iteration
//                     counter and stale-pixel lists.
//   k_render_fitness  one workgroup per (solve, particle): fuses
//                     RenderFixedPointKernelAABB (main.cu:305-419),
//                     ComputeAABBFinalReductionKernel (main.cu:423-546) and
//                     ComputeFitnessKernel (main.cu:550-764).  The depth
//                     buffer is an int32 tile in LDS covering the in-frame
//                     part of the union box, processed in row bands when the
//                     box exceeds the tile; ds_min_i32 replaces the
//                     reference's global atomicMin.
//   k_pso_step        one workgroup per solve: thrust::min_element argmin
//                     (main.cu:1691-1698), UpdatePersonalAndGlobalBest
//                     (main.cu:768-855) and UpdateParticlesEquationKernel
//                     (main.cu:858-1060), with no host round trip.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <type_traits>

#include "qpso_internal.h"
#include "qpso_kernels.h"
#include "qpso_math.h"

using namespace qpso;

namespace {

constexpr int kRows = 480;
constexpr int kCols = 640;

// Global-memory pointers: loads through them cannot alias LDS, so the compiler
// may move them across LDS stores and atomics (generic pointers could).
#define QG __attribute__((address_space(1)))
typedef int32_t qv4 __attribute__((ext_vector_type(4)));
typedef int32_t qv2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ QpsoFaceRef load_ref(const QG QpsoFaceRef* p, int i) {
    const qv4 v = reinterpret_cast<const QG qv4*>(p)[i];
    return QpsoFaceRef{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ QpsoStale load_stale(const QG QpsoStale* p, int i) {
    const qv2 v = reinterpret_cast<const QG qv2*>(p)[i];
    return QpsoStale{v.x, v.y};
}
__device__ __forceinline__ void store_stale(QG QpsoStale* p, int i, int32_t pos, int32_t val) {
    reinterpret_cast<QG qv2*>(p)[i] = qv2{pos, val};
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ unsigned wave_sum(unsigned v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += (unsigned)__shfl_xor((int)v, o);
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned lo = (unsigned)__shfl_xor((int)(unsigned)v, o);
        const unsigned hi = (unsigned)__shfl_xor((int)(unsigned)(v >> 32), o);
        v += ((unsigned long long)hi << 32) | lo;
    }
    return v;
}

#define RT(S, p, s) (S).rand[(size_t)(p) + (size_t)(s) * (size_t)(S).N]

// ------------------------------------------------------------------------
// InitQPSO (main.cu:72-198) + QuatKinematicsPSOStateVectWithInit (main.h:267-307)
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_init(const QpsoSolveDev* __restrict__ solves) {
    const QpsoSolveDev& S = solves[blockIdx.x];
    const int N = S.N;
    for (int p = threadIdx.x; p < N; p += blockDim.x) {
        float P[7], V[7];
        int s = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            P[d] = __builtin_fmaf(S.x_hi[d] - S.x_lo[d], RT(S, p, s), S.x_lo[d]); ++s;
            V[d] = __builtin_fmaf(6.0f, RT(S, p, s), -3.0f); ++s;  // -vdot_ + 2 vdot_ r
        }
        const float qinit[4] = {0.7071f, 0.7071f, 0.0f, 0.0f};
        float wp[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) { wp[d] = __builtin_fmaf(20.0f, RT(S, p, s), -10.0f); ++s; }
        float q[4];
        kinematics_from(qinit, q, wp);
#pragma unroll
        for (int i = 0; i < 4; ++i) P[3 + i] = q[i];
        V[3] = 0.0f;
#pragma unroll
        for (int d = 0; d < 3; ++d) { V[4 + d] = __builtin_fmaf(20.0f, RT(S, p, s), -10.0f); ++s; }
#pragma unroll
        for (int d = 0; d < 7; ++d) {
            S.pos[d * N + p] = P[d];
            S.vel[d * N + p] = V[d];
            S.pbest[d * N + p] = P[d];
        }
        S.pbest_fit[p] = QPSO_BEST_FIT_INIT;
        S.stale_cnt[0][p] = 0;
        S.stale_cnt[1][p] = 0;
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int d = 0; d < 7; ++d) S.gbest[d] = 0.0f;
        S.gbest[7] = QPSO_BEST_FIT_INIT;
        *S.iter = 0;
        *S.flags = 0;
        S.counters[0] = 0;
        S.counters[1] = 0;
        S.counters[2] = 0;
    }
}


// ------------------------------------------------------------------------
// fused render + AABB + fitness
// ------------------------------------------------------------------------
// Number of times the fitness scan (main.cu:632-733) visits the pixel at
// offset (xb, k) of the union box: the row index of t is the fp32 product
// trunc(t * rcp(W)), so for t = kW it falls one row short when that product
// rounds below k; (0, k) is then skipped by t = kW and visited by t = (k+1)W.
// __mul24 (the full-rate 24-bit multiply) forms products of pixel
// coordinates and row lengths: every such operand is below 2^23 in magnitude
// (the union box is at most 2^23 pixels), so the product is exact.
struct Visits {
    int32_t W, H, T;
    float rW;
    __device__ __forceinline__ bool short_row(int32_t k) const {
        return k > 0 && row_of(__mul24(k, W), rW) == k - 1;
    }
    __device__ __forceinline__ int32_t at(int32_t xb, int32_t k) const {
        if (xb != 0) return (k < H || __mul24(k, W) + xb < T) ? 1 : 0;
        return ((__mul24(k, W) < T && !short_row(k)) ? 1 : 0) + ((__mul24(k + 1, W) < T && short_row(k + 1)) ? 1 : 0);
    }
};

// Waves per SIMD when as many workgroups as the LDS allows share a CU (160
// KiB of LDS, 4 SIMDs): the register budget both bodies of a windowed
// kernel must fit, so that its extra code never costs a workgroup.  Left
// free (1) for 256-thread workgroups: the 10,112-word tile would need 4 waves
// and 128 registers, which its bodies exceed without spilling at 3.
constexpr int rf_waves_per_simd(int bt, int cap) {
    return bt >= 512 ? (bt / 256) * ((160 * 1024) / (cap * 4 + 512)) : 1;
}

// The render kernel's LDS, declared once per kernel and shared by the bodies
// it may run.
template <int CAP>
struct RfLds {
    __attribute__((aligned(16))) int32_t lds[CAP];  // [vertex cache: (X, Y) pairs, Z] [depth tile]
    int32_t s_box[4];    // rendered AABB: minX, maxX, minY, maxY
    int32_t s_stale[4];  // carried pixels inside the union box: x0, x1, y0, y1 (inclusive)
    int32_t s_degen;
    int32_t s_nstale;
    int32_t s_nq;        // small non-flat faces queued
    int32_t s_nqb;       // big faces queued
    int32_t s_take;      // next big face to hand to a wave
    int32_t s_later[3];  // a face of the window reaches later bands (slot: windows drawn mod 3)
    uint32_t s_cnt[3];   // seg, tr, n_rendered
    unsigned long long s_px[2];
};

// WIN: the body that can draw a mesh window by window (larger than the vertex
// cache); the other one keeps the cached path free of it.
template <int BT, int CAP, int MODE>
__device__ __forceinline__ void rf_body(const QpsoSolveDev* __restrict__ solves, const QpsoItem* __restrict__ items,
                                        RfLds<CAP>& sh) {
    constexpr bool WIN = MODE >= 1;     // meshes beyond the vertex cache, window by window
    constexpr bool FRAMES = MODE == 2;  // and the frame dump (stage calls)
    constexpr int SPT = 2048 / BT;  // fitness slots per thread
    int32_t* const lds = sh.lds;
    int32_t* const s_box = sh.s_box;
    int32_t* const s_stale = sh.s_stale;
    int32_t& s_degen = sh.s_degen;
    int32_t& s_nstale = sh.s_nstale;
    int32_t& s_nq = sh.s_nq;
    int32_t& s_nqb = sh.s_nqb;
    int32_t& s_take = sh.s_take;
    int32_t* const s_later = sh.s_later;
    uint32_t* const s_cnt = sh.s_cnt;
    unsigned long long* const s_px = sh.s_px;

    const int tid = threadIdx.x;
    // no item table: one solve, workgroup b renders particle b
    const QpsoItem item = items ? items[blockIdx.x] : QpsoItem{0, (int32_t)blockIdx.x};
    const QpsoSolveDev& S = solves[item.solve];
    const int p = item.particle;
    // Start-up loads go out together through global (not flat) pointers, so
    // no barrier waits on them: the iteration, the pose, both carried-pixel
    // counts and the vertices of pass A (kVPT per thread, the whole vertex
    // cache).  Then the carried pixels (at most QPSO_STALE_CAP, kSPT per
    // thread, kept in registers for the whole pass).  All are in bounds
    // whatever the iteration, so they load before the finished-solve test.
    const int N = S.N;
    const int iter_v = ((const QG int32_t*)S.iter)[0];
    float pose[7];
#pragma unroll
    for (int d = 0; d < 7; ++d) pose[d] = ((const QG float*)S.pos)[d * N + p];
    const int n_old0 = ((const QG int32_t*)S.stale_cnt[0])[p];
    const int n_old1 = ((const QG int32_t*)S.stale_cnt[1])[p];
    constexpr int kWin = CAP / 6;  // vertex cache capacity
    constexpr int kVPT = (kWin + BT - 1) / BT;
    const int Ve = S.V_eff;
    const QG float* __restrict__ wx = (const QG float*)S.verts;
    const QG float* __restrict__ wy = wx + S.V;
    const QG float* __restrict__ wz = wy + S.V;
    float ux[kVPT], uy[kVPT], uz[kVPT];
#pragma unroll
    for (int j = 0; j < kVPT; ++j) {
        const int v = min(tid + j * BT, Ve - 1);  // Ve >= 1
        ux[j] = wx[v]; uy[j] = wy[v]; uz[j] = wz[v];
    }
    const int iter = __builtin_amdgcn_readfirstlane(iter_v);
    const int par = iter & 1;
    const int n_old = __builtin_amdgcn_readfirstlane(par ? n_old1 : n_old0);
    const int cap = S.stale_cap;
    constexpr int kSPT = (QPSO_STALE_CAP + BT - 1) / BT;
    QpsoStale carried[kSPT];
    {
        const QG QpsoStale* __restrict__ old = (const QG QpsoStale*)(par ? S.stale[1] : S.stale[0]) + (size_t)p * cap;
#pragma unroll
        for (int j = 0; j < kSPT; ++j) {
            const int k = tid + j * BT;
            carried[j] = k < n_old ? load_stale(old, k) : QpsoStale{0, 0};
        }
    }
    if (iter >= S.I) return;  // this solve has finished (batched runs)
    // each instantiation only for the meshes and outputs it handles
    if ((!WIN && Ve > kWin) || (!FRAMES && S.dbg_frames)) {
        if (tid == 0) {
            S.fit[p] = __builtin_nanf("");
            atomicOr(S.flags, QPSO_FLAG_LAUNCH);
        }
        return;
    }

    const Cam cam{S.fx, S.fy, S.cx, S.cy};
    const int Fe = S.F_eff;
    const QG QpsoFaceRef* __restrict__ faces = (const QG QpsoFaceRef*)S.faces;
    const int n_faces = S.n_faces;
    const int32_t big_face = S.big_face;

    if (tid == 0) {
        s_box[0] = QPSO_AABB_MIN_SENT; s_box[1] = QPSO_AABB_MAX_SENT;
        s_box[2] = QPSO_AABB_MIN_SENT; s_box[3] = QPSO_AABB_MAX_SENT;
        s_stale[0] = INT32_MAX; s_stale[1] = INT32_MIN;
        s_stale[2] = INT32_MAX; s_stale[3] = INT32_MIN;
        s_degen = 0; s_nstale = 0; s_nq = 0; s_nqb = 0; s_take = 0;
        s_later[0] = s_later[1] = s_later[2] = 0;
        s_cnt[0] = s_cnt[1] = s_cnt[2] = 0;
        s_px[0] = s_px[1] = 0;
    }
    __syncthreads();

    // The projected unique vertices of a mesh of at most kWin vertices stay in
    // LDS.  A larger mesh is cut on the host into runs of faces whose corners
    // lie in a short range of vertex indices; each band then takes the runs
    // in windows (consecutive runs whose joint range fits the cache) and
    // projects the windows in turn.  Without usable runs every face's corners
    // are projected where needed.
    //   A vertex takes three words, X and Y read together, then Z; or, packed,
    // one word, X and Y as 16-bit halves, and a drawn face's corner depths
    // are projected again from the face's corner table.  Packing gives the
    // tile two words per vertex for that re-projection.  It is taken when
    // every 28.4 coordinate of the particle fits the halves (checked after
    // pass A) and, for a cached mesh, a depth region covering the whole
    // cluster rect would not fit beside the three-word cache in one band.
    // Windows always pack: a window then holds 2 kWin vertices and the tile
    // gains kWin words, and a face's cull reads one word per corner (measured
    // faster even where the band count stays the same).
    const bool cached = !WIN || Ve <= kWin;  // compile-time true without windows
    const bool runs = WIN && !cached && S.n_runs > 0 && S.n_runs * 256 <= Fe;  // runs of 256 faces on average
    int VC = cached ? Ve : (runs ? kWin : 0);  // cache capacity in vertices (three-word)
    constexpr int kQCap = CAP / 32;   // small faces needing the per-pixel depth
    constexpr int kBQCap = CAP / 128;
    constexpr int kRing = 128;
    constexpr int kFixed = kQCap + 4 * kBQCap + (BT / 64) * kRing + 4;  // queues, rings, alignment
    bool pack = runs;
    if (cached) {
        const int32_t rw = S.rect[2] - S.rect[0], rh = S.rect[3] - S.rect[1];
        pack = rw * rh + rh * (((rw + 31) >> 5) + 1) + rw + 8 > CAP - kFixed - 3 * VC;
    }
    uint32_t* __restrict__ vP = reinterpret_cast<uint32_t*>(lds);  // packed: X | Y << 16
    int2* __restrict__ vXY = reinterpret_cast<int2*>(lds);  // X, Y in one 8-byte read
    float* __restrict__ vZ = reinterpret_cast<float*>(lds + 2 * VC);  // set again below for windows
    int32_t vbase = 0;  // vertex index of cache entry 0
    // windows: the vertices [cbase, chi) the cache holds projected (entry
    // v - cbase), so that a window inside them is not projected again
    int32_t cbase = 0, chi = 0;

    // ---- pass A: per-face bbox and AABB over every face (main.h:1340-1353,
    //      AABBreduction main.h:1190-1304, final reduction main.cu:423-546).
    //      (m + 15) >> 4 is monotone for |m| <= 2^30, so the face-bbox
    //      extremes are those of the vertex coordinates.
    {
        int bx0 = QPSO_AABB_MIN_SENT, bx1 = QPSO_AABB_MAX_SENT;
        int by0 = QPSO_AABB_MIN_SENT, by1 = QPSO_AABB_MAX_SENT;
        int dg = 0;
        {
            // faces [0, Fe) use exactly the vertices [0, Ve); their bbox
            // extremes are those of the vertices
            int32_t xl = INT32_MAX, xh = INT32_MIN, yl = INT32_MAX, yh = INT32_MIN;
            auto take = [&](int v, float x, float y, float z) {
                int32_t X, Y;
                float Z;
                dg |= vertex_fixed(pose, x, y, z, cam, X, Y, Z) ? 1 : 0;
                if (cached) {
                    if (pack) vP[v] = ((uint32_t)X & 0xffffu) | ((uint32_t)Y << 16);  // used only if all fit
                    else { vXY[v] = make_int2(X, Y); vZ[v] = Z; }
                } else if (runs && v < 2 * kWin) {
                    // the packed window from vertex 0, kept if every coordinate fits
                    vP[v] = ((uint32_t)X & 0xffffu) | ((uint32_t)Y << 16);
                }
                xl = min(xl, X); xh = max(xh, X);
                yl = min(yl, Y); yh = max(yh, Y);
            };
#pragma unroll
            for (int j = 0; j < kVPT; ++j) {
                const int v = tid + j * BT;
                if (v < Ve) take(v, ux[j], uy[j], uz[j]);
            }
            if (WIN)
                for (int v = tid + kVPT * BT; v < Ve; v += BT) take(v, wx[v], wy[v], wz[v]);  // beyond the cache
            if (xl <= xh) {
                bx0 = min(bx0, wadd(xl, 15) >> 4); bx1 = max(bx1, wadd(xh, 15) >> 4);
                by0 = min(by0, wadd(yl, 15) >> 4); by1 = max(by1, wadd(yh, 15) >> 4);
            }
        }
        bx0 = wave_min(bx0); bx1 = wave_max(bx1);
        by0 = wave_min(by0); by1 = wave_max(by1);
        dg = wave_max(dg);
        if ((tid & 63) == 0) {
            atomicMin(&s_box[0], bx0); atomicMax(&s_box[1], bx1);
            atomicMin(&s_box[2], by0); atomicMax(&s_box[3], by1);
            if (dg) s_degen = 1;
        }
    }
    __syncthreads();

    const int32_t ax0 = s_box[0], ax1 = s_box[1], ay0 = s_box[2], ay1 = s_box[3];
    const int32_t minX = min(ax0, S.rect[0]), maxX = max(ax1, S.rect[2]);
    const int32_t minY = min(ay0, S.rect[1]), maxY = max(ay1, S.rect[3]);
    const long long W64 = (long long)maxX - minX, H64 = (long long)maxY - minY;
    const bool degen = s_degen != 0 || W64 <= 0 || H64 <= 0 || W64 * H64 > (long long)QPSO_AREA_LIMIT;

    if (S.dbg_aabb && tid == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) S.dbg_aabb[p * 4 + i] = s_box[i];
    }
    QG QpsoStale* __restrict__ nw = (QG QpsoStale*)S.stale[par ^ 1] + (size_t)p * cap;

    if (degen) {  // guard G: NaN fitness, buffer reset
        if (tid == 0) {
            S.fit[p] = __builtin_nanf("");
            S.stale_cnt[par ^ 1][p] = 0;
            atomicAdd(&S.counters[2], 1ull);
        }
        return;
    }

    // (X + 15) >> 4 >= -2047 gives X >= -32767 and <= 2046 gives X <= 32736:
    // every cached vertex fits the 16-bit halves.  A particle with vertices
    // beyond them writes the three-word cache instead (its vertices are still
    // in registers; the packed words are never read) and takes more bands;
    // with windows it projects them in three words.
    if (pack && !(ax0 >= -2047 && ax1 <= 2046 && ay0 >= -2047 && ay1 <= 2046)) {
        pack = false;
#pragma unroll
        for (int j = 0; j < kVPT; ++j) {
            const int v = tid + j * BT;
            if (cached && v < Ve) {
                int32_t X, Y;
                float Z;
                vertex_fixed(pose, ux[j], uy[j], uz[j], cam, X, Y, Z);
                vXY[v] = make_int2(X, Y);
                vZ[v] = Z;
            }
        }
    }
    // queue of small faces needing the per-pixel depth (kQCap entries); big
    // faces of more than S.big_face pixels (non-flat ones: half as many),
    // rasterized by a whole wave so one large face does not hold its wave's
    // other 63 lanes idle: clipped
    // rectangle (x0 | x1 << 16, y0 | y1 << 16), constant depth, face << 1 | flat
    if (WIN && runs) {  // window capacity in vertices
        VC = pack ? 2 * kWin : kWin;
        vZ = reinterpret_cast<float*>(lds + 2 * VC);
        if (pack) chi = min(Ve, 2 * kWin);  // pass A's window [0, chi)
    }
    int32_t* __restrict__ queue = lds + (pack ? VC : 3 * VC);
    int32_t* __restrict__ bq_x = queue + kQCap;
    int32_t* __restrict__ bq_y = bq_x + kBQCap;
    int32_t* __restrict__ bq_k = bq_y + kBQCap;
    int32_t* __restrict__ bq_i = bq_k + kBQCap;
    // per wave: ring of kRing culled face entries waiting to be drawn
    int32_t* __restrict__ rings = bq_i + kBQCap;
    // the tile starts on a 16-byte boundary for 128-bit fills
    const int32_t tile_off = (int32_t)(rings - lds) + (BT / 64) * kRing;
    int32_t* __restrict__ tile = lds + ((tile_off + 3) & ~3);
    const int32_t tcap = CAP - ((tile_off + 3) & ~3);

    const int32_t W = (int32_t)W64, H = (int32_t)H64;
    const int32_t A = W * H;
    const int32_t T = ((A + 1023) >> 10) << 10;  // NumIters * 1024 (main.cu:621-623)
    const Visits vis{W, H, T, rcp((float)W)};
    const int32_t rx0 = S.rect[0], ry0 = S.rect[1], rx1 = S.rect[2], ry1 = S.rect[3];
    const int32_t rw = rx1 - rx0;
    const QG uint16_t* __restrict__ obs = (const QG uint16_t*)S.obs;

    // carried pixels inside the union box join the tile region
    {
        int32_t sx0 = INT32_MAX, sx1 = INT32_MIN, sy0 = INT32_MAX, sy1 = INT32_MIN;
#pragma unroll
        for (int j = 0; j < kSPT; ++j) {
            if (tid + j * BT >= n_old) break;
            const QpsoStale e = carried[j];
            const int32_t ex = e.pos & 0xffff, ey = e.pos >> 16;
            if (ex >= minX && ex < maxX && ey >= minY && ey <= maxY) {
                sx0 = min(sx0, ex); sx1 = max(sx1, ex);
                sy0 = min(sy0, ey); sy1 = max(sy1, ey);
            }
        }
        if (n_old > 0) {
            sx0 = wave_min(sx0); sx1 = wave_max(sx1);
            sy0 = wave_min(sy0); sy1 = wave_max(sy1);
            if ((tid & 63) == 0 && sx0 <= sx1) {
                atomicMin(&s_stale[0], sx0); atomicMax(&s_stale[1], sx1);
                atomicMin(&s_stale[2], sy0); atomicMax(&s_stale[3], sy1);
            }
        }
    }
    __syncthreads();
    // Tile region: rendered AABB plus those carried pixels, inside the frame.
    // Every depth value other than 60000 in the union box lies in it.
    int32_t tx0 = ax0, tx1 = ax1, ty0 = ay0, ty1 = ay1;  // [x0, x1) x [y0, y1)
    if (s_stale[0] <= s_stale[1]) {
        tx0 = min(tx0, s_stale[0]); tx1 = max(tx1, s_stale[1] + 1);
        ty0 = min(ty0, s_stale[2]); ty1 = max(ty1, s_stale[3] + 1);
    }
    tx0 = max(tx0, 0); tx1 = min(tx1, kCols);
    ty0 = max(ty0, 0); ty1 = min(ty1, kRows);
    const bool has_tile = tx0 < tx1 && ty0 < ty1;
    const int32_t TW = has_tile ? tx1 - tx0 : 1;
    // Depth region D: the tile pixels whose depth value the fitness reads, the
    // tile inside the cluster rect.  Elsewhere only "rendered" matters (a bit
    // per pixel, the mask), except in the column x = minX and the row y = maxY,
    // where the scan leaves pixels unvisited and their depth carries to the
    // next iteration: those two lines keep exact depth in their own arrays.
    // A stage call that dumps the rendered frames keeps the depth of the whole
    // tile (D = tile, no mask).
    const bool full = FRAMES && S.dbg_frames != nullptr;
    const int32_t dx0 = full ? tx0 : max(tx0, rx0), dx1 = full ? tx1 : min(tx1, rx1);
    const int32_t dy0 = full ? ty0 : max(ty0, ry0), dy1 = full ? ty1 : min(ty1, ry1);
    const int32_t DW = dx0 < dx1 && dy0 < dy1 ? dx1 - dx0 : 0;
    const int32_t MW = full ? 0 : (TW + 31) >> 5;  // mask words per tile row
    const bool col_line = minX >= tx0 && minX < tx1;

    float acc_e[SPT];
#pragma unroll
    for (int j = 0; j < SPT; ++j) acc_e[j] = 0.0f;
    uint32_t n_seg = 0, n_tr = 0, n_rend = 0;
    unsigned long long px_r = 0;

    auto push_stale = [&](int32_t x, int32_t y, int32_t v) {
        const int k = atomicAdd(&s_nstale, 1);
        if (k < cap) store_stale(nw, k, (y << 16) | x, v);
        else atomicOr(S.flags, QPSO_FLAG_STALE_OVERFLOW);
    };
    // the face's 28.4 vertices: from the LDS vertex cache, or projected again.
    // f is the face's index, e its table entry (the host keeps the table in
    // face order, e.face == f; passing f lets the corner reads go out without
    // waiting for the entry)
    auto load_face = [&](int32_t f, const QpsoFaceRef& e, Face& fc) {
        if (pack) {
            const int32_t vb = WIN ? vbase : 0;
            const uint32_t pa = vP[e.v0 - vb], pb = vP[e.v1 - vb], pc = vP[e.v2 - vb];
            fc.X[0] = (int32_t)(pa << 16) >> 16; fc.Y[0] = (int32_t)pa >> 16;
            fc.X[1] = (int32_t)(pb << 16) >> 16; fc.Y[1] = (int32_t)pb >> 16;
            fc.X[2] = (int32_t)(pc << 16) >> 16; fc.Y[2] = (int32_t)pc >> 16;
            // corner depths: pass A's projection of the same coordinates,
            // read as one 16-byte entry per corner
            typedef float f4 __attribute__((ext_vector_type(4)));
            const QG f4* __restrict__ cr = (const QG f4*)S.corners + 3 * (size_t)f;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const f4 c = cr[j];
                float u, v;
                project(pose, c.x, c.y, c.z, cam, u, v, fc.Z[j]);
            }
            face_bbox(fc);
        } else if (!WIN || VC > 0) {  // always cached without windows
            const int32_t vb = WIN ? vbase : 0;
            const int32_t a = e.v0 - vb, b = e.v1 - vb, c = e.v2 - vb;
            const int2 pa = vXY[a], pb = vXY[b], pc = vXY[c];
            fc.X[0] = pa.x; fc.Y[0] = pa.y; fc.Z[0] = vZ[a];
            fc.X[1] = pb.x; fc.Y[1] = pb.y; fc.Z[1] = vZ[b];
            fc.X[2] = pc.x; fc.Y[2] = pc.y; fc.Z[2] = vZ[c];
            face_bbox(fc);
        } else {
            face_project(S.corners, e.face, pose, cam, fc);
        }
    };
    // a face to draw: with the packed cache its corner entries carry the
    // vertex indices (w), so the face table entry is not read
    auto draw_face = [&](int32_t f, Face& fc) {
        if (pack) {
            typedef float f4 __attribute__((ext_vector_type(4)));
            const QG f4* __restrict__ cr = (const QG f4*)S.corners + 3 * (size_t)f;
            const int32_t vb = WIN ? vbase : 0;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const f4 c = cr[j];
                // the component goes through a scalar first: this compiler
                // lowers __builtin_bit_cast of a vector component to the
                // bits of component 0
                const float wf = c.w;
                const uint32_t w = vP[__float_as_int(wf) - vb];
                fc.X[j] = (int32_t)(w << 16) >> 16;
                fc.Y[j] = (int32_t)w >> 16;
                float u, v;
                project(pose, c.x, c.y, c.z, cam, u, v, fc.Z[j]);
            }
            face_bbox(fc);
        } else {
            load_face(f, load_ref(faces, f), fc);
        }
    };
    // a queued face, drawn after its window (windowed meshes: projected from
    // the corner table, the same arithmetic as the cache's, so the same bits)
    auto queued_face = [&](int32_t f, Face& fc) {
        if (WIN && runs) face_project(S.corners, f, pose, cam, fc);
        else draw_face(f, fc);
    };
    // per-pixel depth over [x0, x1) x [y0, y1): edge values at (x0, y0), per
    // pixel CX -= FDY, from a row's last pixel to the next row's first
    // CX += FDX + (w - 1) FDY
    int32_t db0 = 0;  // first D row of the current band
    auto raster_exact = [&](const Face& fc, const Setup& su, int32_t x0, int32_t x1, int32_t y0, int32_t y1) {
        if (x0 >= x1 || y0 >= y1) return;
        const int32_t w = x1 - x0, n = __mul24(w, y1 - y0);
        int32_t CX2 = wsub(wadd(su.CY2, wmul(y0 - fc.miny, su.FDX23)), wmul(x0 - fc.minx, su.FDY23));
        int32_t CX3 = wsub(wadd(su.CY3, wmul(y0 - fc.miny, su.FDX31)), wmul(x0 - fc.minx, su.FDY31));
        const int32_t step2 = wsub(0, su.FDY23), step3 = wsub(0, su.FDY31);
        const int32_t wrap2 = wadd(su.FDX23, wmul(w - 1, su.FDY23));
        const int32_t wrap3 = wadd(su.FDX31, wmul(w - 1, su.FDY31));
        const int32_t awrap = DW - w + 1;
        int32_t* addr = tile + __mul24(y0 - db0, DW) + (x0 - dx0);
        int32_t c = w;
        for (int32_t m = 0; m < n; ++m) {
            atomicMin(addr, pixel_depth(CX2, CX3, su));
            const bool wrap = --c == 0;
            CX2 = wadd(CX2, wrap ? wrap2 : step2);
            CX3 = wadd(CX3, wrap ? wrap3 : step3);
            addr += wrap ? awrap : 1;
            c = wrap ? w : c;
        }
    };
    // per-pixel depth over [x0, x1) x [y0, y1) by K lanes together: lane j
    // starts at pixel j of the rectangle in row-major order and every step
    // moves K pixels on (sy = K / w rows and sx = K % w columns, one wrap at
    // most); the minima do not depend on which lane writes which pixel
    auto raster_lanes = [&](const Face& fc, const Setup& su, int32_t x0, int32_t x1, int32_t y0, int32_t y1, int32_t j,
                            int32_t K) {
        if (x0 >= x1 || y0 >= y1) return;
        const int32_t w = x1 - x0, n = __mul24(w, y1 - y0);
        const int32_t sy = K / w, sx = K - __mul24(sy, w);
        int32_t dy = j / w, dx = j - __mul24(dy, w);
        int32_t CX2 = wsub(wadd(su.CY2, wmul(y0 + dy - fc.miny, su.FDX23)), wmul(x0 + dx - fc.minx, su.FDY23));
        int32_t CX3 = wsub(wadd(su.CY3, wmul(y0 + dy - fc.miny, su.FDX31)), wmul(x0 + dx - fc.minx, su.FDY31));
        const int32_t step2 = wsub(wmul(sy, su.FDX23), wmul(sx, su.FDY23));
        const int32_t step3 = wsub(wmul(sy, su.FDX31), wmul(sx, su.FDY31));
        const int32_t fix2 = wadd(su.FDX23, wmul(w, su.FDY23));  // x -= w, y += 1
        const int32_t fix3 = wadd(su.FDX31, wmul(w, su.FDY31));
        const int32_t astep = __mul24(sy, DW) + sx, afix = DW - w;
        int32_t off = __mul24(y0 + dy - db0, DW) + (x0 + dx - dx0);
        for (int32_t m = j; m < n; m += K) {
            atomicMin(tile + off, pixel_depth(CX2, CX3, su));
            dx += sx;
            CX2 = wadd(CX2, step2);
            CX3 = wadd(CX3, step3);
            off += astep;
            if (dx >= w) {
                dx -= w;
                CX2 = wadd(CX2, fix2);
                CX3 = wadd(CX3, fix3);
                off += afix;
            }
        }
    };
    // a queued big face walked by the 64 lanes of a wave: lane l starts at
    // pixel l of the rectangle in row-major order and every step moves 64
    // pixels on, i.e. sy = 64 / w rows and sx = 64 % w columns, with one wrap
    // at most.  floor(a / w) for a <= 64, w <= 640 is trunc((a + 0.5) * rcp(w)):
    // (a + 0.5) / w lies at least 1 / 1280 from an integer, far beyond the
    // error of the approximate reciprocal.
    auto raster_wave = [&](int32_t q) {
        const int32_t x0 = bq_x[q] & 0xffff, x1 = bq_x[q] >> 16;
        const int32_t y0 = bq_y[q] & 0xffff, y1 = bq_y[q] >> 16;
        const int32_t kd = bq_k[q], fi = bq_i[q];
        const int32_t w = x1 - x0, n = __mul24(w, y1 - y0);
        const int32_t lane = tid & 63;
        const float rw = __builtin_amdgcn_rcpf((float)w);
        const int32_t sy = (int32_t)(64.5f * rw), sx = 64 - __mul24(sy, w);
        int32_t dy = (int32_t)(((float)lane + 0.5f) * rw), dx = lane - __mul24(dy, w);
        const int32_t astep = __mul24(sy, DW) + sx, afix = DW - w;
        int32_t off = __mul24(y0 + dy - db0, DW) + (x0 + dx - dx0);
        if (fi & 1) {  // flat: one depth over the rectangle
            for (int32_t m = lane; m < n; m += 64) {
                atomicMin(tile + off, kd);
                dx += sx;
                off += astep;
                if (dx >= w) { dx -= w; off += afix; }
            }
            return;
        }
        Face fc;
        Setup su;
        queued_face(fi >> 1, fc);
        face_setup(fc, su);
        int32_t CX2 = wsub(wadd(su.CY2, wmul(y0 + dy - fc.miny, su.FDX23)), wmul(x0 + dx - fc.minx, su.FDY23));
        int32_t CX3 = wsub(wadd(su.CY3, wmul(y0 + dy - fc.miny, su.FDX31)), wmul(x0 + dx - fc.minx, su.FDY31));
        const int32_t step2 = wsub(wmul(sy, su.FDX23), wmul(sx, su.FDY23));
        const int32_t step3 = wsub(wmul(sy, su.FDX31), wmul(sx, su.FDY31));
        const int32_t fix2 = wadd(su.FDX23, wmul(w, su.FDY23));  // x -= w, y += 1
        const int32_t fix3 = wadd(su.FDX31, wmul(w, su.FDY31));
        for (int32_t m = lane; m < n; m += 64) {
            atomicMin(tile + off, pixel_depth(CX2, CX3, su));
            dx += sx;
            CX2 = wadd(CX2, step2);
            CX3 = wadd(CX3, step3);
            off += astep;
            if (dx >= w) {
                dx -= w;
                CX2 = wadd(CX2, fix2);
                CX3 = wadd(CX3, fix3);
                off += afix;
            }
        }
    };

    // ---- cluster pixels outside the tile region: depth buffer reads 60000
    //      there, so only "observed but not rendered" counts (order-free).
    //      The rect minus the tile is at most four strips: rows above and
    //      below the tile, and left and right of it within its rows.  Rect
    //      rows lie above the union box's last row, so every pixel there is
    //      visited once except those of the column x = minX (vis.at(0, k)
    //      times): a strip row adds its observed pixels, counted from the
    //      crop's per-row prefix table, and corrects that column.
    {
        const QG uint16_t* __restrict__ obs_rows = (const QG uint16_t*)S.obs_rows;
        auto strip = [&](int32_t x0, int32_t x1, int32_t y0, int32_t y1) {
            if (x0 >= x1 || y0 >= y1) return;
            const bool col = minX >= x0 && minX < x1;
            for (int32_t y = y0 + tid; y < y1; y += BT) {
                const QG uint16_t* __restrict__ pr = obs_rows + __mul24(y - ry0, rw + 1);
                uint32_t c = (uint32_t)pr[x1 - rx0] - (uint32_t)pr[x0 - rx0];
                if (col && obs[__mul24(y - ry0, rw) + (minX - rx0)] != 0) c = c - 1 + (uint32_t)vis.at(0, y - minY);
                n_seg += c;
            }
        };
        if (has_tile) {
            const int32_t my0 = max(ry0, ty0), my1 = min(ry1, ty1);
            strip(rx0, rx1, ry0, min(ry1, ty0));
            strip(rx0, rx1, max(ry0, ty1), ry1);
            strip(rx0, min(rx1, tx0), my0, my1);
            strip(max(rx0, tx1), rx1, my0, my1);
        } else {
            strip(rx0, rx1, ry0, ry1);
        }
    }

    // Band rows [b0, b1): as many tile rows as fit, a row costing its mask
    // words and its column-line word, plus DW for the rows of D; the row line
    // (TW words) and four words of alignment come once per band.
    auto band_end = [&](int32_t b0) {
        const int32_t cr = MW + 1;
        int32_t budget = tcap - TW - 8, y = b0;
        auto take = [&](int32_t y_end, int32_t cost) {
            if (y < y_end) {
                const int32_t n = min(y_end - y, budget / cost);
                y += n;
                budget -= n * cost;
            }
            return y >= y_end;
        };
        if (take(min(dy0, ty1), cr) && take(DW > 0 ? dy1 : y, cr + DW)) take(ty1, cr);
        return max(y, b0 + 1);  // one D row fits: TW <= 640 and tcap >= 3909
    };

    uint32_t* mask = nullptr;    // [band rows][MW]
    int32_t* col_d = nullptr;    // [band rows]: depth of (minX, y)
    int32_t* row_d = nullptr;    // [TW]: depth of (x, maxY)
    uint64_t fmask = ~0ull;  // bit k: face entry tid + k*BT still reaches later bands (meshes without windows)
    const bool use_mask = !runs && n_faces <= 64 * BT;
    // windows: bit w, window w of a band (the same windows every band) still
    // has faces reaching later bands; a window without is neither projected
    // nor culled again.  wslot: the windows drawn so far mod 3, the slot of
    // s_later the current window uses.
    uint64_t wmask = ~0ull;
    int wslot = 0;
    for (int32_t b0 = ty0, b1; has_tile && b0 < ty1; b0 = b1) {
        b1 = band_end(b0);  // band rows [b0, b1)
        const int32_t nb = b1 - b0;
        db0 = max(b0, dy0);
        const int32_t db1 = min(b1, dy1);
        const int32_t nd = DW > 0 && db0 < db1 ? (db1 - db0) * DW : 0;  // depth words
        mask = reinterpret_cast<uint32_t*>(tile + ((nd + 3) & ~3));
        col_d = reinterpret_cast<int32_t*>(mask) + nb * MW;
        row_d = col_d + nb;
        const bool row_line = maxY >= b0 && maxY < b1;
        {
            typedef int32_t v4i __attribute__((ext_vector_type(4)));
            v4i* t4 = reinterpret_cast<v4i*>(tile);
            const v4i e4 = {QPSO_EMPTY, QPSO_EMPTY, QPSO_EMPTY, QPSO_EMPTY};
            for (int i = tid; i < nd / 4; i += BT) t4[i] = e4;
            if (tid < (nd & 3)) tile[(nd & ~3) + tid] = QPSO_EMPTY;
            for (int i = tid; i < nb * MW; i += BT) mask[i] = 0u;
            for (int i = tid; i < nb + TW; i += BT) col_d[i] = QPSO_EMPTY;
        }
        __syncthreads();
        // pixels carried over from earlier iterations (their values are below 60000)
#pragma unroll
        for (int j = 0; j < kSPT; ++j) {
            if (tid + j * BT >= n_old) break;
            const QpsoStale e = carried[j];
            const int32_t ex = e.pos & 0xffff, ey = e.pos >> 16;
            if (ey >= b0 && ey < b1 && ex >= tx0 && ex < tx1) {
                if (ex >= dx0 && ex < dx1 && ey >= db0 && ey < db1) tile[(ey - db0) * DW + ex - dx0] = e.val;
                if (MW) atomicOr(&mask[(ey - b0) * MW + ((ex - tx0) >> 5)], 1u << ((ex - tx0) & 31));
                if (ex == minX) col_d[ey - b0] = e.val;
                if (ey == maxY) row_d[ex - tx0] = e.val;
            }
        }
        __syncthreads();

        // bits of the mask row y over columns [x0, x1)
        auto mark_rows = [&](int32_t x0, int32_t x1, int32_t y0, int32_t y1) {
            const int32_t w0 = (x0 - tx0) >> 5, w1 = (x1 - 1 - tx0) >> 5;
            const uint32_t m0 = ~0u << ((x0 - tx0) & 31), m1 = ~0u >> (31 - ((x1 - 1 - tx0) & 31));
            for (int32_t y = y0; y < y1; ++y) {
                uint32_t* row = mask + __mul24(y - b0, MW);
                for (int32_t w = w0; w <= w1; ++w) {
                    const uint32_t m = (w == w0 ? m0 : ~0u) & (w == w1 ? m1 : ~0u);
                    atomicOr(&row[w], m);
                }
            }
        };

        // ---- raster: bbox fill (main.h:1355-1488) in two steps per wave.
        //      Cull: the wave scans 64 face entries at a time and appends those
        //      that are front-facing and reach the band and the tile to its
        //      ring in LDS.  Draw: whenever 64 are waiting, one per lane, the
        //      lane writes its face's clipped bbox C: the part in D with
        //      ds_min_i32, as one flat pixel loop (a nested row/column loop
        //      would diverge as max rows x max columns); the rest of C into the
        //      mask where its depth is below 60000; and the column and row
        //      lines.  Faces whose depth over C is provably one integer
        //      (depth_bounds) store that constant; the others are queued
        //      and rasterized with the per-pixel depth after the pass; D parts
        //      of more than big_face pixels (non-flat: big_face / 2) go to the
        //      wave queue.
        auto draw = [&](int32_t i) {
            Face fc;
            Setup su;
            draw_face(i, fc);
            face_setup(fc, su);
            const int32_t x0 = max(fc.minx, tx0), x1 = min(fc.maxx, tx1);
            const int32_t y0 = max(fc.miny, b0), y1 = min(fc.maxy, b1);
            px_r += (unsigned long long)((x1 - x0) * (y1 - y0));
            // part in D
            const int32_t px0 = max(x0, dx0), px1 = min(x1, dx1), py0 = max(y0, db0), py1 = min(y1, db1);
            const bool in_d = px0 < px1 && py0 < py1;
            const bool whole = in_d && px0 == x0 && px1 == x1 && py0 == y0 && py1 == y1;
            // bounds over all of C (for the mask and the lines); one depth over
            // C is one over its D part (a D part of a face not flat over C
            // takes the per-pixel path, exact either way)
            int32_t kmin = 0, kmax = 0;
            const bool bounded = depth_bounds(fc, su, x0, x1, y0, y1, kmin, kmax);
            const bool flat = in_d && bounded && kmin == kmax;
            const int32_t kd = kmin;
            if (MW && !whole) {
                if (bounded && kmax < QPSO_EMPTY) {
                    mark_rows(x0, x1, y0, y1);
                } else if (!bounded || kmin < QPSO_EMPTY) {
                    for (int32_t y = y0; y < y1; ++y)
                        for (int32_t x = x0; x < x1; ++x)
                            if (!(x >= dx0 && x < dx1 && y >= db0 && y < db1) && depth_at(fc, su, x, y) < QPSO_EMPTY)
                                atomicOr(&mask[(y - b0) * MW + ((x - tx0) >> 5)], 1u << ((x - tx0) & 31));
                }
            }
            const bool one = bounded && kmin == kmax;
            if (col_line && x0 == minX) {
                for (int32_t y = y0; y < y1; ++y) atomicMin(&col_d[y - b0], one ? kmin : depth_at(fc, su, minX, y));
            }
            if (row_line && y0 <= maxY && maxY < y1) {
                for (int32_t x = x0; x < x1; ++x) atomicMin(&row_d[x - tx0], one ? kmin : depth_at(fc, su, x, maxY));
            }
            if (!in_d) return;
            const int32_t w = px1 - px0, n = __mul24(w, py1 - py0);
            int32_t qb;
            // a non-flat face costs a per-pixel loop in its lane, so it goes to
            // the wave queue from half the threshold of a flat one; with that
            // queue full, to the small-face queue (drawn by several lanes,
            // per pixel: a flat face's pixels all get kd there too)
            const bool big = (flat ? n : 2 * n) > big_face;
            if (big && (qb = atomicAdd(&s_nqb, 1)) < kBQCap) {
                bq_x[qb] = px0 | (px1 << 16);
                bq_y[qb] = py0 | (py1 << 16);
                bq_k[qb] = kd;
                bq_i[qb] = (i << 1) | (flat ? 1 : 0);
            } else if (flat && !big) {
                // rows in chunks of four pixels; a chunk's positions past the
                // row's end repeat its last pixel (the same minimum again)
                int32_t* row = tile + __mul24(py0 - db0, DW) + (px0 - dx0);
                const int32_t last = w - 1;
                int32_t c = 0, rows = py1 - py0;
                while (rows > 0) {
                    atomicMin(row + c, kd);
                    atomicMin(row + min(c + 1, last), kd);
                    atomicMin(row + min(c + 2, last), kd);
                    atomicMin(row + min(c + 3, last), kd);
                    c += 4;
                    const bool next = c > last;
                    row += next ? DW : 0;
                    rows -= next ? 1 : 0;
                    c = next ? 0 : c;
                }
            } else {
                const int q = atomicAdd(&s_nq, 1);
                if (q < kQCap) queue[q] = i;
                else raster_exact(fc, su, px0, px1, py0, py1);
            }
        };
        int32_t* __restrict__ ring = rings + (tid >> 6) * kRing;
        const int lane = tid & 63;
        const uint64_t below = (1ull << lane) - 1ull;
        int run = 0;  // next run of faces (windowed meshes)
        int wi = -1;  // window index in this band
        for (;;) {
        int f_lo = 0, f_hi = n_faces;
        if (runs) {
            // the runs whose joint vertex window fits the cache; project it
            const int32_t* __restrict__ rt = S.runs;
            int32_t vlo = rt[3 * run + 1], vhi = rt[3 * run + 2];
            f_lo = rt[3 * run];
            int r2 = run + 1;
            while (r2 < S.n_runs && rt[3 * r2] < Fe && max(vhi, rt[3 * r2 + 2]) - min(vlo, rt[3 * r2 + 1]) <= VC) {
                vlo = min(vlo, rt[3 * r2 + 1]);
                vhi = max(vhi, rt[3 * r2 + 2]);
                ++r2;
            }
            f_hi = r2 < S.n_runs ? rt[3 * r2] : n_faces;
            run = r2;
            vbase = vlo;
            ++wi;
            if (wi < 64 && !((wmask >> wi) & 1ull)) {  // uniform: no face of it is left
                if (run >= S.n_runs || rt[3 * run] >= Fe) break;
                continue;
            }
            if (vlo >= cbase && vhi <= chi) {
                vbase = cbase;  // projected already (pass A, or the last window drawn)
            } else {
                const QG float* __restrict__ wx = (const QG float*)S.verts;
                const QG float* __restrict__ wy = wx + S.V;
                const QG float* __restrict__ wz = wy + S.V;
                for (int v = vlo + tid; v < vhi; v += BT) {
                    int32_t X, Y;
                    float Z;
                    vertex_fixed(pose, wx[v], wy[v], wz[v], cam, X, Y, Z);
                    if (pack) vP[v - vlo] = ((uint32_t)X & 0xffffu) | ((uint32_t)Y << 16);
                    else { vXY[v - vlo] = make_int2(X, Y); vZ[v - vlo] = Z; }
                }
                cbase = vlo;
                chi = vhi;
                __syncthreads();
            }
        }
        int head = 0, cnt = 0;  // wave-uniform ring state
        int k = 0;
        QpsoFaceRef nxt{0, 0, 0, INT32_MAX};
        if (f_lo + tid < f_hi) nxt = load_ref(faces, f_lo + tid);
        if (runs) fmask = 0ull;  // windows: non-zero once a face of the window reaches later bands
        for (int base = f_lo + tid - lane; base < f_hi; base += BT, ++k) {
            const int i = base + lane;  // = f_lo + tid + k * BT
            const QpsoFaceRef e = nxt;  // prefetched one iteration ahead
            if (i + BT < f_hi) nxt = load_ref(faces, i + BT);
            bool keep = false;
            if (i < f_hi && (!use_mask || ((fmask >> k) & 1ull))) {
                bool later = false;
                if (e.face < Fe) {
                    Face fc;
                    load_face(i, e, fc);
                    if (max(fc.minx, tx0) < min(fc.maxx, tx1) && fc.miny < ty1 && fc.maxy > b0 && face_front(fc)) {
                        later = fc.maxy > b1;
                        keep = fc.miny < b1;
                    }
                }
                if (use_mask && !later) fmask &= ~(1ull << k);
                if (runs && later) fmask = 1ull;
            }
            const uint64_t m = __ballot(keep);
            if (keep) ring[(head + cnt + __popcll(m & below)) & (kRing - 1)] = i;
            cnt += __popcll(m);
            __builtin_amdgcn_wave_barrier();
            if (cnt >= 64) {
                draw(ring[(head + lane) & (kRing - 1)]);
                head = (head + 64) & (kRing - 1);
                cnt -= 64;
                __builtin_amdgcn_wave_barrier();
            }
        }
        if (lane < cnt) draw(ring[(head + lane) & (kRing - 1)]);
        if (!runs) break;  // one pass over the faces; the queues follow
        // windows: the queued faces wait for the band's end (they are
        // projected again from the corner table there); every wave is done
        // with this window's cache before the next one is projected.  The
        // window's flag goes to slot wslot, read after the barrier; the next
        // slot (last read two windows ago, before the previous barrier) is
        // cleared before it.
        if (__ballot(fmask != 0ull) != 0ull && lane == 0) s_later[wslot] = 1;
        if (tid == 0) s_later[wslot == 2 ? 0 : wslot + 1] = 0;
        __syncthreads();
        if (wi < 64 && s_later[wslot] == 0) wmask &= ~(1ull << wi);
        wslot = wslot == 2 ? 0 : wslot + 1;
        if (run >= S.n_runs || S.runs[3 * run] >= Fe) break;
        }
        __syncthreads();
        {
        // queued small faces: per-pixel depth over their D part, K lanes per
        // face (a power of two, as many as the threads allow, at most 64),
        // lane j of a face taking its pixels j, j + K, j + 2K, ...
        const int nq = min(s_nq, kQCap);
        if (nq > 0) {
            const int share = BT / nq;
            const int K = share >= 64 ? 64 : share >= 1 ? 1 << (31 - __builtin_clz(share)) : 1;
            const int j = tid & (K - 1);
            for (int q = tid / K; q < nq; q += BT / K) {
                Face fc;
                Setup su;
                queued_face(queue[q], fc);
                face_setup(fc, su);
                raster_lanes(fc, su, max(fc.minx, dx0), min(fc.maxx, dx1), max(fc.miny, db0), min(fc.maxy, db1), j, K);
            }
        }
        // queued big faces: one face per wave, taken in turn
        const int nqb = min(s_nqb, kBQCap);
        for (;;) {
            int q = 0;
            if ((tid & 63) == 0) q = atomicAdd(&s_take, 1);
            q = __shfl(q, 0);
            if (q >= nqb) break;
            raster_wave(q);
        }
        __syncthreads();
        if (tid == 0) { s_nq = 0; s_nqb = 0; s_take = 0; }
        }

        if (full) {  // D is the whole tile here
            int32_t* fr = S.dbg_frames + (size_t)p * kRows * kCols;
            for (int i = tid; i < nd; i += BT) fr[(b0 + i / TW) * kCols + tx0 + i % TW] = tile[i];
        }

        // ---- counters of the band's tile pixels outside the cluster rect,
        //      weighted by visits (order-free): nothing is observed there, so
        //      a rendered pixel counts in n_rendered and n_tr.  Pixels of D
        //      inside the rect are counted by the error scan below, once per
        //      visit.  The rect contains D unless D is the whole tile.
        if (full) {
            const int32_t dq = BT / TW, dr = BT % TW;
            int32_t cy = b0 + tid / TW, cx = tx0 + tid % TW;
            for (int i = tid; i < nd; i += BT, cx += dr, cy += dq) {
                if (cx >= tx1) { cx -= TW; ++cy; }
                const int32_t y = cy, x = cx;
                if (tile[i] == QPSO_EMPTY || (x >= rx0 && x < rx1 && y >= ry0 && y < ry1)) continue;
                const uint32_t w = (uint32_t)vis.at(x - minX, y - minY);
                n_rend += w;
                n_tr += w;
            }
        } else {
            // the mask outside D, one row per thread: every pixel but (minX, y)
            // is visited once, except in the row maxY, where only offsets
            // [1, T - H W) are; (minX, y) is visited vis.at(0, k) times
            for (int32_t y = b0 + tid; y < b1; y += BT) {
                const int32_t k = y - minY;
                const uint32_t* row = mask + (y - b0) * MW;
                int32_t ca = tx0, cb = tx1;  // counted columns [ca, cb) besides (minX, y)
                if (k == H) cb = min(cb, minX + T - H * W);
                ca = max(ca, minX + 1);
                const bool row_d_in = y >= db0 && y < db1;
                uint32_t c = 0;
                for (int32_t w = 0; w < MW && ca < cb; ++w) {
                    const int32_t xw = tx0 + 32 * w;  // column of bit 0
                    auto span = [&](int32_t a, int32_t b2) -> uint32_t {  // bits for columns [a, b2)
                        const int32_t lo = max(a - xw, 0), hi = min(b2 - xw, 32);
                        if (lo >= hi) return 0u;
                        return (hi - lo == 32) ? ~0u : (((1u << (hi - lo)) - 1u) << lo);
                    };
                    uint32_t m = row[w] & span(ca, cb);
                    if (row_d_in) m &= ~span(dx0, dx1);
                    c += (uint32_t)__builtin_popcount(m);
                }
                if (col_line && !(row_d_in && minX >= dx0 && minX < dx1)) {
                    const int32_t c0 = minX - tx0;
                    if ((row[c0 >> 5] >> (c0 & 31)) & 1u) c += (uint32_t)vis.at(0, k);
                }
                n_rend += c;
                n_tr += c;
            }
        }

        // ---- squared error: the only order-sensitive sum.  Slot t & 2047
        //      accumulates in increasing t (main.cu:693-717); contributions come
        //      only from D pixels inside the cluster rect, so scan the t whose
        //      row lies in this band and whose pixel lies in that intersection.
        //      Each t is one visit, so the counters of those pixels are summed
        //      here too.
        const int32_t qx0 = max(dx0, rx0), qx1 = min(dx1, rx1);
        const int32_t qy0 = max(db0, ry0), qy1 = min(db1, ry1);
        if (DW > 0 && qx0 < qx1 && qy0 < qy1) {
            const int32_t k0 = qy0 - minY, k1 = qy1 - minY;
            const int32_t tlo = min(k0 * W + (qx0 - minX), T);
            const int32_t thi = min(k1 * W + 1, T);
            // U steps of 2048 per batch: all their loads go out before the
            // first sum (a pixel outside Q reads as empty and observed 0,
            // which adds nothing), then each slot sums in increasing t
            constexpr int U = SPT >= 8 ? 1 : 8 / SPT;
            // Q in union-box offsets (x, y), and the tile and crop indices as
            // y * pitch + x + constant
            const int32_t qxa = qx0 - minX, qxb = qx1 - minX, qya = qy0 - minY, qyb = qy1 - minY;
            // (minY - db0 can reach -2^23 - 480, beyond __mul24: wrapping 32-bit
            // products, exact modulo 2^32, as is each index they add up to)
            const int32_t tc = (minY - db0) * DW + (minX - dx0), oc = (minY - ry0) * rw + (minX - rx0);
            // Thread tid walks t = t0 + m BT (m = 0, 1, ...; slot j = m mod SPT)
            // and keeps the true quotient and remainder y, x of t by W from
            // step to step (BT = sq W + sr).  The fp32 row index differs from
            // the true one only at t = yW (x = 0): for t below 2^23 the
            // product t rcp(W) carries a relative error under 2^-23, less
            // than x / t for 0 < x < W.  There it is y - 1 when row y is
            // "short", and t visits (0, y - 1); column 0 of the union box lies
            // in Q only when qxa = 0.  A t inside Q in (x, y) is at least tlo
            // and below k1 W + 1, so of [tlo, thi) only t < T is left to test,
            // and only in the last batch: every t of an earlier one is below
            // its end, at most thi.  The four forms (column 0 in Q or not,
            // last batch or not) are compiled apart.
            const uint32_t xn = (uint32_t)(qxb - qxa), yn = (uint32_t)(qyb - qya);
            const int32_t sq = BT / W, sr = BT - sq * W;
            const bool col0 = qxa == 0;
            // the crop addressed by a 32-bit byte offset from its base
            const QG char* __restrict__ ob = (const QG char*)obs;
            int32_t y, x;
            {
                const int32_t t = (tlo / 2048) * 2048 + tid;
                y = row_of(t, vis.rW);
                x = t - y * W;
                if (x >= W) { x -= W; ++y; }
            }
            // the three counters per wave, from ballots (scalar adds)
            uint32_t w_rend = 0, w_tr = 0, w_seg = 0;
            auto batch = [&](int32_t base, auto c0, auto last) {
                constexpr bool kCol0 = decltype(c0)::value, kLast = decltype(last)::value;
                int32_t zr[U][SPT], zk[U][SPT];
#pragma unroll
                for (int u = 0; u < U; ++u) {
#pragma unroll
                    for (int j = 0; j < SPT; ++j) {
                        // t = base + u 2048 + j BT + tid = y W + x, slot j BT +
                        // tid; for t < thi <= T <= 2^23, y < 2^23 fits the
                        // 24-bit products
                        int32_t yv = y;
                        if (kCol0) {
                            // lanes hold consecutive t: a wave meets x = 0 once
                            // per W values of t, so the fp32 row is formed only then
                            const bool z = x == 0;
                            if (__ballot(z) != 0ull) {
                                // kept as a branch (not evaluated for every t)
                                __asm__ volatile("" ::: "memory");
                                if (z && row_of(y * W, vis.rW) != y) yv = y - 1;
                            }
                        }
                        zr[u][j] = QPSO_EMPTY;
                        zk[u][j] = 0;
                        if ((!kLast || (uint32_t)(base + u * 2048 + j * BT + tid) < (uint32_t)thi) &&
                            (uint32_t)(x - qxa) < xn && (uint32_t)(yv - qya) < yn) {
                            zr[u][j] = tile[__mul24(yv, DW) + x + tc];
                            const uint32_t oi = (uint32_t)(__mul24(yv, rw) + x + oc);  // >= 0: a rect pixel
                            zk[u][j] = (int32_t)*(const QG uint16_t*)(ob + 2u * oi);
                        }
                        x += sr;
                        y += sq;
                        if (x >= W) { x -= W; ++y; }
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
#pragma unroll
                    for (int j = 0; j < SPT; ++j) {
                        // an empty or unobserved pixel adds fma(0, 0, acc) = acc
                        // (acc >= +0)
                        const bool r = zr[u][j] != QPSO_EMPTY, o = zk[u][j] != 0;
                        const uint64_t mr = __ballot(r), mo = __ballot(o);
                        w_rend += (uint32_t)__popcll(mr);
                        w_tr += (uint32_t)__popcll(mr & ~mo);
                        w_seg += (uint32_t)__popcll(mo & ~mr);
                        const float e = (float)((r && o) ? wsub(zr[u][j], zk[u][j]) : 0) * 0.001f;
                        acc_e[j] = __builtin_fmaf(e, e, acc_e[j]);
                    }
                }
            };
            using yes = std::true_type;
            using no = std::false_type;
            for (int32_t base = (tlo / 2048) * 2048; base < thi; base += 2048 * U) {
                const bool last = base + 2048 * U > thi;
                if (col0) {
                    if (last) batch(base, yes{}, yes{});
                    else batch(base, yes{}, no{});
                } else {
                    if (last) batch(base, no{}, yes{});
                    else batch(base, no{}, no{});
                }
            }
            if ((tid & 63) == 0) {
                n_rend += w_rend;
                n_tr += w_tr;
                n_seg += w_seg;
            }
        }

        // ---- tile pixels the scan never visits keep their depth
        if (col_line) {  // column x-offset 0
            for (int32_t py = b0 + tid; py < b1; py += BT) {
                if (py > maxY) continue;
                if (vis.at(0, py - minY) == 0) {
                    const int32_t v = col_d[py - b0];
                    if (v != QPSO_EMPTY) push_stale(minX, py, v);
                }
            }
        }
        if (row_line) {  // last row, tail beyond t = T
            const int32_t xs = max(max(1, T - H * W), tx0 - minX), xe = min(W, tx1 - minX);
            for (int32_t xb = xs + tid; xb < xe; xb += BT) {
                const int32_t v = row_d[minX + xb - tx0];
                if (v != QPSO_EMPTY) push_stale(minX + xb, maxY, v);
            }
        }
        // the next band refills the tile; after the last one nothing reads
        // or writes it
        if (b1 < ty1) __syncthreads();
    }

    // carried pixels outside this iteration's union box stay as they are
#pragma unroll
    for (int j = 0; j < kSPT; ++j) {
        if (tid + j * BT >= n_old) break;
        const QpsoStale e = carried[j];
        const int32_t ex = e.pos & 0xffff, ey = e.pos >> 16;
        if (!(ex >= minX && ex < maxX && ey >= minY && ey <= maxY)) push_stale(ex, ey, e.val);
    }

    // ---- in-frame visits of the scan (work counter P_f), one row per thread
    unsigned long long px_f = 0;
    {
        const int32_t fx0 = max(minX, 0), fx1 = min(maxX, kCols);
        const int32_t fy0 = max(minY, 0), fy1 = min(maxY, kRows - 1);
        const bool col0 = minX >= 0;
        for (int32_t py = fy0 + tid; py <= fy1; py += BT) {
            const int32_t k = py - minY;
            int32_t n;
            if (k < H) n = fx1 - fx0 - (col0 ? 1 : 0);
            else n = max(0, min(fx1 - minX, T - H * W) - max(fx0 - minX, 1));
            if (col0) n += vis.at(0, k);
            px_f += (unsigned long long)n;
        }
    }

    // ---- reductions (FitnessReductionBallot, main.h:1522-1573): halving tree
    //      over the 2048 error slots; thread tid holds slots tid + j*BT
#pragma unroll
    for (int stride = 1024; stride >= BT; stride >>= 1) {
#pragma unroll
        for (int j = 0; j < stride / BT; ++j) acc_e[j] = acc_e[j] + acc_e[j + stride / BT];
    }
    const uint32_t cs = wave_sum(n_seg), ct = wave_sum(n_tr), cr = wave_sum(n_rend);
    if ((tid & 63) == 0) {
        atomicAdd(&s_cnt[0], cs); atomicAdd(&s_cnt[1], ct); atomicAdd(&s_cnt[2], cr);
    }
    px_r = wave_sum64(px_r);
    px_f = wave_sum64(px_f);
    if ((tid & 63) == 0) {
        atomicAdd(&s_px[0], px_r);
        atomicAdd(&s_px[1], px_f);
    }

    // the wave rings (2 BT words) are dead since the last draw; one barrier,
    // then wave 0 runs the rest of the tree: slot t = lane + 64 k sits in
    // v[k], so level "t += t + stride" is v[k] += v[k + stride / 64]
    float* red = reinterpret_cast<float*>(rings);
    red[tid] = acc_e[0];
    __syncthreads();
    if (tid < 64) {
        float r[BT / 64];
#pragma unroll
        for (int k = 0; k < BT / 64; ++k) r[k] = red[tid + 64 * k];
#pragma unroll
        for (int s = BT / 128; s >= 1; s >>= 1) {
#pragma unroll
            for (int k = 0; k < s; ++k) r[k] = r[k] + r[k + s];
        }
        float v = r[0];
#pragma unroll
        for (int stride = 32; stride >= 1; stride >>= 1) {
            const float o = __shfl_down(v, stride);
            v = v + o;
        }
        if (tid == 0) {
            const uint32_t seg = s_cnt[0], tr = s_cnt[1], nr = s_cnt[2];
            const float inv_r = rcp((float)nr);
            const float ratio = (float)seg * rcp((float)S.num_points);
            const float ratioT = (float)tr * inv_r;
            const float zerr = v * inv_r;
            S.fit[p] = (float)__builtin_fma((double)(ratio + ratioT), 0.05, (double)zerr);
            const int ns = s_nstale;
            S.stale_cnt[par ^ 1][p] = ns < cap ? ns : cap;
            atomicAdd(&S.counters[0], s_px[0]);
            atomicAdd(&S.counters[1], s_px[1]);
        }
    }
}

// One workgroup per (solve, particle).  A launch whose solves include a mesh
// beyond the vertex cache runs the windowed body (MODE 1); its particles of
// cached meshes take the windowless body, which is faster for them.
template <int BT, int CAP, int MODE>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(rf_waves_per_simd(BT, CAP))))
void k_render_fitness(const QpsoSolveDev* __restrict__ solves, const QpsoItem* __restrict__ items) {
    __shared__ RfLds<CAP> sh;
    if (MODE == 1) {
        const QpsoItem item = items ? items[blockIdx.x] : QpsoItem{0, (int32_t)blockIdx.x};
        if (solves[item.solve].V_eff <= CAP / 6) {
            rf_body<BT, CAP, 0>(solves, items, sh);
            return;
        }
    }
    rf_body<BT, CAP, MODE>(solves, items, sh);
}

// ------------------------------------------------------------------------
// reference-faithful off-frame mode (qpso_params.offframe = 1).  The
// reference's own per-particle int32 buffer of the whole frame lives in
// device memory and is addressed as the reference does: a bbox pixel (x, y)
// goes to idx = y*640 + x in wrapping 32-bit arithmetic (main.h:1442), so a
// pixel past the left or right edge lands in the neighbouring row, and the
// fitness scan reads and clears pxY*640 + pxX the same way (main.cu:674,
// 688).  An index outside [0, 640*480) belongs to another particle's buffer
// in the reference (or lies outside the allocation): dropped / read as empty.
// Pixels the scan does not visit keep their depth in the buffer.  One face
// per thread, global atomicMin, the scan over the 2048 slots: a slow path for
// exact emulation.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fill_frames(int32_t* __restrict__ frames, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        frames[i] = QPSO_EMPTY;
}

// Observed depth at frame index L in [0, 640*480): the cluster crop inside
// the rect, nothing elsewhere (the cluster depth is zero outside the cluster).
__device__ __forceinline__ int32_t observed_at(const QpsoSolveDev& S, int32_t L) {
    const int32_t y = L / kCols, x = L - y * kCols;
    if (x < S.rect[0] || x >= S.rect[2] || y < S.rect[1] || y >= S.rect[3]) return 0;
    const int32_t rw = S.rect[2] - S.rect[0];
    return (int32_t)((const QG uint16_t*)S.obs)[(y - S.rect[1]) * rw + (x - S.rect[0])];
}

template <int BT>
__global__ __launch_bounds__(BT) void k_render_fitness_wrap(const QpsoSolveDev* __restrict__ solves,
                                                           const QpsoItem* __restrict__ items) {
    constexpr int SPT = 2048 / BT;  // fitness slots per thread: tid + j BT
    constexpr int32_t kFrame = kRows * kCols;
    __shared__ int32_t s_box[4];
    __shared__ int32_t s_degen;
    __shared__ uint32_t s_cnt[3];
    __shared__ unsigned long long s_px[2];
    __shared__ float red[BT];
    const int tid = threadIdx.x;
    const QpsoItem item = items ? items[blockIdx.x] : QpsoItem{0, (int32_t)blockIdx.x};
    const QpsoSolveDev& S = solves[item.solve];
    const int p = item.particle;
    const int iter = __builtin_amdgcn_readfirstlane(((const QG int32_t*)S.iter)[0]);
    if (iter >= S.I) return;  // this solve has finished
    if (!S.frames) {  // a solve without the buffers (the host never launches this)
        if (tid == 0) {
            S.fit[p] = __builtin_nanf("");
            atomicOr(S.flags, QPSO_FLAG_LAUNCH);
        }
        return;
    }
    const int N = S.N;
    float pose[7];
#pragma unroll
    for (int d = 0; d < 7; ++d) pose[d] = ((const QG float*)S.pos)[d * N + p];
    const Cam cam{S.fx, S.fy, S.cx, S.cy};
    if (tid == 0) {
        s_box[0] = QPSO_AABB_MIN_SENT; s_box[1] = QPSO_AABB_MAX_SENT;
        s_box[2] = QPSO_AABB_MIN_SENT; s_box[3] = QPSO_AABB_MAX_SENT;
        s_degen = 0;
        s_cnt[0] = s_cnt[1] = s_cnt[2] = 0;
        s_px[0] = s_px[1] = 0;
    }
    __syncthreads();
    // AABB of every face (main.h:1190-1304, main.cu:423-546): faces
    // [0, F_eff) use exactly the vertices [0, V_eff)
    {
        const QG float* __restrict__ wx = (const QG float*)S.verts;
        const QG float* __restrict__ wy = wx + S.V;
        const QG float* __restrict__ wz = wy + S.V;
        int32_t xl = INT32_MAX, xh = INT32_MIN, yl = INT32_MAX, yh = INT32_MIN;
        int dg = 0;
        for (int v = tid; v < S.V_eff; v += BT) {
            int32_t X, Y;
            float Z;
            dg |= vertex_fixed(pose, wx[v], wy[v], wz[v], cam, X, Y, Z) ? 1 : 0;
            xl = min(xl, X); xh = max(xh, X);
            yl = min(yl, Y); yh = max(yh, Y);
        }
        int bx0 = QPSO_AABB_MIN_SENT, bx1 = QPSO_AABB_MAX_SENT, by0 = QPSO_AABB_MIN_SENT, by1 = QPSO_AABB_MAX_SENT;
        if (xl <= xh) {
            bx0 = wadd(xl, 15) >> 4; bx1 = wadd(xh, 15) >> 4;
            by0 = wadd(yl, 15) >> 4; by1 = wadd(yh, 15) >> 4;
        }
        bx0 = wave_min(bx0); bx1 = wave_max(bx1);
        by0 = wave_min(by0); by1 = wave_max(by1);
        dg = wave_max(dg);
        if ((tid & 63) == 0) {
            atomicMin(&s_box[0], bx0); atomicMax(&s_box[1], bx1);
            atomicMin(&s_box[2], by0); atomicMax(&s_box[3], by1);
            if (dg) s_degen = 1;
        }
    }
    __syncthreads();
    const int32_t ax0 = s_box[0], ax1 = s_box[1], ay0 = s_box[2], ay1 = s_box[3];
    const int32_t minX = min(ax0, S.rect[0]), maxX = max(ax1, S.rect[2]);
    const int32_t minY = min(ay0, S.rect[1]), maxY = max(ay1, S.rect[3]);
    const long long W64 = (long long)maxX - minX, H64 = (long long)maxY - minY;
    const bool degen = s_degen != 0 || W64 <= 0 || H64 <= 0 || W64 * H64 > (long long)QPSO_AREA_LIMIT;
    if (S.dbg_aabb && tid == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) S.dbg_aabb[p * 4 + i] = s_box[i];
    }
    int32_t* __restrict__ buf = S.frames + (size_t)p * kFrame;
    if (degen) {  // guard G: NaN fitness, buffer reset
        for (int i = tid; i < kFrame; i += BT) buf[i] = QPSO_EMPTY;
        if (S.dbg_frames)
            for (int i = tid; i < kFrame; i += BT) S.dbg_frames[(size_t)p * kFrame + i] = QPSO_EMPTY;
        if (tid == 0) {
            S.fit[p] = __builtin_nanf("");
            atomicAdd(&S.counters[2], 1ull);
        }
        return;
    }

    // raster: bbox fill (main.h:1355-1488), one face per thread, the whole bbox
    unsigned long long px_r = 0;
    for (int f = tid; f < S.F_eff; f += BT) {
        Face fc;
        face_project(S.corners, f, pose, cam, fc);
        if (!face_front(fc)) continue;
        Setup su;
        face_setup(fc, su);
        for (int32_t y = fc.miny; y < fc.maxy; ++y) {
            // where y*640 + x cannot wrap (|y| <= 3e6, |x| <= 2^26) the
            // columns that land in the buffer are one interval; elsewhere
            // (union boxes of 2^23 pixels at most are then a few columns
            // wide) every pixel is checked
            int32_t xa = fc.minx, xb = fc.maxx;
            if (y >= -3000000 && y <= 3000000) {
                const int32_t base = y * kCols;
                xa = max(xa, -base);
                xb = min(xb, kFrame - base);
                if (xa >= xb) continue;
            }
            int32_t CX2 = wsub(wadd(su.CY2, wmul(y - fc.miny, su.FDX23)), wmul(xa - fc.minx, su.FDY23));
            int32_t CX3 = wsub(wadd(su.CY3, wmul(y - fc.miny, su.FDX31)), wmul(xa - fc.minx, su.FDY31));
            const int32_t row = wmul(y, kCols);
            for (int32_t x = xa; x < xb; ++x) {
                const int32_t L = wadd(row, x);
                if ((uint32_t)L < (uint32_t)kFrame) {
                    atomicMin(buf + L, pixel_depth(CX2, CX3, su));
                    ++px_r;
                }
                CX2 = wsub(CX2, su.FDY23);
                CX3 = wsub(CX3, su.FDY31);
            }
        }
    }
    __syncthreads();
    if (S.dbg_frames) {  // the rendered buffer, before the fitness pass
        int32_t* fr = S.dbg_frames + (size_t)p * kFrame;
        for (int i = tid; i < kFrame; i += BT) fr[i] = __hip_atomic_load(buf + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    // fitness (main.cu:596-760): slot t & 2047 sums its pixels in increasing t
    const int32_t W = (int32_t)W64, H = (int32_t)H64;
    const int32_t T = ((W * H + 1023) >> 10) << 10;  // NumIters * 1024
    const float rW = rcp((float)W);
    float acc_e[SPT];
    uint32_t n_seg = 0, n_tr = 0, n_rend = 0;
    unsigned long long px_f = 0;
    auto pixel = [&](int32_t t, int32_t& L) {  // false past the last row or outside the buffer
        const int32_t yb = row_of(t, rW);
        const int32_t xb = t % W;  // x_bar = tIdxLoop % W_bar
        const int32_t py = minY + yb;
        if (py > maxY) return 0;
        L = wadd(wmul(py, kCols), minX + xb);
        return (uint32_t)L < (uint32_t)kFrame ? 1 : 2;
    };
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        acc_e[j] = 0.0f;
        for (int32_t t = tid + j * BT; t < T; t += 2048) {
            int32_t L = 0;
            const int r = pixel(t, L);
            if (r == 0) break;  // the row index does not decrease with t
            if (r == 2) continue;
            ++px_f;
            const int32_t zr = __hip_atomic_load(buf + L, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int32_t zk = observed_at(S, L);
            if (zr != QPSO_EMPTY) {
                ++n_rend;
                if (zk == 0) {
                    ++n_tr;
                } else {
                    const float e = (float)wsub(zr, zk) * 0.001f;
                    acc_e[j] = __builtin_fmaf(e, e, acc_e[j]);
                }
            } else if (zk != 0) {
                ++n_seg;
            }
        }
    }
    __syncthreads();
    // visited pixels are cleared after every read (a pixel visited twice reads
    // its rendered value both times); the others carry their depth
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        for (int32_t t = tid + j * BT; t < T; t += 2048) {
            int32_t L = 0;
            const int r = pixel(t, L);
            if (r == 0) break;
            if (r == 1) buf[L] = QPSO_EMPTY;
        }
    }

    // FitnessReductionBallot's halving tree (main.h:1522-1573)
#pragma unroll
    for (int stride = 1024; stride >= BT; stride >>= 1) {
#pragma unroll
        for (int j = 0; j < stride / BT; ++j) acc_e[j] = acc_e[j] + acc_e[j + stride / BT];
    }
    const uint32_t cs = wave_sum(n_seg), ct = wave_sum(n_tr), cr = wave_sum(n_rend);
    px_r = wave_sum64(px_r);
    px_f = wave_sum64(px_f);
    if ((tid & 63) == 0) {
        atomicAdd(&s_cnt[0], cs); atomicAdd(&s_cnt[1], ct); atomicAdd(&s_cnt[2], cr);
        atomicAdd(&s_px[0], px_r);
        atomicAdd(&s_px[1], px_f);
    }
    red[tid] = acc_e[0];
    __syncthreads();
    if (tid < 64) {
        float r[BT / 64];
#pragma unroll
        for (int k = 0; k < BT / 64; ++k) r[k] = red[tid + 64 * k];
#pragma unroll
        for (int s2 = BT / 128; s2 >= 1; s2 >>= 1) {
#pragma unroll
            for (int k = 0; k < s2; ++k) r[k] = r[k] + r[k + s2];
        }
        float v = r[0];
#pragma unroll
        for (int stride = 32; stride >= 1; stride >>= 1) {
            const float o = __shfl_down(v, stride);
            v = v + o;
        }
        if (tid == 0) {
            const uint32_t seg = s_cnt[0], tr = s_cnt[1], nr = s_cnt[2];
            const float inv_r = rcp((float)nr);
            const float ratio = (float)seg * rcp((float)S.num_points);
            const float ratioT = (float)tr * inv_r;
            const float zerr = v * inv_r;
            S.fit[p] = (float)__builtin_fma((double)(ratio + ratioT), 0.05, (double)zerr);
            atomicAdd(&S.counters[0], s_px[0]);
            atomicAdd(&S.counters[1], s_px[1]);
        }
    }
}

// ------------------------------------------------------------------------
// argmin + pbest/gbest + particle update, one workgroup per solve
// ------------------------------------------------------------------------
__device__ __forceinline__ bool better(float av, int ai, float bv, int bi) {
    // invalid (NaN) entries carry index INT32_MAX
    if (ai == INT32_MAX) return false;
    if (bi == INT32_MAX) return true;
    return av < bv || (av == bv && ai < bi);
}

// The state one particle update reads, held in registers.
struct Particle {
    float P[7], V[7], B[7];  // pos, vel, pbest
    float R[12];             // the iteration's random numbers, in order of use
    float fit, pbest_fit;
};

__device__ __forceinline__ void load_particle(const QpsoSolveDev& S, int p, int iter, Particle& q) {
    const int N = S.N;
    const QG float* __restrict__ pos = (const QG float*)S.pos;
    const QG float* __restrict__ vel = (const QG float*)S.vel;
    const QG float* __restrict__ pb = (const QG float*)S.pbest;
#pragma unroll
    for (int d = 0; d < 7; ++d) {
        q.P[d] = pos[d * N + p];
        q.V[d] = vel[d * N + p];
        q.B[d] = pb[d * N + p];
    }
    q.fit = ((const QG float*)S.fit)[p];
    q.pbest_fit = ((const QG float*)S.pbest_fit)[p];
    // d_randIdx: 12 at init + 12 per iteration; a finished solve (iter = I,
    // values unused) reads the table's last rows instead of past its end
    const int s0 = min(12 + 12 * iter, S.S - 12);
    const QG float* __restrict__ rt = (const QG float*)S.rand;
#pragma unroll
    for (int k = 0; k < 12; ++k) q.R[k] = rt[(size_t)p + (size_t)(s0 + k) * (size_t)N];
}

// UpdateParticlesEquationKernel (main.cu:858-1060) for one particle, gb = gbest
__device__ void step_particle(const QpsoSolveDev& S, Particle& q, const float* gb) {
    float* P = q.P;
    float* V = q.V;
    const float* B = q.B;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float r1 = S.c1 * q.R[2 * d];
        const float r2 = S.c2 * q.R[2 * d + 1];
        const float iV = (float)(S.w * (double)V[d]);
        float v = __builtin_fmaf(r1, B[d] - P[d], iV);
        v = __builtin_fmaf(r2, gb[d] - P[d], v);
        V[d] = v;
        P[d] += v;
        if (P[d] < S.x_lo[d]) { P[d] = S.x_lo[d]; V[d] = 0.0f; }
        else if (P[d] > S.x_hi[d]) { P[d] = S.x_hi[d]; V[d] = 0.0f; }
    }
    float wc[3] = {0.0f, 0.0f, 0.0f}, ws[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        float qt[4];
        qprod_conj(k == 0 ? &B[3] : &gb[3], &P[3], qt);
        top_half(qt);
        float n2 = __builtin_fmaf(qt[1], qt[1], qt[2] * qt[2]);
        n2 = __builtin_fmaf(qt[3], qt[3], n2);
        const float n = __builtin_sqrtf(n2);
        if (n > 1.0e-7f) {
            const float sc = (2.0f * acosf_q(qt[0])) * rcp(n);
            float* o = k == 0 ? wc : ws;
            o[0] = qt[1] * sc; o[1] = qt[2] * sc; o[2] = qt[3] * sc;
        }
    }
    V[3] = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float r1 = S.c1 * q.R[6 + 2 * i];
        const float r2 = S.c2 * q.R[7 + 2 * i];
        const float iW = (float)(S.w * (double)V[4 + i]);
        const float v = __builtin_fmaf(r1, wc[i], iW);
        V[4 + i] = __builtin_fmaf(r2, ws[i], v);
    }
    float qq[4] = {P[3], P[4], P[5], P[6]};
    kinematics_from(qq, &P[3], &V[4]);
}

// personal best (strict <, main.cu:786-811), then the update; stores the particle
__device__ __forceinline__ void finish_particle(const QpsoSolveDev& S, int p, Particle& q, const float* gb) {
    const int N = S.N;
    if (q.fit < q.pbest_fit) {
        S.pbest_fit[p] = q.fit;
#pragma unroll
        for (int d = 0; d < 7; ++d) {
            q.B[d] = q.P[d];
            S.pbest[d * N + p] = q.P[d];
        }
    }
    step_particle(S, q, gb);
#pragma unroll
    for (int d = 0; d < 7; ++d) {
        S.pos[d * N + p] = q.P[d];
        S.vel[d * N + p] = q.V[d];
    }
}

// Particle tid's state loads at once with the iteration count and gbest, so
// the step waits on one round of loads (more when N > BT).  Every thread
// finishes the argmin from the per-wave winners, and the owner of the winner
// hands its pose to the others through LDS.
template <int BT>
__global__ __launch_bounds__(BT) void k_pso_step(const QpsoSolveDev* __restrict__ solves) {
    constexpr int kWaves = BT / 64;
    __shared__ float s_v[kWaves];
    __shared__ int s_i[kWaves];
    __shared__ float s_pos[7];
    const QpsoSolveDev& S = solves[blockIdx.x];
    const int tid = threadIdx.x;
    const int N = S.N;
    const int iter = __builtin_amdgcn_readfirstlane(((const QG int32_t*)S.iter)[0]);
    Particle q0;
    if (tid < N) load_particle(S, tid, iter, q0);
    float g[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) g[d] = ((const QG float*)S.gbest)[d];
    if (iter >= S.I) return;

    // argmin, first index on ties, NaN never chosen (thrust::min_element,
    // main.cu:1691-1698)
    float bv = 0.0f;
    int bi = INT32_MAX;
    if (tid < N) {
        if (q0.fit == q0.fit) { bv = q0.fit; bi = tid; }
        if (S.fit_trace) S.fit_trace[(size_t)iter * N + tid] = q0.fit;
    }
    for (int p = tid + BT; p < N; p += BT) {
        const float v = S.fit[p];
        if (v == v && better(v, p, bv, bi)) { bv = v; bi = p; }
        if (S.fit_trace) S.fit_trace[(size_t)iter * N + p] = v;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float ov = __shfl_xor(bv, o);
        const int oi = __shfl_xor(bi, o);
        if (better(ov, oi, bv, bi)) { bv = ov; bi = oi; }
    }
    if ((tid & 63) == 0) { s_v[tid >> 6] = bv; s_i[tid >> 6] = bi; }
    __syncthreads();
    bv = s_v[0];
    bi = s_i[0];
#pragma unroll
    for (int k = 1; k < kWaves; ++k)
        if (better(s_v[k], s_i[k], bv, bi)) { bv = s_v[k]; bi = s_i[k]; }
    const int best = bi == INT32_MAX ? 0 : bi;

    // global best (main.cu:824-851): fit[best] is bv unless every value is
    // NaN, when no comparison succeeds
    const bool gnew = bi != INT32_MAX && bv < g[7];
    if (gnew) {
        if (tid == best % BT) {
#pragma unroll
            for (int d = 0; d < 7; ++d) s_pos[d] = best < BT ? q0.P[d] : S.pos[d * N + best];
        }
        __syncthreads();
#pragma unroll
        for (int d = 0; d < 7; ++d) g[d] = s_pos[d];
        g[7] = bv;
    }
    if (tid == 0) {
        if (gnew) {
#pragma unroll
            for (int d = 0; d < 8; ++d) S.gbest[d] = g[d];
        }
        if (S.gbest_trace) {
#pragma unroll
            for (int d = 0; d < 8; ++d) S.gbest_trace[(size_t)iter * 8 + d] = g[d];
        }
        if (S.argmin_trace) S.argmin_trace[iter] = best;
    }
    if (tid < N) finish_particle(S, tid, q0, g);
    for (int p = tid + BT; p < N; p += BT) {
        Particle q;
        load_particle(S, p, iter, q);
        finish_particle(S, p, q, g);
    }
    if (tid == 0) *S.iter = iter + 1;
}

}  // namespace

// ------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------
namespace qpso_launch {

struct RfCfg {
    int bt, cap;
};

// ------------------------------------------------------------------------
// scratch of the allocation-free entry point: a solve's arena and random
// table, one copy per GPU, held until the process ends.  Host code never
// passes `scratch` itself (in host code it names a stand-in, not device
// memory); it uses the address hipGetSymbolAddress gives for the current
// device.
// ------------------------------------------------------------------------
static __device__ __align__(256) char scratch[QPSO_NOALLOC_SCRATCH_BYTES];

// The address is taken once per device (with that device current) and kept.
hipError_t scratch_address(int device, void** ptr) {
    static std::mutex mu;
    static void* addr[64] = {};
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(mu);
    if (!addr[device]) {
        void* a = nullptr;
        const hipError_t e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[device] = a;
    }
    *ptr = addr[device];
    return hipSuccess;
}

// Per crop row, the number of observed (non-zero) pixels left of each column
// ((rw + 1) counts a row), one thread per row.
__global__ __launch_bounds__(256) void k_obs_rows(const uint16_t* __restrict__ crop, uint16_t* __restrict__ rows,
                                                  int rw, int rh) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= rh) return;
    const uint16_t* c = crop + (size_t)y * rw;
    uint16_t* r = rows + (size_t)y * (rw + 1);
    uint16_t n = 0;
    r[0] = 0;
    for (int x = 0; x < rw; ++x) {
        n = (uint16_t)(n + (c[x] != 0 ? 1 : 0));
        r[x + 1] = n;
    }
}

hipError_t obs_rows(const uint16_t* d_crop, uint16_t* d_rows, int rw, int rh, hipStream_t st) {
    if (rw <= 0 || rh <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_obs_rows, dim3((rh + 255) / 256), dim3(256), 0, st, d_crop, d_rows, rw, rh);
    return hipGetLastError();
}

static const RfCfg kRfCfgs[] = {
    {1024, 40448},  // 1 workgroup / CU, 158 KiB tile
    {512, 40448},   // 1 workgroup / CU, 8 waves
    {512, 20224},   // 2 workgroups / CU
    {256, 20224},   // 2 workgroups / CU, 4 waves each
    {256, 10112},   // 4 workgroups / CU
};

int rf_num_configs() { return (int)(sizeof(kRfCfgs) / sizeof(kRfCfgs[0])); }
int rf_config_threads(int cfg) { return kRfCfgs[cfg].bt; }
int rf_config_tile(int cfg) { return kRfCfgs[cfg].cap; }
bool rf_needs_windows(int cfg, int v_eff) { return v_eff > kRfCfgs[cfg].cap / 6; }

hipError_t init(const QpsoSolveDev* d_solves, int n_solves, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    if (n_solves <= 0) return hipSuccess;
    if (ev0 || ev1)
        hipExtLaunchKernelGGL(k_init, dim3(n_solves), dim3(256), 0, st, ev0, ev1, 0, d_solves);
    else
        hipLaunchKernelGGL(k_init, dim3(n_solves), dim3(256), 0, st, d_solves);
    return hipGetLastError();
}

// With events, the launch stamps them with the dispatch's own start and end
// (hipExtLaunchKernel), so a timed kernel carries no event-record gaps.
template <int BT, int CAP, int MODE>
static void launch_rf_mode(const QpsoSolveDev* d_solves, const QpsoItem* d_items, int n_items, hipStream_t st,
                           hipEvent_t ev0, hipEvent_t ev1) {
    if (ev0 || ev1)
        hipExtLaunchKernelGGL((k_render_fitness<BT, CAP, MODE>), dim3(n_items), dim3(BT), 0, st, ev0, ev1, 0,
                              d_solves, d_items);
    else
        hipLaunchKernelGGL((k_render_fitness<BT, CAP, MODE>), dim3(n_items), dim3(BT), 0, st, d_solves, d_items);
}

template <int BT, int CAP>
static void launch_rf(int mode, const QpsoSolveDev* d_solves, const QpsoItem* d_items, int n_items, hipStream_t st,
                      hipEvent_t ev0, hipEvent_t ev1) {
    if (mode == 2) launch_rf_mode<BT, CAP, 2>(d_solves, d_items, n_items, st, ev0, ev1);
    else if (mode == 1) launch_rf_mode<BT, CAP, 1>(d_solves, d_items, n_items, st, ev0, ev1);
    else launch_rf_mode<BT, CAP, 0>(d_solves, d_items, n_items, st, ev0, ev1);
}

hipError_t render_fitness(int cfg, int mode, const QpsoSolveDev* d_solves, const QpsoItem* d_items, int n_items,
                          hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    if (n_items <= 0) return hipSuccess;
    switch (cfg) {
        case 0: launch_rf<1024, 40448>(mode, d_solves, d_items, n_items, st, ev0, ev1); break;
        case 1: launch_rf<512, 40448>(mode, d_solves, d_items, n_items, st, ev0, ev1); break;
        case 2: launch_rf<512, 20224>(mode, d_solves, d_items, n_items, st, ev0, ev1); break;
        case 3: launch_rf<256, 20224>(mode, d_solves, d_items, n_items, st, ev0, ev1); break;
        case 4: launch_rf<256, 10112>(mode, d_solves, d_items, n_items, st, ev0, ev1); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t render_fitness_wrap(const QpsoSolveDev* d_solves, const QpsoItem* d_items, int n_items, hipStream_t st,
                               hipEvent_t ev0, hipEvent_t ev1) {
    if (n_items <= 0) return hipSuccess;
    if (ev0 || ev1)
        hipExtLaunchKernelGGL(k_render_fitness_wrap<256>, dim3(n_items), dim3(256), 0, st, ev0, ev1, 0, d_solves,
                              d_items);
    else
        hipLaunchKernelGGL(k_render_fitness_wrap<256>, dim3(n_items), dim3(256), 0, st, d_solves, d_items);
    return hipGetLastError();
}

hipError_t fill_frames(int32_t* d_frames, size_t n_words, hipStream_t st) {
    if (n_words == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((n_words + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_frames, dim3((unsigned)blocks), dim3(256), 0, st, d_frames, n_words);
    return hipGetLastError();
}

hipError_t pso_step(const QpsoSolveDev* d_solves, int n_solves, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    if (n_solves <= 0) return hipSuccess;
    if (ev0 || ev1)
        hipExtLaunchKernelGGL(k_pso_step<256>, dim3(n_solves), dim3(256), 0, st, ev0, ev1, 0, d_solves);
    else
        hipLaunchKernelGGL(k_pso_step<256>, dim3(n_solves), dim3(256), 0, st, d_solves);
    return hipGetLastError();
}

}  // namespace qpso_launch
