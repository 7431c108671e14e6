// qpso_api.cpp -- C ABI of the gfx950 Q-PSO hot path (include/qpso.h).
//
// Host orchestration replacing the reference's estimate-pose block
// (main.cu:1269-1733): parameters (main.cu:1303-1331), random table
// (main.cu:1337-1352), uploads (main.cu:1366-1557), InitQPSO (main.cu:1607)
// and the iteration loop (main.cu:1625-1727).  All device memory of a solve
// lives in one arena; the mesh and the random table are shared by solves.
#include "qpso.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "qpso_internal.h"
#include "qpso_kernels.h"
#include "qpso_sched.h"
#include "qpso_tables.h"

namespace {

constexpr int kRows = QPSO_ROWS;
constexpr int kCols = QPSO_COLS;
constexpr int kFrame = kRows * kCols;
constexpr int kStaleCap = QPSO_STALE_CAP;
constexpr int kMaxIters = 1 << 20;          // iterations per solve
constexpr size_t kMaxRand = size_t(1) << 30;  // random-table entries (4 GiB)

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }


// One device random table, shared by the solves of equal (seed, N, S).
struct RandTable {
    float* d = nullptr;
    size_t bytes = 0;
    int refs = 0;           // solves using it
    uint64_t last_use = 0;  // for eviction of unused tables
};
// unused tables kept for later solves, at most this many bytes
constexpr size_t kRandIdleBytes = size_t(256) << 20;

struct Mesh {
    float* verts = nullptr;        // unique vertices [3][V]
    QpsoFaceRef* faces = nullptr;  // interleaved raster order [F]
    float* corners = nullptr;      // [F][3][4]: a face's corners (x, y, z, vertex index bits), 16 bytes each
    uint32_t F = 0, V = 0;
    std::vector<uint32_t> v_prefix;  // unique vertices referenced by faces [0, f+1)
    int32_t* runs = nullptr;         // (first face, first vertex, end vertex) per run
    int n_runs = 0;
    int refs = 0;           // solves created on the mesh and not yet destroyed
    bool released = false;  // qpso_mesh_release called; freed when refs reaches 0
};

void free_mesh(Mesh& m) {
    hipFree(m.verts);
    hipFree(m.faces);
    hipFree(m.corners);
    hipFree(m.runs);
    m = Mesh{};
}

struct ArenaLayout {
    size_t desc, obs, obs_rows, pos, vel, pbest, pbest_fit, fit, gbest, iter, flags, counters;
    size_t stale0, stale1, cnt0, cnt1, fit_tr, gbest_tr, argmin_tr, frames, total;
};

// `wrap`: the reference-faithful off-frame mode's full per-particle buffers
ArenaLayout layout(int N, int I, int rw, int rh, bool trace, bool wrap = false) {
    ArenaLayout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) { size_t r = o; o = align_up(o + bytes, 256); return r; };
    L.desc = take(sizeof(QpsoSolveDev));
    L.obs = take(sizeof(uint16_t) * (size_t)rw * rh);
    L.obs_rows = take(sizeof(uint16_t) * (size_t)(rw + 1) * rh);
    L.pos = take(sizeof(float) * 7 * N);
    L.vel = take(sizeof(float) * 7 * N);
    L.pbest = take(sizeof(float) * 7 * N);
    L.pbest_fit = take(sizeof(float) * N);
    L.fit = take(sizeof(float) * N);
    L.gbest = take(sizeof(float) * 8);
    L.iter = take(sizeof(int32_t));
    L.flags = take(sizeof(int32_t));
    L.counters = take(sizeof(unsigned long long) * 3);
    L.stale0 = take(sizeof(QpsoStale) * (size_t)N * kStaleCap);
    L.stale1 = take(sizeof(QpsoStale) * (size_t)N * kStaleCap);
    L.cnt0 = take(sizeof(int32_t) * N);
    L.cnt1 = take(sizeof(int32_t) * N);
    if (trace) {
        L.fit_tr = take(sizeof(float) * (size_t)I * N);
        L.gbest_tr = take(sizeof(float) * 8 * (size_t)I);
        L.argmin_tr = take(sizeof(int32_t) * (size_t)I);
    }
    if (wrap) L.frames = take(sizeof(int32_t) * (size_t)N * kFrame);
    L.total = o;
    return L;
}

// The device descriptor of a solve whose arena starts at A (layout L).
void fill_desc(QpsoSolveDev& h, const Mesh& m, const qpso_cluster* cl, const qpso_params* p, char* A,
               const ArenaLayout& L, const float* rand, bool trace, int big_face) {
    h = QpsoSolveDev{};
    const int N = p->n_particles, I = p->n_iters;
    h.F = (int32_t)m.F;
    h.F_eff = (p->face_cap > 0 && (uint32_t)p->face_cap < m.F) ? p->face_cap : (int32_t)m.F;
    h.verts = m.verts;
    h.V = (int32_t)m.V;
    h.V_eff = (int32_t)m.v_prefix[h.F_eff - 1];
    h.faces = m.faces;
    h.corners = m.corners;
    h.n_faces = (int32_t)m.F;
    h.runs = m.runs;
    h.n_runs = m.n_runs;
    h.fx = p->cam.fx; h.fy = p->cam.fy; h.cx = p->cam.cx; h.cy = p->cam.cy;
    h.N = N; h.I = I; h.S = qpso_host::rand_stride(I);
    h.w = p->w; h.c1 = p->c1; h.c2 = p->c2;
    // translation bounds around the centroid in metres (main.cu:1290-1331)
    const float cxm = cl->centroid_mm[0] * 0.001f, cym = cl->centroid_mm[1] * 0.001f,
                czm = cl->centroid_mm[2] * 0.001f;
    h.x_lo[0] = cxm - p->t_bound;   h.x_hi[0] = cxm + p->t_bound;
    h.x_lo[1] = cym - p->t_bound;   h.x_hi[1] = cym + p->t_bound;
    h.x_lo[2] = czm - p->t_bound_z; h.x_hi[2] = czm + p->t_bound_z;
    std::memcpy(h.rect, cl->rect, sizeof(h.rect));
    h.num_points = cl->num_points;
    h.obs = reinterpret_cast<const uint16_t*>(A + L.obs);
    h.obs_rows = reinterpret_cast<const uint16_t*>(A + L.obs_rows);
    h.rand = rand;
    h.pos = reinterpret_cast<float*>(A + L.pos);
    h.vel = reinterpret_cast<float*>(A + L.vel);
    h.pbest = reinterpret_cast<float*>(A + L.pbest);
    h.pbest_fit = reinterpret_cast<float*>(A + L.pbest_fit);
    h.fit = reinterpret_cast<float*>(A + L.fit);
    h.gbest = reinterpret_cast<float*>(A + L.gbest);
    h.iter = reinterpret_cast<int32_t*>(A + L.iter);
    h.stale[0] = reinterpret_cast<QpsoStale*>(A + L.stale0);
    h.stale[1] = reinterpret_cast<QpsoStale*>(A + L.stale1);
    h.stale_cnt[0] = reinterpret_cast<int32_t*>(A + L.cnt0);
    h.stale_cnt[1] = reinterpret_cast<int32_t*>(A + L.cnt1);
    h.stale_cap = kStaleCap;
    h.big_face = big_face;
    h.flags = reinterpret_cast<int32_t*>(A + L.flags);
    h.counters = reinterpret_cast<unsigned long long*>(A + L.counters);
    if (trace) {
        h.fit_trace = reinterpret_cast<float*>(A + L.fit_tr);
        h.gbest_trace = reinterpret_cast<float*>(A + L.gbest_tr);
        h.argmin_trace = reinterpret_cast<int32_t*>(A + L.argmin_tr);
    }
    if (p->offframe) {
        h.wrap = 1;
        h.frames = reinterpret_cast<int32_t*>(A + L.frames);
    }
}

}  // namespace

struct qpso_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;  // serialises API calls that use `stream`
    std::string err;
    std::map<uint32_t, Mesh> meshes;
    uint32_t next_mesh = 1;
    std::map<qpso_sched::TableKey, RandTable> rand_cache;
    uint64_t rand_clock = 0;
    int rf_cfg = 0;
    int n_cu = 0;
    // cached batch plan (graph of a whole batched run)
    struct Plan {
        std::vector<qpso_solve*> solves;
        int max_iters = 0;
        QpsoSolveDev* d_descs = nullptr;
        QpsoItem* d_items = nullptr;
        int n_items = 0;
        int rf_cfg = 0;
        hipGraphExec_t exec = nullptr;
    } plan;
    int big_face = 128;  // default big-face threshold of new solves
    // this device's copy of the allocation-free entry point's scratch
    // (hipGetSymbolAddress at create; null if unavailable)
    char* scratch = nullptr;
    // solve arenas come from this stream-ordered pool where the device has
    // one: a solve's memory is then released in the order of the stream that
    // ran it, with no wait (hipFree waits for the whole device); freed
    // memory stays in the pool for reuse until the context is destroyed
    hipMemPool_t pool = nullptr;
};

struct qpso_solve {
    qpso_ctx* ctx = nullptr;
    QpsoSolveDev h{};             // host copy of the device descriptor
    QpsoSolveDev* d_desc = nullptr;
    char* arena = nullptr;
    ArenaLayout L{};
    bool trace = false;
    qpso_params params{};
    int32_t rect[4]{};
    uint32_t mesh_id = 0;
    qpso_sched::TableKey rand_key{};  // its (shared) random table
    int rf_cfg = 0;                  // render+fitness configuration (qpso_autotune)
    hipGraphExec_t exec = nullptr;   // captured plain run: init + I x (render+fitness, PSO)
    std::vector<hipEvent_t> timed_ev;     // qpso_solve_run_timed's events, kept with the solve
};

namespace {

int set_err(qpso_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

#define HIP_TRY(ctx, expr)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return set_err((ctx), QPSO_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

int validate(const qpso_cluster* cl, const qpso_params* p) {
    if (!cl || !p || !cl->depth) return QPSO_E_INVAL;
    if (p->n_particles <= 0 || p->n_particles > 65535 || p->n_iters <= 0 || p->n_iters > kMaxIters)
        return QPSO_E_INVAL;
    const int32_t* r = cl->rect;
    if (r[0] < 0 || r[1] < 0 || r[2] > kCols || r[3] > kRows || r[2] <= r[0] || r[3] <= r[1])
        return QPSO_E_INVAL;
    if (cl->num_points == 0) return QPSO_E_INVAL;
    if (!(p->t_bound >= 0.0f) || !(p->t_bound_z >= 0.0f)) return QPSO_E_INVAL;
    if (p->offframe != 0 && p->offframe != 1) return QPSO_E_INVAL;
    return QPSO_OK;
}

// The random table of (seed, N, S), generated and uploaded on first use; the
// caller holds a reference until release_rand.  Caller holds ctx->mu.
// A blocking copy on the given stream.  The library never uses the null
// stream: that would take one of the device's hardware queues (four per
// process by default, handed to the least used first), which the serving
// lanes need one each to run side by side.
hipError_t copy_on(hipStream_t st, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, st);
    return e == hipSuccess ? hipStreamSynchronize(st) : e;
}

// A solve arena: from the context's pool in the order of its stream (the
// caller synchronises that stream before another one uses it), or from
// hipMalloc.
hipError_t alloc_arena(qpso_ctx* ctx, size_t bytes, char** p) {
    if (ctx->pool) return hipMallocFromPoolAsync(reinterpret_cast<void**>(p), bytes, ctx->pool, ctx->stream);
    return hipMalloc(reinterpret_cast<void**>(p), bytes);
}

// Release an arena after the work queued on `st` (pooled: without waiting;
// otherwise hipFree, which waits for the whole device).
void free_arena(qpso_ctx* ctx, char* p, hipStream_t st) {
    if (ctx->pool) hipFreeAsync(p, st);
    else hipFree(p);
}

int get_rand(qpso_ctx* ctx, const qpso_sched::TableKey& key, float** out) {
    auto it = ctx->rand_cache.find(key);
    if (it != ctx->rand_cache.end()) {
        ++it->second.refs;
        it->second.last_use = ++ctx->rand_clock;
        *out = it->second.d;
        return QPSO_OK;
    }
    const uint32_t seed = std::get<0>(key);
    const int N = std::get<1>(key), S = std::get<2>(key);
    if ((size_t)S * N > kMaxRand) return set_err(ctx, QPSO_E_NOMEM, "random table too large");
    std::vector<float> h;
    qpso_host::rand_table(seed, (size_t)S * N, h);
    float* d = nullptr;
    HIP_TRY(ctx, hipMalloc(&d, h.size() * sizeof(float)));
    if (copy_on(ctx->stream, d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(d);
        return set_err(ctx, QPSO_E_HIP, "random table upload failed");
    }
    RandTable& t = ctx->rand_cache[key];
    t.d = d;
    t.bytes = h.size() * sizeof(float);
    t.refs = 1;
    t.last_use = ++ctx->rand_clock;
    *out = d;
    return QPSO_OK;
}

// Drop a solve's reference; unused tables beyond kRandIdleBytes are freed,
// least recently used first.  Caller holds ctx->mu, and no queued work reads
// an unused table (its last solve was synchronised before release).
void release_rand(qpso_ctx* ctx, const qpso_sched::TableKey& key) {
    auto it = ctx->rand_cache.find(key);
    if (it == ctx->rand_cache.end() || --it->second.refs > 0) return;
    for (;;) {
        size_t idle = 0;
        auto lru = ctx->rand_cache.end();
        for (auto j = ctx->rand_cache.begin(); j != ctx->rand_cache.end(); ++j) {
            if (j->second.refs > 0) continue;
            idle += j->second.bytes;
            if (lru == ctx->rand_cache.end() || j->second.last_use < lru->second.last_use) lru = j;
        }
        if (idle <= kRandIdleBytes || lru == ctx->rand_cache.end()) return;
        hipFree(lru->second.d);
        ctx->rand_cache.erase(lru);
    }
}

void invalidate_plan(qpso_ctx* ctx) {
    auto& P = ctx->plan;
    if (P.exec) hipGraphExecDestroy(P.exec);
    if (P.d_descs) hipFree(P.d_descs);
    if (P.d_items) hipFree(P.d_items);
    P = qpso_ctx::Plan{};
}

// A finished solve's result, read on stream `st`; touches no context state
// (the serving thread calls it without the context lock).
// A solve's result words as they lie in its arena: gbest pose and fitness,
// the work counters, the error flags.
struct ResultWords {
    float g[8];
    unsigned long long c[3];
    int32_t flags;
    int32_t pad;
};

// Enqueue the copies of the result words into `w` on `st`.
hipError_t copy_result(const qpso_solve* s, ResultWords* w, hipStream_t st) {
    hipError_t e = hipMemcpyAsync(w->g, s->arena + s->L.gbest, sizeof(w->g), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(w->c, s->arena + s->L.counters, sizeof(w->c), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&w->flags, s->arena + s->L.flags, sizeof(w->flags), hipMemcpyDeviceToHost, st);
    return e;
}

// The result from its words; returns its status.
int result_from(const ResultWords& w, qpso_result* out) {
    const int status = (w.flags & QPSO_FLAG_LAUNCH) ? QPSO_E_HIP
                       : (w.flags & QPSO_FLAG_STALE_OVERFLOW) ? QPSO_E_STALE : QPSO_OK;
    if (out) {
        std::memcpy(out->pose, w.g, sizeof(float) * 7);
        out->fit = w.g[7];
        out->raster_px = w.c[0];
        out->visit_px = w.c[1];
        out->n_degenerate = (uint32_t)w.c[2];
        out->status = status;
        out->latency_ms = 0.0;
    }
    return status;
}

int read_result_status(qpso_solve* s, qpso_result* out, hipStream_t st) {
    ResultWords w{};
    if (copy_result(s, &w, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) return QPSO_E_HIP;
    return result_from(w, out);
}

int read_result(qpso_solve* s, qpso_result* out, hipStream_t st) {
    const int rc = read_result_status(s, out, st);
    switch (rc) {
        case QPSO_OK: return QPSO_OK;
        case QPSO_E_STALE: return set_err(s->ctx, rc, "stale-pixel list overflow");
        default: return set_err(s->ctx, rc, "result readback failed, or a render kernel launched without mesh windows");
    }
}

bool needs_windows(const qpso_solve* s, int cfg) { return qpso_launch::rf_needs_windows(cfg, s->h.V_eff); }

// One render+fitness launch of a solve's particles (all, or the items given):
// the reference-faithful off-frame mode has its own kernel.
hipError_t launch_rf(const qpso_solve* s, int cfg, const QpsoSolveDev* d_desc, const QpsoItem* d_items, int n,
                     hipStream_t st, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr) {
    if (s->h.wrap) return qpso_launch::render_fitness_wrap(d_desc, d_items, n, st, ev0, ev1);
    return qpso_launch::render_fitness(cfg, needs_windows(s, cfg), d_desc, d_items, n, st, ev0, ev1);
}

bool any_wrap(qpso_solve* const* solves, int n) {
    for (int i = 0; i < n; ++i)
        if (solves[i]->h.wrap) return true;
    return false;
}
bool any_needs_windows(qpso_solve* const* solves, int n, int cfg) {
    for (int i = 0; i < n; ++i)
        if (needs_windows(solves[i], cfg)) return true;
    return false;
}

// Launch the whole iteration loop of one solve (plain form).
int launch_solve(qpso_solve* s, hipStream_t st) {
    qpso_ctx* ctx = s->ctx;
    if (s->h.wrap) HIP_TRY(ctx, qpso_launch::fill_frames(s->h.frames, (size_t)s->h.N * kFrame, st));
    HIP_TRY(ctx, qpso_launch::init(s->d_desc, 1, st));
    for (int it = 0; it < s->h.I; ++it) {
        HIP_TRY(ctx, launch_rf(s, s->rf_cfg, s->d_desc, nullptr, s->h.N, st));
        HIP_TRY(ctx, qpso_launch::pso_step(s->d_desc, 1, st));
    }
    return QPSO_OK;
}

// The plain run as one graph launch (captured at the first run, and again
// after the configuration changes): 2I + 1 kernels without host launch cost.
int run_solve_graph(qpso_solve* s, hipStream_t st) {
    qpso_ctx* ctx = s->ctx;
    if (!s->exec) {
        hipGraph_t graph = nullptr;
        HIP_TRY(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = launch_solve(s, st);
        const hipError_t e = hipStreamEndCapture(st, &graph);
        if (rc || e != hipSuccess) {
            if (graph) hipGraphDestroy(graph);
            return rc ? rc : set_err(ctx, QPSO_E_HIP, "solve graph capture failed");
        }
        const hipError_t e2 = hipGraphInstantiate(&s->exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e2 != hipSuccess) { s->exec = nullptr; return set_err(ctx, QPSO_E_HIP, "solve graph instantiate failed"); }
    }
    HIP_TRY(ctx, hipGraphLaunch(s->exec, st));
    return QPSO_OK;
}

void drop_solve_graph(qpso_solve* s) {
    if (s->exec) hipGraphExecDestroy(s->exec);
    s->exec = nullptr;
}

}  // namespace

extern "C" {

const char* qpso_status_str(int status) {
    switch (status) {
        case QPSO_OK: return "ok";
        case QPSO_E_INVAL: return "invalid argument";
        case QPSO_E_HIP: return "HIP runtime error";
        case QPSO_E_NOMEM: return "out of memory";
        case QPSO_E_BUDGET: return "request exceeds the HBM budget";
        case QPSO_E_NOTFOUND: return "unknown id";
        case QPSO_E_STALE: return "stale-pixel list overflow";
        case QPSO_E_PENDING: return "pending";
        case QPSO_E_CANCELLED: return "cancelled";
        case QPSO_E_TIMEOUT: return "timed out";
        default: return "unknown status";
    }
}

void qpso_params_default(qpso_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->n_particles = 1024;       // Nparticle_ (main.h:84)
    p->n_iters = 20;             // max_steps_ (main.h:99)
    p->w = 0.3;                  // psoW_ (main.h:105)
    p->c1 = 1.0f;                // c1_ (main.h:103)
    p->c2 = 1.0f;                // c2_ (main.h:104)
    p->t_bound = 0.3f;           // main.cu:1324
    p->t_bound_z = 0.3f;         // main.cu:1325
    p->seed = 0;                 // gsl default seed
    p->cam = {572.41140, 573.57043, 325.26110, 242.04899};  // FX, FY, CX, CY (main.h:30-33)
    p->face_cap = 0;
}

int qpso_create(int device, qpso_ctx** out) {
    if (!out) return QPSO_E_INVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0 || device < 0 || device >= n) return QPSO_E_HIP;
    if (hipSetDevice(device) != hipSuccess) return QPSO_E_HIP;
    auto* ctx = new qpso_ctx();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return QPSO_E_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        ctx->n_cu = prop.multiProcessorCount;
        if (prop.memoryPoolsSupported) {
            hipMemPoolProps pp{};
            pp.allocType = hipMemAllocationTypePinned;
            pp.handleTypes = hipMemHandleTypeNone;
            pp.location.type = hipMemLocationTypeDevice;
            pp.location.id = device;
            hipMemPool_t pool = nullptr;
            uint64_t keep = UINT64_MAX;
            if (hipMemPoolCreate(&pool, &pp) == hipSuccess) {
                if (hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess) ctx->pool = pool;
                else hipMemPoolDestroy(pool);
            }
            hipGetLastError();  // without a pool the arenas come from hipMalloc
        }
    }
    // with this device current: its copy of the no-allocation scratch
    void* scratch = nullptr;
    if (qpso_launch::scratch_address(device, &scratch) == hipSuccess) ctx->scratch = static_cast<char*>(scratch);
    else hipGetLastError();  // leaves the entry point unavailable, nothing else
    if (const char* e = std::getenv("QPSO_RF_CFG")) {
        const int c = std::atoi(e);
        if (c >= 0 && c < qpso_launch::rf_num_configs()) ctx->rf_cfg = c;
    }
    *out = ctx;
    return QPSO_OK;
}

int qpso_destroy(qpso_ctx* ctx) {
    if (!ctx) return QPSO_E_INVAL;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    invalidate_plan(ctx);
    for (auto& kv : ctx->meshes) free_mesh(kv.second);
    for (auto& kv : ctx->rand_cache) hipFree(kv.second.d);
    if (ctx->pool) {
        hipDeviceSynchronize();
        hipMemPoolDestroy(ctx->pool);
    }
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return QPSO_OK;
}

const char* qpso_last_error(qpso_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int qpso_mesh_upload(qpso_ctx* ctx, const float* soa, uint32_t n_faces, uint32_t* mesh_id) {
    if (!ctx || !soa || n_faces == 0 || !mesh_id || n_faces > (1u << 26)) return QPSO_E_INVAL;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    Mesh m;
    m.F = n_faces;
    std::vector<float> verts;
    std::vector<QpsoFaceRef> order;
    std::vector<int32_t> runs;
    qpso_host::build_mesh_tables(soa, n_faces, verts, order, m.v_prefix, runs);
    m.n_runs = (int)(runs.size() / 3);
    m.V = (uint32_t)(verts.size() / 3);
    std::vector<float> corners;
    qpso_host::build_corner_table(soa, n_faces, order, corners);
    hipError_t e = hipMalloc(&m.corners, sizeof(float) * corners.size());
    if (e == hipSuccess) e = hipMalloc(&m.verts, sizeof(float) * verts.size());
    if (e == hipSuccess) e = hipMalloc(&m.faces, sizeof(QpsoFaceRef) * order.size());
    if (e == hipSuccess) e = hipMalloc(&m.runs, sizeof(int32_t) * std::max<size_t>(runs.size(), 1));
    if (e == hipSuccess) e = copy_on(ctx->stream, m.verts, verts.data(), sizeof(float) * verts.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = copy_on(ctx->stream, m.corners, corners.data(), sizeof(float) * corners.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = copy_on(ctx->stream, m.faces, order.data(), sizeof(QpsoFaceRef) * order.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !runs.empty())
        e = copy_on(ctx->stream, m.runs, runs.data(), sizeof(int32_t) * runs.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_mesh(m);
        return set_err(ctx, e == hipErrorOutOfMemory ? QPSO_E_NOMEM : QPSO_E_HIP,
                       std::string("mesh upload: ") + hipGetErrorString(e));
    }
    *mesh_id = ctx->next_mesh++;
    ctx->meshes[*mesh_id] = std::move(m);
    return QPSO_OK;
}

int qpso_mesh_release(qpso_ctx* ctx, uint32_t mesh_id) {
    if (!ctx) return QPSO_E_INVAL;
    std::lock_guard<std::mutex> g(ctx->mu);
    auto it = ctx->meshes.find(mesh_id);
    if (it == ctx->meshes.end() || it->second.released) return QPSO_E_NOTFOUND;
    // solves still created on the mesh keep it until the last is destroyed
    it->second.released = true;
    if (it->second.refs > 0) return QPSO_OK;
    hipStreamSynchronize(ctx->stream);
    free_mesh(it->second);
    ctx->meshes.erase(it);
    return QPSO_OK;
}

uint64_t qpso_solve_bytes(const qpso_params* p, const int32_t rect[4]) {
    if (!p || !rect) return 0;
    const int rw = std::max(0, rect[2] - rect[0]), rh = std::max(0, rect[3] - rect[1]);
    return layout(p->n_particles, p->n_iters, rw, rh, false, p->offframe != 0).total;
}

int qpso_solve_create(qpso_ctx* ctx, uint32_t mesh_id, const qpso_cluster* cl, const qpso_params* p,
                      int with_trace, qpso_solve** out) {
    if (!ctx || !out) return QPSO_E_INVAL;
    *out = nullptr;
    if (int v = validate(cl, p)) return set_err(ctx, v, "invalid cluster or parameters");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    auto mit = ctx->meshes.find(mesh_id);
    if (mit == ctx->meshes.end() || mit->second.released) return set_err(ctx, QPSO_E_NOTFOUND, "unknown mesh id");
    const int N = p->n_particles, I = p->n_iters, S = qpso_host::rand_stride(I);
    const int rw = cl->rect[2] - cl->rect[0], rh = cl->rect[3] - cl->rect[1];

    std::unique_ptr<qpso_solve> s(new qpso_solve());
    s->ctx = ctx;
    s->trace = with_trace != 0;
    s->params = *p;
    s->mesh_id = mesh_id;
    std::memcpy(s->rect, cl->rect, sizeof(s->rect));
    s->L = layout(N, I, rw, rh, s->trace, p->offframe != 0);
    if (alloc_arena(ctx, s->L.total, &s->arena) != hipSuccess)
        return set_err(ctx, QPSO_E_NOMEM, "solve arena allocation failed");
    float* rand = nullptr;
    s->rand_key = qpso_sched::TableKey(qpso_host::rand_seed(p->seed), N, S);
    if (int r = get_rand(ctx, s->rand_key, &rand)) {
        free_arena(ctx, s->arena, ctx->stream);
        hipStreamSynchronize(ctx->stream);
        return r;
    }

    fill_desc(s->h, mit->second, cl, p, s->arena, s->L, rand, s->trace, ctx->big_face);
    s->rf_cfg = ctx->rf_cfg;
    char* A = s->arena;
    s->d_desc = reinterpret_cast<QpsoSolveDev*>(A + s->L.desc);

    // cluster depth: only its rect is non-zero, keep that crop resident, with
    // per crop row the number of observed pixels left of each column
    std::vector<uint16_t> crop, rows;
    qpso_host::build_crop(cl->depth, kCols, cl->rect, crop, rows);
    hipStream_t st = ctx->stream;
    int rc = QPSO_OK;
    if (hipMemcpyAsync(A + s->L.obs, crop.data(), crop.size() * 2, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(A + s->L.obs_rows, rows.data(), rows.size() * 2, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(s->d_desc, &s->h, sizeof(QpsoSolveDev), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        rc = set_err(ctx, QPSO_E_HIP, "solve upload failed");
    if (rc) {
        free_arena(ctx, s->arena, st);
        hipStreamSynchronize(st);
        release_rand(ctx, s->rand_key);
        return rc;
    }
    ++mit->second.refs;
    *out = s.release();
    return QPSO_OK;
}

namespace {

// Destroy a solve whose work all lies on stream `st`; its arena is released
// in that stream's order.
void solve_release(qpso_solve* s, hipStream_t st) {
    qpso_ctx* ctx = s->ctx;
    auto& P = ctx->plan;
    if (std::find(P.solves.begin(), P.solves.end(), s) != P.solves.end()) invalidate_plan(ctx);
    drop_solve_graph(s);
    for (auto& e : s->timed_ev) hipEventDestroy(e);
    free_arena(ctx, s->arena, st);
    release_rand(ctx, s->rand_key);
    auto mit = ctx->meshes.find(s->mesh_id);
    if (mit != ctx->meshes.end() && --mit->second.refs == 0 && mit->second.released) {
        free_mesh(mit->second);
        ctx->meshes.erase(mit);
    }
    delete s;
}

}  // namespace

int qpso_solve_destroy(qpso_solve* s) {
    if (!s) return QPSO_E_INVAL;
    qpso_ctx* ctx = s->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    solve_release(s, ctx->stream);
    return QPSO_OK;
}

int qpso_solve_run(qpso_solve* s, qpso_result* out) {
    if (!s) return QPSO_E_INVAL;
    qpso_ctx* ctx = s->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (int r = run_solve_graph(s, ctx->stream)) return r;
    return read_result(s, out, ctx->stream);
}

int qpso_solve_run_reference_launches(qpso_solve* s, qpso_result* out) {
    if (!s) return QPSO_E_INVAL;
    qpso_ctx* ctx = s->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int N = s->h.N, I = s->h.I;
    constexpr int kStreams = 16;  // NumStreams (main.h): particle p on stream p % 16
    std::vector<QpsoItem> items(N);
    for (int p = 0; p < N; ++p) items[p] = QpsoItem{0, p};
    QpsoItem* d_items = nullptr;
    hipStream_t ps[kStreams] = {};
    hipEvent_t join[kStreams] = {}, fork = nullptr;
    auto cleanup = [&]() {
        for (auto& x : ps) if (x) hipStreamDestroy(x);
        for (auto& x : join) if (x) hipEventDestroy(x);
        if (fork) hipEventDestroy(fork);
        if (d_items) hipFree(d_items);
    };
    hipError_t e = hipMalloc(&d_items, sizeof(QpsoItem) * N);
    if (e == hipSuccess) e = copy_on(st, d_items, items.data(), sizeof(QpsoItem) * N, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
    for (int k = 0; k < kStreams && e == hipSuccess; ++k) {
        e = hipStreamCreateWithFlags(&ps[k], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&join[k], hipEventDisableTiming);
    }
    float best = 0.0f;
    if (e == hipSuccess && s->h.wrap) e = qpso_launch::fill_frames(s->h.frames, (size_t)N * kFrame, st);
    if (e == hipSuccess) e = qpso_launch::init(s->d_desc, 1, st);
    for (int it = 0; it < I && e == hipSuccess; ++it) {
        e = hipEventRecord(fork, st);
        for (int k = 0; k < kStreams && e == hipSuccess; ++k) e = hipStreamWaitEvent(ps[k], fork, 0);
        for (int p = 0; p < N && e == hipSuccess; ++p)
            e = launch_rf(s, s->rf_cfg, s->d_desc, d_items + p, 1, ps[p % kStreams]);
        for (int k = 0; k < kStreams && e == hipSuccess; ++k) {
            e = hipEventRecord(join[k], ps[k]);
            if (e == hipSuccess) e = hipStreamWaitEvent(st, join[k], 0);
        }
        if (e == hipSuccess) e = qpso_launch::pso_step(s->d_desc, 1, st);
        // the reference reads the argmin back to the host every iteration
        if (e == hipSuccess) e = hipMemcpyAsync(&best, s->arena + s->L.gbest + 7 * sizeof(float), sizeof(float),
                                                hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    cleanup();
    if (e != hipSuccess) return set_err(ctx, QPSO_E_HIP, std::string("reference-shaped run: ") + hipGetErrorString(e));
    return read_result(s, out, st);
}

int qpso_solve_run_timed(qpso_solve* s, qpso_result* out, qpso_timing* tm) {
    if (!s || !tm) return QPSO_E_INVAL;
    qpso_ctx* ctx = s->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int I = s->h.I;
    // a start and a stop event per kernel, stamped by the kernel's dispatch
    // itself (no event records between the kernels)
    const size_t n_ev = 2 * (1 + 2 * (size_t)I);
    if (s->timed_ev.size() != n_ev) {
        for (auto& e : s->timed_ev) hipEventDestroy(e);
        s->timed_ev.assign(n_ev, nullptr);
        for (auto& e : s->timed_ev) HIP_TRY(ctx, hipEventCreate(&e));
    }
    const std::vector<hipEvent_t>& ev = s->timed_ev;
    hipError_t e = hipSuccess;
    if (s->h.wrap) e = qpso_launch::fill_frames(s->h.frames, (size_t)s->h.N * kFrame, st);
    if (e == hipSuccess) e = qpso_launch::init(s->d_desc, 1, st, ev[0], ev[1]);
    for (int it = 0; it < I && e == hipSuccess; ++it) {
        e = launch_rf(s, s->rf_cfg, s->d_desc, nullptr, s->h.N, st, ev[2 + 4 * it], ev[3 + 4 * it]);
        if (e == hipSuccess) e = qpso_launch::pso_step(s->d_desc, 1, st, ev[4 + 4 * it], ev[5 + 4 * it]);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return set_err(ctx, QPSO_E_HIP, std::string("timed run: ") + hipGetErrorString(e));
    auto span = [&](size_t a, size_t b) {
        float ms = 0.0f;
        return hipEventElapsedTime(&ms, ev[a], ev[b]) == hipSuccess ? (double)ms : 0.0;
    };
    std::memset(tm, 0, sizeof(*tm));
    tm->ms_init = span(0, 1);
    for (int it = 0; it < I; ++it) {
        tm->ms_render_fitness += span(2 + 4 * it, 3 + 4 * it);
        tm->ms_pso += span(4 + 4 * it, 5 + 4 * it);
    }
    tm->ms_solve = span(0, n_ev - 1);
    tm->launches = I;
    return read_result(s, out, st);
}

int qpso_estimate_pose(qpso_ctx* ctx, uint32_t mesh_id, const qpso_cluster* cl, const qpso_params* p,
                       qpso_result* out) {
    qpso_solve* s = nullptr;
    int r = qpso_solve_create(ctx, mesh_id, cl, p, 0, &s);
    if (r) return r;
    {
        // a solve run once: its launches go straight to the stream (capturing
        // and instantiating a graph would cost more than replaying it saves)
        std::lock_guard<std::mutex> g(ctx->mu);
        r = hipSetDevice(ctx->device) == hipSuccess ? launch_solve(s, ctx->stream)
                                                   : set_err(ctx, QPSO_E_HIP, "hipSetDevice failed");
        if (!r) r = read_result(s, out, ctx->stream);
    }
    qpso_solve_destroy(s);
    return r;
}

// ---- the plain per-cluster call without device allocation -------------------
namespace {
// calls on one GPU share its scratch: they run one at a time
std::mutex& scratch_mutex(int device) {
    static std::mutex mu[64];
    return mu[device & 63];
}
}  // namespace

int qpso_estimate_pose_noalloc(qpso_ctx* ctx, uint32_t mesh_id, const qpso_cluster* cl, const qpso_params* p,
                               qpso_result* out) {
    if (!ctx || !out) return QPSO_E_INVAL;
    if (int v = validate(cl, p)) return set_err(ctx, v, "invalid cluster or parameters");
    if (p->offframe) return set_err(ctx, QPSO_E_INVAL, "the off-frame mode needs its own buffers: not allocation-free");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (!ctx->scratch) return set_err(ctx, QPSO_E_HIP, "no-allocation scratch unavailable on this device");
    auto mit = ctx->meshes.find(mesh_id);
    if (mit == ctx->meshes.end() || mit->second.released) return set_err(ctx, QPSO_E_NOTFOUND, "unknown mesh id");
    const int N = p->n_particles, I = p->n_iters, S = qpso_host::rand_stride(I);
    const int rw = cl->rect[2] - cl->rect[0], rh = cl->rect[3] - cl->rect[1];
    // the arena as a solve's, then its random table
    const ArenaLayout L = layout(N, I, rw, rh, false);
    const size_t rand_off = align_up(L.total, 256);
    const size_t n_rand = (size_t)S * N;
    if (rand_off + sizeof(float) * n_rand > QPSO_NOALLOC_SCRATCH_BYTES)
        return set_err(ctx, QPSO_E_NOMEM,
                       "solve needs " + std::to_string((rand_off + sizeof(float) * n_rand) >> 10) +
                           " KiB, more than the no-allocation scratch (" +
                           std::to_string(QPSO_NOALLOC_SCRATCH_BYTES >> 10) + " KiB)");
    std::lock_guard<std::mutex> sg(scratch_mutex(ctx->device));
    char* A = ctx->scratch;
    float* rand = reinterpret_cast<float*>(A + rand_off);
    QpsoSolveDev h;
    fill_desc(h, mit->second, cl, p, A, L, rand, false, ctx->big_face);
    QpsoSolveDev* d_desc = reinterpret_cast<QpsoSolveDev*>(A + L.desc);
    hipStream_t st = ctx->stream;
    // uploads from the caller's buffers and the stack: the descriptor, the
    // cluster rect of the depth frame, the random table in chunks
    HIP_TRY(ctx, hipMemcpyAsync(d_desc, &h, sizeof(h), hipMemcpyHostToDevice, st));
    HIP_TRY(ctx, hipMemcpy2DAsync(A + L.obs, sizeof(uint16_t) * rw,
                                  cl->depth + (size_t)cl->rect[1] * kCols + cl->rect[0], sizeof(uint16_t) * kCols,
                                  sizeof(uint16_t) * rw, rh, hipMemcpyHostToDevice, st));
    {
        qpso_host::Mt19937 mt(qpso_host::rand_seed(p->seed));
        float chunk[4096];
        for (size_t k = 0; k < n_rand; k += 4096) {
            const size_t n = std::min<size_t>(4096, n_rand - k);
            for (size_t i = 0; i < n; ++i) chunk[i] = (float)((double)mt.next() / 4294967296.0);
            HIP_TRY(ctx, hipMemcpyAsync(rand + k, chunk, sizeof(float) * n, hipMemcpyHostToDevice, st));
            HIP_TRY(ctx, hipStreamSynchronize(st));  // the chunk is reused
        }
    }
    HIP_TRY(ctx, qpso_launch::obs_rows(reinterpret_cast<const uint16_t*>(A + L.obs),
                                       reinterpret_cast<uint16_t*>(A + L.obs_rows), rw, rh, st));
    // InitQPSO clears the iteration, flags, counters and carried-pixel counts
    HIP_TRY(ctx, qpso_launch::init(d_desc, 1, st));
    const bool win = qpso_launch::rf_needs_windows(ctx->rf_cfg, h.V_eff);
    for (int it = 0; it < I; ++it) {
        HIP_TRY(ctx, qpso_launch::render_fitness(ctx->rf_cfg, win, d_desc, nullptr, N, st));
        HIP_TRY(ctx, qpso_launch::pso_step(d_desc, 1, st));
    }
    float gb[8];
    unsigned long long c[3];
    int32_t flags = 0;
    HIP_TRY(ctx, hipMemcpyAsync(gb, h.gbest, sizeof(gb), hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipMemcpyAsync(c, h.counters, sizeof(c), hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipMemcpyAsync(&flags, h.flags, sizeof(flags), hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    std::memcpy(out->pose, gb, sizeof(float) * 7);
    out->fit = gb[7];
    out->raster_px = c[0];
    out->visit_px = c[1];
    out->n_degenerate = (uint32_t)c[2];
    out->status = (flags & QPSO_FLAG_LAUNCH) ? QPSO_E_HIP : (flags & QPSO_FLAG_STALE_OVERFLOW) ? QPSO_E_STALE : QPSO_OK;
    out->latency_ms = 0.0;
    if (out->status == QPSO_E_STALE) return set_err(ctx, QPSO_E_STALE, "stale-pixel list overflow");
    if (out->status != QPSO_OK) return set_err(ctx, out->status, "render kernel launched without mesh windows");
    return QPSO_OK;
}

int qpso_solve_get_trace(qpso_solve* s, qpso_trace* tr) {
    if (!s || !tr) return QPSO_E_INVAL;
    qpso_ctx* ctx = s->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const int N = s->h.N, I = s->h.I;
    const char* A = s->arena;
    if ((tr->fit_trace || tr->gbest_trace || tr->argmin_trace) && !s->trace)
        return set_err(ctx, QPSO_E_INVAL, "solve was created without traces");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (tr->fit_trace) HIP_TRY(ctx, copy_on(ctx->stream, tr->fit_trace, A + s->L.fit_tr, sizeof(float) * (size_t)I * N, hipMemcpyDeviceToHost));
    if (tr->gbest_trace) HIP_TRY(ctx, copy_on(ctx->stream, tr->gbest_trace, A + s->L.gbest_tr, sizeof(float) * 8 * (size_t)I, hipMemcpyDeviceToHost));
    if (tr->argmin_trace) HIP_TRY(ctx, copy_on(ctx->stream, tr->argmin_trace, A + s->L.argmin_tr, sizeof(int32_t) * (size_t)I, hipMemcpyDeviceToHost));
    if (tr->pos) HIP_TRY(ctx, copy_on(ctx->stream, tr->pos, A + s->L.pos, sizeof(float) * 7 * N, hipMemcpyDeviceToHost));
    if (tr->vel) HIP_TRY(ctx, copy_on(ctx->stream, tr->vel, A + s->L.vel, sizeof(float) * 7 * N, hipMemcpyDeviceToHost));
    if (tr->pbest) HIP_TRY(ctx, copy_on(ctx->stream, tr->pbest, A + s->L.pbest, sizeof(float) * 7 * N, hipMemcpyDeviceToHost));
    if (tr->pbest_fit) HIP_TRY(ctx, copy_on(ctx->stream, tr->pbest_fit, A + s->L.pbest_fit, sizeof(float) * N, hipMemcpyDeviceToHost));
    if (tr->frames && s->h.wrap) {  // the buffers themselves
        HIP_TRY(ctx, copy_on(ctx->stream, tr->frames, s->h.frames, sizeof(int32_t) * (size_t)N * kFrame, hipMemcpyDeviceToHost));
    } else if (tr->frames) {
        // rebuild the reference's per-particle buffers after the last fitness pass:
        // empty except for the carried (never visited) pixels
        int32_t it = 0;
        HIP_TRY(ctx, copy_on(ctx->stream, &it, A + s->L.iter, sizeof(int32_t), hipMemcpyDeviceToHost));
        const int par = it & 1;
        std::vector<int32_t> cnt(N);
        std::vector<QpsoStale> ent((size_t)N * kStaleCap);
        HIP_TRY(ctx, copy_on(ctx->stream, cnt.data(), A + (par ? s->L.cnt1 : s->L.cnt0), sizeof(int32_t) * N, hipMemcpyDeviceToHost));
        HIP_TRY(ctx, copy_on(ctx->stream, ent.data(), A + (par ? s->L.stale1 : s->L.stale0), sizeof(QpsoStale) * ent.size(), hipMemcpyDeviceToHost));
        for (int p = 0; p < N; ++p) {
            int32_t* fr = tr->frames + (size_t)p * kFrame;
            std::fill(fr, fr + kFrame, (int32_t)QPSO_EMPTY);
            for (int k = 0; k < cnt[p]; ++k) {
                const QpsoStale e = ent[(size_t)p * kStaleCap + k];
                fr[(e.pos >> 16) * kCols + (e.pos & 0xffff)] = e.val;
            }
        }
    }
    return QPSO_OK;
}

int qpso_solve_eval_poses(qpso_solve* s, int n, const float* poses, float* fit, int32_t* aabb,
                          int32_t* frames) {
    if (!s || !poses || n <= 0 || n > s->h.N) return QPSO_E_INVAL;
    qpso_ctx* ctx = s->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int N = s->h.N;
    char* A = s->arena;
    for (int d = 0; d < 7; ++d)
        HIP_TRY(ctx, hipMemcpyAsync(A + s->L.pos + sizeof(float) * (size_t)d * N, poses + (size_t)d * n,
                                    sizeof(float) * n, hipMemcpyHostToDevice, st));
    HIP_TRY(ctx, hipMemsetAsync(A + s->L.iter, 0, sizeof(int32_t), st));
    HIP_TRY(ctx, hipMemsetAsync(A + s->L.flags, 0, sizeof(int32_t), st));
    HIP_TRY(ctx, hipMemsetAsync(A + s->L.cnt0, 0, sizeof(int32_t) * N, st));
    HIP_TRY(ctx, hipMemsetAsync(A + s->L.counters, 0, sizeof(unsigned long long) * 3, st));
    int32_t* d_frames = nullptr;
    int32_t* d_aabb = nullptr;
    QpsoSolveDev* d_tmp = nullptr;
    auto cleanup = [&]() {
        hipStreamSynchronize(st);
        if (d_frames) hipFree(d_frames);
        if (d_aabb) hipFree(d_aabb);
        if (d_tmp) hipFree(d_tmp);
    };
    QpsoSolveDev tmp = s->h;
    if (frames) {
        if (hipMalloc(&d_frames, sizeof(int32_t) * (size_t)n * kFrame) != hipSuccess) { cleanup(); return QPSO_E_NOMEM; }
        hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(d_frames), QPSO_EMPTY, (size_t)n * kFrame, st);
        tmp.dbg_frames = d_frames;
    }
    if (hipMalloc(&d_aabb, sizeof(int32_t) * 4 * (size_t)n) != hipSuccess) { cleanup(); return QPSO_E_NOMEM; }
    tmp.dbg_aabb = d_aabb;
    if (hipMalloc(&d_tmp, sizeof(QpsoSolveDev)) != hipSuccess) { cleanup(); return QPSO_E_NOMEM; }
    hipError_t e = hipMemcpyAsync(d_tmp, &tmp, sizeof(tmp), hipMemcpyHostToDevice, st);
    if (s->h.wrap) {  // the particles' own buffers start empty too
        if (e == hipSuccess) e = qpso_launch::fill_frames(s->h.frames, (size_t)n * kFrame, st);
        if (e == hipSuccess) e = qpso_launch::render_fitness_wrap(d_tmp, nullptr, n, st);
    } else if (e == hipSuccess) {
        // frames come from their own instantiation (it handles every mesh)
        e = qpso_launch::render_fitness(s->rf_cfg, frames ? 2 : needs_windows(s, s->rf_cfg) ? 1 : 0, d_tmp, nullptr, n, st);
    }
    if (e == hipSuccess && fit) e = hipMemcpyAsync(fit, A + s->L.fit, sizeof(float) * n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && aabb) e = hipMemcpyAsync(aabb, d_aabb, sizeof(int32_t) * 4 * n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && frames)
        e = hipMemcpyAsync(frames, d_frames, sizeof(int32_t) * (size_t)n * kFrame, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    cleanup();
    if (e != hipSuccess) return set_err(ctx, QPSO_E_HIP, std::string("eval_poses: ") + hipGetErrorString(e));
    return QPSO_OK;
}

// ---- autotuner --------------------------------------------------------------
int qpso_autotune(qpso_ctx* ctx, int n, qpso_solve* const* solves, int iters, qpso_tune_entry* table, int cap,
                  int* n_out) {
    if (!ctx || n <= 0 || !solves || iters <= 0 || (cap > 0 && !table)) return QPSO_E_INVAL;
    for (int i = 0; i < n; ++i)
        if (!solves[i] || solves[i]->ctx != ctx) return QPSO_E_INVAL;
    if (any_wrap(solves, n)) return set_err(ctx, QPSO_E_INVAL, "the reference-faithful off-frame mode has one kernel: nothing to tune");
    static const int kBigFaces[] = {48, 96, 128, 160};
    const int n_cfg = qpso_launch::rf_num_configs();
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    // the solves run together as in a batch: descriptors with the candidate
    // threshold, particles of the largest meshes first
    std::vector<QpsoSolveDev> descs(n);
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) { descs[i] = solves[i]->h; order[i] = i; }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return descs[a].F_eff > descs[b].F_eff; });
    std::vector<QpsoItem> items;
    for (int i : order)
        for (int q = 0; q < descs[i].N; ++q) items.push_back(QpsoItem{i, q});
    QpsoSolveDev* d_descs = nullptr;
    QpsoItem* d_items = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        hipStreamSynchronize(st);
        if (d_descs) hipFree(d_descs);
        if (d_items) hipFree(d_items);
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
    };
    if (hipMalloc(&d_descs, sizeof(QpsoSolveDev) * n) != hipSuccess ||
        hipMalloc(&d_items, sizeof(QpsoItem) * items.size()) != hipSuccess) {
        cleanup();
        return set_err(ctx, QPSO_E_NOMEM, "autotune scratch allocation failed");
    }
    hipError_t e = copy_on(st, d_items, items.data(), sizeof(QpsoItem) * items.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    auto run = [&](int cfg) {
        hipError_t r = qpso_launch::init(d_descs, n, st);
        for (int it = 0; it < iters && r == hipSuccess; ++it) {
            r = qpso_launch::render_fitness(cfg, any_needs_windows(solves, n, cfg), d_descs, d_items, (int)items.size(), st);
            if (r == hipSuccess) r = qpso_launch::pso_step(d_descs, n, st);
        }
        return r;
    };
    int best_cfg = 0, best_big = kBigFaces[0], k = 0;
    float best_ms = 0.0f;
    constexpr int kNBig = (int)(sizeof(kBigFaces) / sizeof(kBigFaces[0]));
    for (int b = 0; b < kNBig && e == hipSuccess; ++b) {
        for (auto& d : descs) d.big_face = kBigFaces[b];
        e = copy_on(st, d_descs, descs.data(), sizeof(QpsoSolveDev) * n, hipMemcpyHostToDevice);
        if (b == 0 && e == hipSuccess) e = run(0);  // warm-up
        for (int c = 0; c < n_cfg && e == hipSuccess; ++c) {
            float ms = 0.0f;
            for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {  // the faster of two runs
                e = hipEventRecord(e0, st);
                if (e == hipSuccess) e = run(c);
                if (e == hipSuccess) e = hipEventRecord(e1, st);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                float t = 0.0f;
                if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
                ms = rep == 0 ? t : std::min(ms, t);
            }
            if (e != hipSuccess) break;
            if (k < cap) {
                table[k] = qpso_tune_entry{c, qpso_launch::rf_config_threads(c), qpso_launch::rf_config_tile(c),
                                           kBigFaces[b], (double)ms};
            }
            if (k == 0 || ms < best_ms) { best_ms = ms; best_cfg = c; best_big = kBigFaces[b]; }
            ++k;
        }
    }
    cleanup();
    if (e != hipSuccess) return set_err(ctx, QPSO_E_HIP, std::string("autotune: ") + hipGetErrorString(e));
    if (n_out) *n_out = k;
    // apply the choice to these solves (cached graphs hold the old one) and
    // make it the default of solves and servers created from now on
    invalidate_plan(ctx);
    ctx->rf_cfg = best_cfg;
    ctx->big_face = best_big;
    for (int i = 0; i < n; ++i) {
        qpso_solve* s = solves[i];
        s->rf_cfg = best_cfg;
        s->h.big_face = best_big;
        drop_solve_graph(s);
        HIP_TRY(ctx, copy_on(ctx->stream, s->d_desc, &s->h, sizeof(QpsoSolveDev), hipMemcpyHostToDevice));
    }
    return QPSO_OK;
}

int qpso_set_tuning(qpso_ctx* ctx, int config, int big_face) {
    if (!ctx || config < 0 || config >= qpso_launch::rf_num_configs() || big_face < 1) return QPSO_E_INVAL;
    ctx->rf_cfg = config;
    ctx->big_face = big_face;
    return QPSO_OK;
}

// ---- batched form ---------------------------------------------------------
int qpso_solve_run_batch(qpso_ctx* ctx, int n, qpso_solve* const* solves, qpso_result* out) {
    if (!ctx || n <= 0 || !solves) return QPSO_E_INVAL;
    for (int i = 0; i < n; ++i)
        if (!solves[i] || solves[i]->ctx != ctx) return QPSO_E_INVAL;
    if (any_wrap(solves, n)) return set_err(ctx, QPSO_E_INVAL, "the reference-faithful off-frame mode runs in plain solves only");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    auto& P = ctx->plan;
    // the batch runs with its first solve's configuration
    const int cfg = solves[0]->rf_cfg;
    const bool same = P.exec && P.rf_cfg == cfg && (int)P.solves.size() == n &&
                      std::equal(P.solves.begin(), P.solves.end(), solves);
    if (!same) {
        invalidate_plan(ctx);
        P.rf_cfg = cfg;
        P.solves.assign(solves, solves + n);
        std::vector<QpsoSolveDev> descs(n);
        std::vector<int> order(n);
        for (int i = 0; i < n; ++i) { descs[i] = solves[i]->h; order[i] = i; P.max_iters = std::max(P.max_iters, solves[i]->h.I); }
        // the largest meshes' workgroups first (a workgroup's time grows with
        // its mesh), so the launch ends on short ones
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return descs[a].F_eff > descs[b].F_eff; });
        std::vector<QpsoItem> items;
        for (int i : order)
            for (int q = 0; q < descs[i].N; ++q) items.push_back(QpsoItem{i, q});
        P.n_items = (int)items.size();
        HIP_TRY(ctx, hipMalloc(&P.d_descs, sizeof(QpsoSolveDev) * n));
        HIP_TRY(ctx, hipMalloc(&P.d_items, sizeof(QpsoItem) * items.size()));
        HIP_TRY(ctx, copy_on(ctx->stream, P.d_descs, descs.data(), sizeof(QpsoSolveDev) * n, hipMemcpyHostToDevice));
        HIP_TRY(ctx, copy_on(ctx->stream, P.d_items, items.data(), sizeof(QpsoItem) * items.size(), hipMemcpyHostToDevice));
        hipGraph_t graph = nullptr;
        HIP_TRY(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t e = qpso_launch::init(P.d_descs, n, st);
        for (int it = 0; it < P.max_iters && e == hipSuccess; ++it) {
            e = qpso_launch::render_fitness(cfg, any_needs_windows(solves, n, cfg), P.d_descs, P.d_items, P.n_items, st);
            if (e == hipSuccess) e = qpso_launch::pso_step(P.d_descs, n, st);
        }
        hipError_t e2 = hipStreamEndCapture(st, &graph);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (graph) hipGraphDestroy(graph);
            invalidate_plan(ctx);
            return set_err(ctx, QPSO_E_HIP, "batch graph capture failed");
        }
        e = hipGraphInstantiate(&P.exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e != hipSuccess) { invalidate_plan(ctx); return set_err(ctx, QPSO_E_HIP, "graph instantiate failed"); }
    }
    HIP_TRY(ctx, hipGraphLaunch(P.exec, st));
    int rc = QPSO_OK;
    for (int i = 0; i < n; ++i) {
        int r = read_result(solves[i], out ? &out[i] : nullptr, st);
        if (r && !rc) rc = r;
    }
    return rc;
}

}  // extern "C"

// ---- serving: continuous batching under an HBM budget -----------------------
namespace {

// The serving loop's device: the scheduler (qpso_sched.h) decides what is
// resident and which lane a request joins; this enqueues the work.  One device
// step of a lane is one PSO iteration of each of its solves: a fused
// render+fitness launch over all their particles and one PSO launch.  The two
// lanes have their own streams and work tables; they run side by side, each
// stepped as soon as it has fewer than kInFlight steps queued (an event
// query), so the GPU overlaps one lane's step tail with the other's work and
// a lane of light solves advances at its own pace.  Within a lane the largest
// meshes' workgroups start first.  The host counts iterations itself, so
// steps are enqueued without a round trip.  A solve's last step is followed
// on its lane by the copies of its result words into pinned host memory and
// an event; its arena goes back to the context's pool in the lane's order.
// Nothing in the loop waits for the device except a failure path.
struct HipDev {
    using Solve = qpso_solve;
    struct Job {
        uint32_t mesh_id = 0;
        qpso_params p{};
        qpso_cluster cl{};
        std::vector<uint16_t> depth;  // the caller's frame, copied at submit
    };
    static constexpr int kInFlight = 2;
    // Two lanes (requests join them by size).  The lanes run side by side only on
    // hardware queues of their own: the process has four by default, handed
    // to the least used stream first; the context's stream holds one and,
    // once a graph has been launched, the runtime another.
    static constexpr int kLanes = 2;
    struct Lane {
        hipStream_t st = nullptr;
        QpsoSolveDev* d_descs = nullptr;
        QpsoItem* d_items = nullptr;
        size_t cap_descs = 0, cap_items = 0;
        std::vector<QpsoSolveDev> descs;
        std::vector<QpsoItem> items;
        hipEvent_t ev[kInFlight] = {};
        uint64_t launched = 0;
    };
    qpso_ctx* ctx;
    int rf_cfg, big_face;  // the context's tuning when the server was created
    Lane lanes[kLanes];
    // A finished solve's last step: its result words, copied into pinned
    // host memory on its lane right after that step, and the event after
    // the copies.  Reading a result then needs no call into the runtime.
    struct End {
        hipEvent_t ev;
        ResultWords* words;
    };
    std::unordered_map<qpso_solve*, End> ends;
    std::vector<hipEvent_t> spare;       // events to reuse
    std::vector<ResultWords*> free_words;  // pinned result slots to reuse
    std::vector<End> draining;             // a cancelled solve's copies may still be queued
    std::vector<ResultWords*> blocks;      // their allocations (kBlock slots each)
    static constexpr size_t kBlock = 256;

    explicit HipDev(qpso_ctx* c) : ctx(c), rf_cfg(c->rf_cfg), big_face(c->big_face) {}

    int open() {
        if (hipSetDevice(ctx->device) != hipSuccess) return QPSO_E_HIP;
        for (auto& l : lanes) {
            if (hipStreamCreateWithFlags(&l.st, hipStreamNonBlocking) != hipSuccess) return QPSO_E_HIP;
            for (auto& e : l.ev)
                if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return QPSO_E_HIP;
        }
        return QPSO_OK;
    }
    void close() {
        hipSetDevice(ctx->device);
        for (auto& l : lanes) {
            if (l.st) hipStreamSynchronize(l.st);
            for (auto& e : l.ev)
                if (e) hipEventDestroy(e);
            if (l.d_descs) hipFree(l.d_descs);
            if (l.d_items) hipFree(l.d_items);
            if (l.st) hipStreamDestroy(l.st);
            l = Lane{};
        }
        for (auto& kv : ends) spare.push_back(kv.second.ev);
        ends.clear();
        for (auto& d : draining) spare.push_back(d.ev);
        draining.clear();
        for (hipEvent_t e : spare) hipEventDestroy(e);
        spare.clear();
        free_words.clear();
        for (ResultWords* b : blocks) hipHostFree(b);
        blocks.clear();
    }
    void on_thread_start() { hipSetDevice(ctx->device); }
    void on_thread_end() {}

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    }

    int create(Job& j, qpso_solve*& s) {
        j.cl.depth = j.depth.data();
        const int rc = qpso_solve_create(ctx, j.mesh_id, &j.cl, &j.p, 0, &s);
        if (rc) return rc;
        s->rf_cfg = rf_cfg;  // the server's tuning, not the context's current one
        s->h.big_face = big_face;
        return QPSO_OK;
    }
    int init(qpso_solve* s, int k) {
        return qpso_launch::init(s->d_desc, 1, lanes[k].st) == hipSuccess ? QPSO_OK : QPSO_E_HIP;
    }
    // the lane's work tables: its solves' descriptors and one item per
    // particle, the largest meshes' particles first
    int rebuild(Lane& l, const std::vector<qpso_solve*>& ss) {
        l.descs.clear();
        l.items.clear();
        std::vector<int> order(ss.size());
        for (size_t i = 0; i < ss.size(); ++i) {
            l.descs.push_back(ss[i]->h);
            order[i] = (int)i;
        }
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return l.descs[a].F_eff > l.descs[b].F_eff; });
        for (int i : order)
            for (int q = 0; q < l.descs[i].N; ++q) l.items.push_back(QpsoItem{(int32_t)i, q});
        if (l.descs.size() > l.cap_descs || l.items.size() > l.cap_items) {
            // the old tables may still be read by queued steps
            if (hipStreamSynchronize(l.st) != hipSuccess) return QPSO_E_HIP;
            if (l.descs.size() > l.cap_descs) {
                if (l.d_descs) hipFree(l.d_descs);
                l.d_descs = nullptr;
                l.cap_descs = 0;
                if (hipMalloc(&l.d_descs, sizeof(QpsoSolveDev) * l.descs.size() * 2) != hipSuccess) return QPSO_E_NOMEM;
                l.cap_descs = l.descs.size() * 2;
            }
            if (l.items.size() > l.cap_items) {
                if (l.d_items) hipFree(l.d_items);
                l.d_items = nullptr;
                l.cap_items = 0;
                if (hipMalloc(&l.d_items, sizeof(QpsoItem) * l.items.size() * 2) != hipSuccess) return QPSO_E_NOMEM;
                l.cap_items = l.items.size() * 2;
            }
        }
        // stream-ordered behind the steps that read the old tables
        if (hipMemcpyAsync(l.d_descs, l.descs.data(), sizeof(QpsoSolveDev) * l.descs.size(), hipMemcpyHostToDevice,
                           l.st) != hipSuccess ||
            hipMemcpyAsync(l.d_items, l.items.data(), sizeof(QpsoItem) * l.items.size(), hipMemcpyHostToDevice,
                           l.st) != hipSuccess)
            return QPSO_E_HIP;
        return QPSO_OK;
    }
    // the lane's oldest queued step of the last kInFlight has completed
    bool ready(int k) {
        const Lane& l = lanes[k];
        return l.launched < (uint64_t)kInFlight || hipEventQuery(l.ev[l.launched % kInFlight]) != hipErrorNotReady;
    }
    // A step with no more workgroups than CUs is latency-bound: it takes the
    // widest configuration (one 1024-thread workgroup per CU); fuller steps
    // take the server's tuned one.  Results do not depend on it.
    int step(int k, const std::vector<qpso_solve*>& ss, bool changed, int64_t total) {
        Lane& l = lanes[k];
        if (changed)
            if (const int rc = rebuild(l, ss)) return rc;
        const int cfg = total <= (int64_t)std::max(ctx->n_cu, 1) ? 0 : rf_cfg;
        bool win = false;
        for (auto* s : ss) win |= needs_windows(s, cfg);
        if (l.launched >= (uint64_t)kInFlight && hipEventSynchronize(l.ev[l.launched % kInFlight]) != hipSuccess)
            return QPSO_E_HIP;
        if (qpso_launch::render_fitness(cfg, win, l.d_descs, l.d_items, (int)l.items.size(), l.st) != hipSuccess ||
            qpso_launch::pso_step(l.d_descs, (int)l.descs.size(), l.st) != hipSuccess ||
            hipEventRecord(l.ev[l.launched % kInFlight], l.st) != hipSuccess)
            return QPSO_E_HIP;
        ++l.launched;
        return QPSO_OK;
    }
    int mark(int k, qpso_solve* s) {
        // slots whose copies have landed can be used again
        for (size_t i = 0; i < draining.size();) {
            if (hipEventQuery(draining[i].ev) == hipErrorNotReady) {
                ++i;
                continue;
            }
            spare.push_back(draining[i].ev);
            free_words.push_back(draining[i].words);
            draining[i] = draining.back();
            draining.pop_back();
        }
        if (free_words.empty()) {
            ResultWords* b = nullptr;
            if (hipHostMalloc(reinterpret_cast<void**>(&b), sizeof(ResultWords) * kBlock) != hipSuccess) return QPSO_E_NOMEM;
            blocks.push_back(b);
            for (size_t i = 0; i < kBlock; ++i) free_words.push_back(b + i);
        }
        End end{nullptr, free_words.back()};
        free_words.pop_back();
        if (!spare.empty()) {
            end.ev = spare.back();
            spare.pop_back();
        } else if (hipEventCreateWithFlags(&end.ev, hipEventDisableTiming) != hipSuccess) {
            free_words.push_back(end.words);
            return QPSO_E_HIP;
        }
        ends[s] = end;
        if (copy_result(s, end.words, lanes[k].st) != hipSuccess) return QPSO_E_HIP;
        return hipEventRecord(end.ev, lanes[k].st) == hipSuccess ? QPSO_OK : QPSO_E_HIP;
    }
    int done(qpso_solve* s) {
        auto it = ends.find(s);
        if (it == ends.end()) return QPSO_E_HIP;
        const hipError_t e = hipEventQuery(it->second.ev);
        return e == hipSuccess ? 1 : e == hipErrorNotReady ? 0 : QPSO_E_HIP;
    }
    int sync(int k) { return hipStreamSynchronize(lanes[k].st) == hipSuccess ? QPSO_OK : QPSO_E_HIP; }
    // after done(): the result words are in host memory
    int read(qpso_solve* s, qpso_result* out) {
        auto it = ends.find(s);
        if (it == ends.end()) return QPSO_E_HIP;
        return result_from(*it->second.words, out);
    }
    // the solve's work all lies on lane k: its memory is released in that
    // lane's order, with no wait
    void destroy(qpso_solve* s, int k) {
        auto it = ends.find(s);
        if (it != ends.end()) {
            if (hipEventQuery(it->second.ev) == hipErrorNotReady) {
                draining.push_back(it->second);
            } else {
                spare.push_back(it->second.ev);
                free_words.push_back(it->second.words);
            }
            ends.erase(it);
        }
        std::lock_guard<std::mutex> g(ctx->mu);
        hipSetDevice(ctx->device);
        solve_release(s, lanes[k].st);
    }
    void idle_wait() { std::this_thread::sleep_for(std::chrono::microseconds(20)); }
};

}  // namespace

struct qpso_server {
    HipDev dev;  // declared first: outlives the scheduler that drives it
    qpso_sched::Server<HipDev> core;
    qpso_server(qpso_ctx* ctx, uint64_t budget) : dev(ctx), core(dev, budget) {}
};

extern "C" {

int qpso_server_create(qpso_ctx* ctx, uint64_t hbm_budget_bytes, qpso_server** out) {
    if (!ctx || !out || hbm_budget_bytes == 0) return QPSO_E_INVAL;
    *out = nullptr;
    auto* sv = new qpso_server(ctx, hbm_budget_bytes);
    if (sv->dev.open() != QPSO_OK) {
        sv->dev.close();
        delete sv;
        return QPSO_E_HIP;
    }
    sv->core.start();
    *out = sv;
    return QPSO_OK;
}

int qpso_server_set_slots(qpso_server* sv, int64_t max_resident_particles) {
    if (!sv) return QPSO_E_INVAL;
    return sv->core.set_slots(max_resident_particles);
}

int qpso_server_set_timeout(qpso_server* sv, double timeout_ms) {
    if (!sv) return QPSO_E_INVAL;
    return sv->core.set_timeout(timeout_ms * 1e-3);
}

int qpso_server_submit(qpso_server* sv, uint32_t mesh_id, const qpso_cluster* cl, const qpso_params* p,
                       uint64_t* ticket) {
    if (!sv || !ticket) return QPSO_E_INVAL;
    if (int v = validate(cl, p)) return v;
    if (p->offframe) return QPSO_E_INVAL;  // the off-frame mode runs in plain solves only
    qpso_sched::Server<HipDev>::Request r;
    r.job.mesh_id = mesh_id;
    r.job.p = *p;
    r.job.cl = *cl;
    r.job.depth.assign(cl->depth, cl->depth + kFrame);
    r.bytes = qpso_solve_bytes(p, cl->rect);
    const int S = qpso_host::rand_stride(p->n_iters);
    r.table = qpso_sched::TableKey(qpso_host::rand_seed(p->seed), p->n_particles, S);
    r.table_bytes = sizeof(float) * (uint64_t)S * (uint64_t)p->n_particles;
    r.particles = p->n_particles;
    r.iters = p->n_iters;
    {
        // a step's work: particles x faces rendered (an unknown mesh fails at
        // admission)
        qpso_ctx* ctx = sv->dev.ctx;
        std::lock_guard<std::mutex> g(ctx->mu);
        auto mit = ctx->meshes.find(mesh_id);
        const int64_t F = mit == ctx->meshes.end() ? 1
                          : p->face_cap > 0 ? std::min<int64_t>(mit->second.F, p->face_cap)
                                            : (int64_t)mit->second.F;
        r.cost = (int64_t)p->n_particles * std::max<int64_t>(F, 1);
    }
    return sv->core.submit(std::move(r), ticket);
}

int qpso_server_poll(qpso_server* sv, uint64_t ticket, qpso_result* out) {
    if (!sv) return QPSO_E_INVAL;
    return sv->core.poll(ticket, out);
}

int qpso_server_wait(qpso_server* sv, uint64_t ticket, qpso_result* out) {
    if (!sv) return QPSO_E_INVAL;
    return sv->core.wait(ticket, out);
}

int qpso_server_cancel(qpso_server* sv, uint64_t ticket) {
    if (!sv) return QPSO_E_INVAL;
    return sv->core.cancel(ticket);
}

int qpso_server_stats_get(qpso_server* sv, qpso_server_stats* st) {
    if (!sv || !st) return QPSO_E_INVAL;
    sv->core.stats(st);
    return QPSO_OK;
}

int qpso_server_destroy(qpso_server* sv) {
    if (!sv) return QPSO_E_INVAL;
    sv->core.shutdown();
    sv->dev.close();
    delete sv;
    return QPSO_OK;
}

}  // extern "C"
