// host_asan.cpp -- the host-only code of the library (q-pso_amd/csrc:
// qpso_tables.h, qpso_roi.h, qpso_io.cpp, qpso_scene.cpp) and the CPU oracle
// (oracle/qpso_oracle.c) exercised under AddressSanitizer and
// UndefinedBehaviorSanitizer, built with the flags the library uses.  Built
// and run by tests/test_sanitizers.py.  Each check prints a line; exits
// non-zero on the first failure (a sanitizer report aborts earlier).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "qpso_io.h"
#include "qpso_oracle.h"
#include "qpso_roi.h"
#include "qpso_scene.h"
#include "qpso_tables.h"

namespace {

int g_fail = 0;
#define CHECK(cond)                                                          \
    do {                                                                     \
        if (!(cond)) {                                                       \
            std::printf("  check failed at line %d: %s\n", __LINE__, #cond); \
            ++g_fail;                                                        \
            return false;                                                    \
        }                                                                    \
    } while (0)

std::vector<float> icosphere(int subdiv) {
    const int F = qpso_mesh_icosphere(subdiv, 50.f, 35.f, 25.f, nullptr);
    std::vector<float> m((size_t)9 * F);
    qpso_mesh_icosphere(subdiv, 50.f, 35.f, 25.f, m.data());
    return m;
}

std::vector<float> uv_mesh(int sl, int st, int boxes, uint64_t seed) {
    const int F = qpso_mesh_uv_ellipsoid(sl, st, 60.f, 40.f, 30.f, boxes, 5.f, 9.f, seed, nullptr);
    std::vector<float> m((size_t)9 * F);
    qpso_mesh_uv_ellipsoid(sl, st, 60.f, 40.f, 30.f, boxes, 5.f, 9.f, seed, m.data());
    return m;
}

// Faces of a SoA mesh in a new order (perm[f] = old face index).
std::vector<float> permuted(const std::vector<float>& m, const std::vector<int>& perm) {
    const size_t F = m.size() / 9;
    std::vector<float> o(m.size());
    for (size_t f = 0; f < F; ++f)
        for (size_t k = 0; k < 9; ++k) o[k * F + f] = m[k * F + perm[f]];
    return o;
}

bool check_tables(const std::vector<float>& soa) {
    const uint32_t F = (uint32_t)(soa.size() / 9);
    std::vector<float> verts;
    std::vector<QpsoFaceRef> order;
    std::vector<uint32_t> v_prefix;
    std::vector<int32_t> runs;
    qpso_host::build_mesh_tables(soa.data(), F, verts, order, v_prefix, runs);
    const size_t V = verts.size() / 3;
    CHECK(order.size() == F && v_prefix.size() == F && V >= 3 && V <= 3 * (size_t)F);
    int32_t max_id = -1;
    for (uint32_t f = 0; f < F; ++f) {
        CHECK(order[f].face == (int32_t)f);
        const int32_t ids[3] = {order[f].v0, order[f].v1, order[f].v2};
        for (int j = 0; j < 3; ++j) {
            CHECK(ids[j] >= 0 && (size_t)ids[j] < V);
            for (int c = 0; c < 3; ++c)  // the vertex holds the corner's coordinates bit for bit
                CHECK(std::memcmp(&verts[(size_t)c * V + ids[j]], &soa[(size_t)c * 3 * F + f + (size_t)j * F], 4) == 0);
            max_id = std::max(max_id, ids[j]);
        }
        CHECK(v_prefix[f] == (uint32_t)max_id + 1);  // first-use numbering: a prefix per face prefix
    }
    if (!runs.empty()) {
        CHECK(runs.size() % 3 == 0 && runs[0] == 0);
        for (size_t r = 0; r < runs.size(); r += 3) {
            const int32_t f0 = runs[r], lo = runs[r + 1], hi = runs[r + 2];
            const int32_t f1 = r + 3 < runs.size() ? runs[r + 3] : (int32_t)F;
            CHECK(f0 < f1 && lo < hi && hi - lo <= qpso_host::kRunVerts);
            for (int32_t f = f0; f < f1; ++f) {
                const int32_t ids[3] = {order[f].v0, order[f].v1, order[f].v2};
                for (int j = 0; j < 3; ++j) CHECK(ids[j] >= lo && ids[j] < hi);
            }
        }
    }
    std::vector<float> corners;
    qpso_host::build_corner_table(soa.data(), F, order, corners);
    CHECK(corners.size() == (size_t)F * 12);
    for (uint32_t f = 0; f < F; f += 7)
        for (int j = 0; j < 3; ++j) {
            int32_t id;
            std::memcpy(&id, &corners[(size_t)f * 12 + j * 4 + 3], 4);
            const int32_t want[3] = {order[f].v0, order[f].v1, order[f].v2};
            CHECK(id == want[j]);
            for (int c = 0; c < 3; ++c) CHECK(corners[(size_t)f * 12 + j * 4 + c] == soa[(size_t)c * 3 * F + f + (size_t)j * F]);
        }
    return true;
}

bool mesh_tables() {
    for (int s = 0; s <= 3; ++s)
        if (!check_tables(icosphere(s))) return false;
    std::vector<float> uv = uv_mesh(120, 90, 4, 3);
    if (!check_tables(uv)) return false;
    // the same faces shuffled: runs no longer hold (one face far from its neighbours)
    std::vector<int> perm(uv.size() / 9);
    for (size_t i = 0; i < perm.size(); ++i) perm[i] = (int)i;
    std::mt19937 rng(5);
    std::shuffle(perm.begin(), perm.end(), rng);
    if (!check_tables(permuted(uv, perm))) return false;
    // every face the same triangle: three vertices
    std::vector<float> same((size_t)9 * 50);
    for (size_t f = 0; f < 50; ++f)
        for (int k = 0; k < 9; ++k) same[(size_t)k * 50 + f] = (float)(k + 1);
    return check_tables(same);
}

bool crop_and_rows() {
    std::mt19937 rng(9);
    std::vector<uint16_t> depth((size_t)QPSO_ROWS * QPSO_COLS);
    for (auto& d : depth) d = (rng() % 4 == 0) ? 0 : (uint16_t)(300 + rng() % 2000);
    const int32_t rects[][4] = {{0, 0, 640, 480}, {639, 479, 640, 480}, {0, 0, 1, 1}, {17, 33, 250, 301}, {600, 10, 640, 11}};
    for (const auto& r : rects) {
        std::vector<uint16_t> crop, rows;
        qpso_host::build_crop(depth.data(), QPSO_COLS, r, crop, rows);
        const int rw = r[2] - r[0], rh = r[3] - r[1];
        CHECK(crop.size() == (size_t)rw * rh && rows.size() == (size_t)(rw + 1) * rh);
        for (int y = 0; y < rh; ++y) {
            int n = 0;
            CHECK(rows[(size_t)y * (rw + 1)] == 0);
            for (int x = 0; x < rw; ++x) {
                const uint16_t v = depth[(size_t)(r[1] + y) * QPSO_COLS + r[0] + x];
                CHECK(crop[(size_t)y * rw + x] == v);
                n += v != 0;
                CHECK(rows[(size_t)y * (rw + 1) + x + 1] == n);
            }
        }
    }
    return true;
}

bool random_table() {
    std::vector<float> t;
    qpso_host::rand_table(qpso_host::rand_seed(0), 5, t);  // GSL: seed 0 is 4357, first output 4293858116
    CHECK(t.size() == 5 && t[0] == (float)(4293858116.0 / 4294967296.0));
    std::vector<float> o(5);
    or_rand_table(0, o.data(), 5);
    CHECK(std::memcmp(o.data(), t.data(), sizeof(float) * 5) == 0);
    CHECK(qpso_host::rand_stride(20) == 512 && qpso_host::rand_stride(200) == 2412);
    return true;
}

bool io_formats(const std::string& dir) {
    const std::string dpt = dir + "/a.dpt", obj = dir + "/a.obj", csv = dir + "/a.csv";
    std::vector<uint16_t> d((size_t)QPSO_ROWS * QPSO_COLS);
    for (size_t i = 0; i < d.size(); ++i) d[i] = (uint16_t)(i * 7);
    CHECK(qpso_write_dpt(dpt.c_str(), d.data(), QPSO_ROWS, QPSO_COLS) == QPSO_OK);
    uint16_t* back = nullptr;
    int32_t r = 0, c = 0;
    CHECK(qpso_read_dpt(dpt.c_str(), &back, &r, &c) == QPSO_OK && r == QPSO_ROWS && c == QPSO_COLS);
    CHECK(std::memcmp(back, d.data(), d.size() * 2) == 0);
    qpso_free(back);
    // truncated data and impossible headers are refused
    {
        FILE* f = std::fopen(dpt.c_str(), "r+b");
        const int32_t hdr[2] = {1000, 1000};
        std::fwrite(hdr, 4, 2, f);
        std::fclose(f);
        CHECK(qpso_read_dpt(dpt.c_str(), &back, &r, &c) == QPSO_E_INVAL);
        f = std::fopen(dpt.c_str(), "wb");
        const int32_t bad[2] = {-5, 3};
        std::fwrite(bad, 4, 2, f);
        std::fclose(f);
        CHECK(qpso_read_dpt(dpt.c_str(), &back, &r, &c) == QPSO_E_INVAL);
    }
    auto obj_rc = [&](const char* text, uint32_t* nf) {
        FILE* f = std::fopen(obj.c_str(), "w");
        std::fputs(text, f);
        std::fclose(f);
        float* soa = nullptr;
        const int rc = qpso_read_obj(obj.c_str(), &soa, nf);
        if (rc == QPSO_OK) qpso_free(soa);
        return rc;
    };
    uint32_t nf = 0;
    CHECK(obj_rc("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nf 1 2 4 3\nf -4 -3 -2\n", &nf) == QPSO_OK && nf == 3);
    CHECK(obj_rc("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n", &nf) == QPSO_E_INVAL);    // index 0
    CHECK(obj_rc("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 9\n", &nf) == QPSO_E_INVAL);    // beyond the vertices
    CHECK(obj_rc("v 0 0 0\nv 1 0 0\nf -7 1 2\n", &nf) == QPSO_E_INVAL);            // before the first vertex
    CHECK(obj_rc("v 0 0 0\n", &nf) == QPSO_E_INVAL);                               // no face
    // CPU visualisation render: ordinary, off-frame, behind the camera and
    // near the camera plane (huge bounding boxes, clipped to the frame)
    std::vector<float> mesh = icosphere(2);
    for (auto& v : mesh) v *= 1e-3f;
    const qpso_camera cam = {525.0, 525.0, 319.5, 239.5};
    const float poses[][7] = {{0, 0, 0.7f, 1, 0, 0, 0},     {-0.4f, 0, 0.6f, 0.9f, 0.1f, 0.3f, 0.3f},
                              {0, 0, -0.5f, 1, 0, 0, 0},    {0, 0, 0.01f, 0.7f, 0.7f, 0, 0},
                              {5, 5, 0.001f, 1, 0, 0, 0},   {0, 0, 1e-7f, 1, 0, 0, 0}};
    std::vector<float> buf((size_t)QPSO_ROWS * QPSO_COLS);
    for (const auto& p : poses) {
        std::fill(buf.begin(), buf.end(), 60000.0f);
        CHECK(qpso_render_cpu(p, mesh.data(), (uint32_t)(mesh.size() / 9), &cam, buf.data()) == QPSO_OK);
    }
    CHECK(qpso_render_cpu_csv(csv.c_str(), poses[0], mesh.data(), (uint32_t)(mesh.size() / 9), &cam) == QPSO_OK);
    CHECK(qpso_write_bestpose((dir + "/b.csv").c_str(), 0.5f, poses[1]) == QPSO_OK);
    // cluster provider at the frame's borders and an absent label
    std::vector<uint16_t> labels((size_t)QPSO_ROWS * QPSO_COLS, 0), cd(labels.size());
    for (int y = 470; y < 480; ++y)
        for (int x = 0; x < 8; ++x) labels[(size_t)y * QPSO_COLS + x] = 3;
    int32_t rect[4];
    float cen[3];
    uint32_t np = 0;
    CHECK(qpso_clusters_from_labels(labels.data(), d.data(), 2, &cam, rect, cen, &np, cd.data()) == QPSO_OK);
    CHECK(rect[0] == 0 && rect[1] == 470 && rect[2] == 8 && rect[3] == 480 && np == 80);
    CHECK(qpso_clusters_from_labels(labels.data(), d.data(), 5, &cam, rect, cen, &np, cd.data()) == QPSO_E_NOTFOUND);
    return true;
}

bool scenes() {
    std::vector<float> m = uv_mesh(40, 30, 2, 1);
    const qpso_scene_object objs[2] = {{m.data(), (int)(m.size() / 9), {0, 0, 700, 1, 0, 0, 0}},
                                       {m.data(), (int)(m.size() / 9), {-420, 0, 650, 0.7f, 0.7f, 0, 0}}};
    std::vector<uint16_t> depth((size_t)QPSO_ROWS * QPSO_COLS), cd(depth.size());
    std::vector<int32_t> label(depth.size());
    CHECK(qpso_scene_render(objs, 2, 525, 525, 319.5, 239.5, 1100.f, 2.f, 0.05f, 7, depth.data(), label.data()) == 0);
    qpso_cluster_roi roi;
    CHECK(qpso_scene_cluster(depth.data(), label.data(), 0, 525, 525, 319.5, 239.5, &roi, cd.data()) == 0);
    CHECK(qpso_scene_cluster(depth.data(), label.data(), 1, 525, 525, 319.5, 239.5, &roi, cd.data()) == 0);
    CHECK(roi.rect[0] == 0);  // cut by the frame's left edge
    return true;
}

bool oracle_paths() {
    std::vector<float> mesh = icosphere(2);
    for (auto& v : mesh) v *= 1e-3f;
    const int F = (int)(mesh.size() / 9);
    const or_cam cam = {525.0, 525.0, 319.5, 239.5};
    const int rect[4] = {280, 200, 360, 280};
    std::vector<uint16_t> obs((size_t)OR_FRAME, 0);
    for (int y = rect[1]; y < rect[3]; ++y)
        for (int x = rect[0]; x < rect[2]; ++x) obs[(size_t)y * OR_COLS + x] = (uint16_t)(690 + (x + y) % 20);
    std::vector<int32_t> frame((size_t)OR_FRAME);
    const float poses[][7] = {{0, 0, 0.7f, 1, 0, 0, 0},      {-0.4f, 0, 0.6f, 0.9f, 0.1f, 0.3f, 0.3f},
                              {0.35f, 0, 0.6f, 1, 0, 0, 0},  {0, 0, -0.5f, 1, 0, 0, 0},
                              {0, 0, 0.01f, 0.7f, 0.7f, 0, 0}, {5, 5, 0.001f, 1, 0, 0, 0}};
    for (int wrap = 0; wrap < 2; ++wrap)
        for (const auto& p : poses) {
            std::fill(frame.begin(), frame.end(), OR_EMPTY);
            or_render_info ri;
            or_fitness_info fi;
            or_render_ex(mesh.data(), F, 0, p, &cam, rect, wrap, frame.data(), &ri);
            const float fit = or_fitness_ex(&ri, rect, 6400, wrap, frame.data(), obs.data(), &fi);
            CHECK(ri.degenerate ? std::isnan(fit) : fi.visit_px > 0);
        }
    for (int wrap = 0; wrap < 2; ++wrap) {
        or_params pr{};
        pr.n_particles = 8;
        pr.n_iters = 3;
        pr.w = 0.7;
        pr.c1 = 1.5f;
        pr.c2 = 1.5f;
        const float lo[3] = {-0.3f, -0.3f, 0.4f}, hi[3] = {0.3f, 0.3f, 1.0f};
        for (int i = 0; i < 3; ++i) { pr.x_lo[i] = lo[i]; pr.x_hi[i] = hi[i]; }
        std::memcpy(pr.rect, rect, sizeof(rect));
        pr.num_points = 6400;
        pr.cam = cam;
        pr.wrap = wrap;
        std::vector<float> rand(or_rand_stride(3) * 8);
        or_rand_table(0, rand.data(), rand.size());
        std::vector<float> fit_trace(3 * 8), gbest(3 * 8), pos(56);
        std::vector<int> argmin(3);
        or_solve_out out{};
        out.fit_trace = fit_trace.data();
        out.gbest_trace = gbest.data();
        out.argmin_trace = argmin.data();
        out.pos = pos.data();
        CHECK(or_solve(mesh.data(), F, obs.data(), &pr, rand.data(), 2, &out) == 0);
        CHECK(out.best_fit < 1e10f);
    }
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    struct Case {
        const char* name;
        bool (*fn)();
    };
    const Case cases[] = {{"mesh_tables", mesh_tables}, {"crop_and_rows", crop_and_rows},
                          {"random_table", random_table}, {"scenes", scenes}, {"oracle_paths", oracle_paths}};
    for (const Case& c : cases) std::printf("%s %s\n", c.fn() ? "PASS" : "FAIL", c.name);
    std::printf("%s io_formats\n", io_formats(dir) ? "PASS" : "FAIL");
    if (g_fail) return 1;
    std::printf("ALL PASS\n");
    return 0;
}
