// qpso_io.cpp -- mesh/depth readers, result writers, CPU visualisation render.
#include "qpso_io.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <climits>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "qpso_roi.h"

namespace {
constexpr int kRows = QPSO_ROWS;
constexpr int kCols = QPSO_COLS;
}

extern "C" {

void qpso_free(void* p) { std::free(p); }

int qpso_read_obj(const char* path, float** soa_out, uint32_t* n_faces) {
    if (!path || !soa_out || !n_faces) return QPSO_E_INVAL;
    std::ifstream in(path);
    if (!in) return QPSO_E_NOTFOUND;
    std::vector<float> v;                // all vertices (global indexing)
    std::vector<int> tri;                // 3 vertex indices per face
    std::string line;
    int objects = 0;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string tag;
        if (!(ss >> tag)) continue;
        if (tag == "o") {
            if (++objects > 1 && !tri.empty()) break;  // first object only (loadOBJ3: mMeshes[0])
        } else if (tag == "v") {
            float x = 0, y = 0, z = 0;
            ss >> x >> y >> z;
            v.push_back(x); v.push_back(y); v.push_back(z);
        } else if (tag == "f") {
            std::vector<int> poly;
            std::string tok;
            while (ss >> tok) {
                const int idx = std::atoi(tok.c_str());  // "a", "a/b", "a//c", "a/b/c"
                if (idx == 0) return QPSO_E_INVAL;
                poly.push_back(idx > 0 ? idx - 1 : (int)(v.size() / 3) + idx);
            }
            for (size_t k = 2; k < poly.size(); ++k) {  // fan triangulation
                tri.push_back(poly[0]); tri.push_back(poly[k - 1]); tri.push_back(poly[k]);
            }
        }
    }
    const size_t F = tri.size() / 3;
    if (F == 0) return QPSO_E_INVAL;
    float* A = (float*)std::malloc(sizeof(float) * 9 * F);
    if (!A) return QPSO_E_NOMEM;
    for (size_t f = 0; f < F; ++f)
        for (int j = 0; j < 3; ++j) {
            const int i = tri[f * 3 + j];
            if (i < 0 || (size_t)i * 3 + 2 >= v.size()) { std::free(A); return QPSO_E_INVAL; }
            A[f + j * F] = v[i * 3 + 0];
            A[3 * F + f + j * F] = v[i * 3 + 1];
            A[6 * F + f + j * F] = v[i * 3 + 2];
        }
    *soa_out = A;
    *n_faces = (uint32_t)F;
    return QPSO_OK;
}

int qpso_read_dpt(const char* path, uint16_t** data_out, int32_t* rows, int32_t* cols) {
    if (!path || !data_out || !rows || !cols) return QPSO_E_INVAL;
    FILE* f = std::fopen(path, "rb");
    if (!f) return QPSO_E_NOTFOUND;
    int32_t r = 0, c = 0;
    if (std::fread(&r, 4, 1, f) != 1 || std::fread(&c, 4, 1, f) != 1 || r <= 0 || c <= 0 || r > 16384 ||
        c > 16384) {
        std::fclose(f);
        return QPSO_E_INVAL;
    }
    uint16_t* d = (uint16_t*)std::malloc(sizeof(uint16_t) * (size_t)r * c);
    if (!d) { std::fclose(f); return QPSO_E_NOMEM; }
    const size_t n = std::fread(d, sizeof(uint16_t), (size_t)r * c, f);
    std::fclose(f);
    if (n != (size_t)r * c) { std::free(d); return QPSO_E_INVAL; }
    *data_out = d;
    *rows = r;
    *cols = c;
    return QPSO_OK;
}

int qpso_write_dpt(const char* path, const uint16_t* data, int32_t rows, int32_t cols) {
    if (!path || !data || rows <= 0 || cols <= 0) return QPSO_E_INVAL;
    FILE* f = std::fopen(path, "wb");
    if (!f) return QPSO_E_NOTFOUND;
    bool ok = std::fwrite(&rows, 4, 1, f) == 1 && std::fwrite(&cols, 4, 1, f) == 1 &&
              std::fwrite(data, sizeof(uint16_t), (size_t)rows * cols, f) == (size_t)rows * cols;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? QPSO_OK : QPSO_E_INVAL;
}

int qpso_write_bestpose(const char* path, float fit, const float pose[7]) {
    FILE* out = std::fopen(path, "w");
    if (!out) return QPSO_E_NOTFOUND;
    std::fprintf(out, "%f,%f,%f,%f,%f,%f,%f,%f", fit, pose[0], pose[1], pose[2], pose[3], pose[4], pose[5],
                 pose[6]);
    std::fclose(out);
    return QPSO_OK;
}

// RenderCPU / CPUfixedPointRender (main.h:1810-1991): host visualisation,
// float depth buffer, iround = (int)(v + 0.5f), inside test dead as on the GPU.
int qpso_render_cpu(const float pose[7], const float* A, uint32_t n_faces, const qpso_camera* cam,
                    float* buf) {
    if (!pose || !A || !cam || !buf) return QPSO_E_INVAL;
    const size_t F = n_faces;
    // (int)(v + 0.5f), truncating; out of range (undefined in C++) the
    // reference's x86 host conversion gives INT32_MIN
    auto iround = [](float v) {
        const float t = v + 0.5f;
        return (t > -2147483904.0f && t < 2147483648.0f) ? (int32_t)t : INT32_MIN;
    };
    // 32-bit integer arithmetic wraps, as on the reference's hardware
    auto sub = [](int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); };
    auto mul = [](int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); };
    for (size_t i = 0; i < F; ++i) {
        float U[3], V[3], Z[3];
        for (int j = 0; j < 3; ++j) {
            const float vx = A[i + j * F], vy = A[3 * F + i + j * F], vz = A[6 * F + i + j * F];
            const float q0 = pose[3], q1 = pose[4], q2 = pose[5], q3 = pose[6];
            const float t0 = 2.0f * (q2 * vz - q3 * vy);
            const float t1 = 2.0f * -(q1 * vz - q3 * vx);
            const float t2 = 2.0f * (q1 * vy - q2 * vx);
            const float px = vx + q0 * t0 + (q2 * t2 - q3 * t1) + pose[0];
            const float py = vy + q0 * t1 + -(q1 * t2 - q3 * t0) + pose[1];
            const float pz = vz + q0 * t2 + (q1 * t1 - q2 * t0) + pose[2];
            const float r = 1.f / pz;
            U[j] = (float)(cam->cx + cam->fx * (px * r));
            V[j] = (float)(cam->cy + cam->fy * (py * r));
            Z[j] = 1000.0f * pz;
        }
        const int X1 = iround(16.f * U[0]), X2 = iround(16.f * U[1]), X3 = iround(16.f * U[2]);
        const int Y1 = iround(16.f * V[0]), Y2 = iround(16.f * V[1]), Y3 = iround(16.f * V[2]);
        const int32_t DX12 = sub(X1, X2), DX23 = sub(X2, X3), DX31 = sub(X3, X1);
        const int32_t DY12 = sub(Y1, Y2), DY23 = sub(Y2, Y3), DY31 = sub(Y3, Y1);
        const int32_t tri = sub(mul(DX12, DY31), mul(DX31, DY12));
        if (tri <= 0) continue;
        const float inv = 1.0f / float(tri);
        auto bound = [](int32_t v) { return (int32_t)((uint32_t)v + 0xFu) >> 4; };
        const int minx = bound(std::min(X1, std::min(X2, X3))), maxx = bound(std::max(X1, std::max(X2, X3)));
        const int miny = bound(std::min(Y1, std::min(Y2, Y3))), maxy = bound(std::max(Y1, std::max(Y2, Y3)));
        if ((int64_t)maxx - minx <= 0 || (int64_t)maxy - miny <= 0) continue;
        // edge values (mod 2^32, as the reference's int arithmetic wraps on its
        // hardware) evaluated directly at the first in-frame pixel: equal to
        // the reference's walk from (minx, miny); off-frame pixels are clipped
        const uint32_t C2 = (uint32_t)DY23 * (uint32_t)X2 - (uint32_t)DX23 * (uint32_t)Y2 +
                            (DY23 < 0 || (DY23 == 0 && DX23 > 0));
        const uint32_t C3 = (uint32_t)DY31 * (uint32_t)X3 - (uint32_t)DX31 * (uint32_t)Y3 +
                            (DY31 < 0 || (DY31 == 0 && DX31 > 0));
        const int x0 = std::max(minx, 0), x1 = std::min(maxx, kCols);
        const int y0 = std::max(miny, 0), y1 = std::min(maxy, kRows);
        const float z2 = (Z[1] - Z[0]) * inv, z3 = (Z[2] - Z[0]) * inv;
        for (int y = y0; y < y1; ++y) {
            uint32_t CX2 = C2 + (uint32_t)DX23 * ((uint32_t)y << 4) - (uint32_t)DY23 * ((uint32_t)x0 << 4);
            uint32_t CX3 = C3 + (uint32_t)DX31 * ((uint32_t)y << 4) - (uint32_t)DY31 * ((uint32_t)x0 << 4);
            for (int x = x0; x < x1; ++x) {
                const int cx2 = (int32_t)(CX2 + 0xFu) >> 8, cx3 = (int32_t)(CX3 + 0xFu) >> 8;
                float d = Z[0];
                d += cx2 * z2;
                d += cx3 * z3;
                float& b = buf[y * kCols + x];
                if (d < b) b = d;
                CX2 -= (uint32_t)DY23 << 4;
                CX3 -= (uint32_t)DY31 << 4;
            }
        }
    }
    return QPSO_OK;
}

int qpso_render_cpu_csv(const char* path, const float pose[7], const float* soa, uint32_t n_faces,
                        const qpso_camera* cam) {
    std::vector<float> buf(kRows * kCols, 60000.0f);  // maxVal_depthBuffer
    int rc = qpso_render_cpu(pose, soa, n_faces, cam, buf.data());
    if (rc) return rc;
    std::ofstream ofs(path);
    if (!ofs) return QPSO_E_NOTFOUND;
    for (int r = 0; r < kRows; ++r) {
        for (int c = 0; c < kCols; ++c) {
            ofs << buf[r * kCols + c];
            if (c < kCols - 1) ofs << ", ";
        }
        ofs << "\n";
    }
    return QPSO_OK;
}

int qpso_clusters_from_labels(const uint16_t* labels, const uint16_t* depth, int k, const qpso_camera* cam,
                              int32_t rect[4], float centroid_mm[3], uint32_t* num_points,
                              uint16_t* depth_out) {
    if (!labels || !depth || !cam || !rect || !centroid_mm || !num_points || k < 0) return QPSO_E_INVAL;
    const uint16_t want = (uint16_t)(k + 1);
    int minx = kCols, miny = kRows, maxx = -1, maxy = -1;
    uint32_t n = 0;
    qpso_roi::Centroid cen(qpso_roi::kinv(cam->fx, cam->fy, cam->cx, cam->cy));
    for (int y = 0; y < kRows; ++y)
        for (int x = 0; x < kCols; ++x) {
            const int i = y * kCols + x;
            const bool in = labels[i] == want;
            if (depth_out) depth_out[i] = in ? depth[i] : 0;
            if (!in) continue;
            ++n;  // every cluster pixel, holes included (GraphCannySeg.h:1824)
            minx = std::min(minx, x); maxx = std::max(maxx, x);
            miny = std::min(miny, y); maxy = std::max(maxy, y);
            cen.add(x, y, depth[i]);
        }
    if (n == 0 || !cen.result(centroid_mm)) return QPSO_E_NOTFOUND;
    rect[0] = minx; rect[1] = miny; rect[2] = maxx + 1; rect[3] = maxy + 1;  // cv::boundingRect, br exclusive
    *num_points = n;
    return QPSO_OK;
}

}  // extern "C"
