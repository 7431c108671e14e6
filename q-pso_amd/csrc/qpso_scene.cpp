// qpso_scene.cpp -- synthetic RGB-D scenes and cluster ROIs (host only).
//
// Stands in for the RGB-D capture and the GraphCannySeg segmentation that
// feed the reference's estimate-pose block (main.cu:1269-1322).  The observed
// depth comes from an independent ray caster (pixel-centre rays, Moller-
// Trumbore intersection in binary64), not from the fixed-point rasterizer
// under test.
#include "qpso_scene.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

#include "qpso_roi.h"

namespace {

constexpr int kRows = 480;
constexpr int kCols = 640;

// splitmix64: seeds and uniform/normal draws (deterministic on every host).
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double normal() {  // Box-Muller
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct V3 { double x, y, z; };
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 mul(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 normalize(V3 a) { double n = std::sqrt(dot(a, a)); return mul(a, 1.0 / n); }

using Tri = std::array<V3, 3>;

// Orient a triangle so that its normal (v1-v0)x(v2-v0) points away from `centre`:
// the rasterizer treats that winding as front-facing (main.h:1363-1367).
Tri outward(Tri t, V3 centre) {
    V3 n = cross(sub(t[1], t[0]), sub(t[2], t[0]));
    V3 c = mul(add(add(t[0], t[1]), t[2]), 1.0 / 3.0);
    if (dot(n, sub(c, centre)) < 0) std::swap(t[1], t[2]);
    return t;
}

void write_soa(const std::vector<Tri>& tris, float* out) {
    const size_t F = tris.size();
    for (size_t f = 0; f < F; ++f)
        for (int j = 0; j < 3; ++j) {
            out[f + j * F] = (float)tris[f][j].x;
            out[3 * F + f + j * F] = (float)tris[f][j].y;
            out[6 * F + f + j * F] = (float)tris[f][j].z;
        }
}

std::vector<Tri> icosphere(int subdiv, double ax, double ay, double az) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<V3> v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0},
                         {0, -1, t}, {0, 1, t}, {0, -1, -t}, {0, 1, -t},
                         {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& p : v) p = normalize(p);
    std::vector<std::array<int, 3>> f = {
        {0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
        {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
        {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    for (int s = 0; s < subdiv; ++s) {
        std::map<std::pair<int, int>, int> mid;
        auto midpoint = [&](int a, int b) {
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            v.push_back(normalize(mul(add(v[a], v[b]), 0.5)));
            return mid[key] = (int)v.size() - 1;
        };
        std::vector<std::array<int, 3>> nf;
        nf.reserve(f.size() * 4);
        for (auto& tri : f) {
            int a = midpoint(tri[0], tri[1]), b = midpoint(tri[1], tri[2]), c = midpoint(tri[2], tri[0]);
            nf.push_back({tri[0], a, c});
            nf.push_back({tri[1], b, a});
            nf.push_back({tri[2], c, b});
            nf.push_back({a, b, c});
        }
        f.swap(nf);
    }
    std::vector<Tri> out;
    out.reserve(f.size());
    for (auto& tri : f) {
        Tri T;
        for (int j = 0; j < 3; ++j) T[j] = {v[tri[j]].x * ax, v[tri[j]].y * ay, v[tri[j]].z * az};
        out.push_back(outward(T, {0, 0, 0}));
    }
    return out;
}

void box(std::vector<Tri>& out, V3 c, V3 h) {
    V3 p[8];
    for (int i = 0; i < 8; ++i)
        p[i] = {c.x + ((i & 1) ? h.x : -h.x), c.y + ((i & 2) ? h.y : -h.y), c.z + ((i & 4) ? h.z : -h.z)};
    const int q[6][4] = {{0, 1, 3, 2}, {4, 5, 7, 6}, {0, 1, 5, 4}, {2, 3, 7, 6}, {0, 2, 6, 4}, {1, 3, 7, 5}};
    for (auto& f : q) {
        out.push_back(outward({p[f[0]], p[f[1]], p[f[2]]}, c));
        out.push_back(outward({p[f[0]], p[f[2]], p[f[3]]}, c));
    }
}

std::vector<Tri> uv_ellipsoid(int slices, int stacks, double ax, double ay, double az,
                              int n_boxes, double bmin, double bmax, uint64_t seed) {
    auto P = [&](int i, int j) {
        const double th = M_PI * i / stacks, ph = 2.0 * M_PI * (j % slices) / slices;
        return V3{ax * std::sin(th) * std::cos(ph), ay * std::sin(th) * std::sin(ph), az * std::cos(th)};
    };
    std::vector<Tri> out;
    const V3 o{0, 0, 0};
    for (int j = 0; j < slices; ++j) out.push_back(outward({P(0, 0), P(1, j), P(1, j + 1)}, o));
    for (int i = 1; i < stacks - 1; ++i)
        for (int j = 0; j < slices; ++j) {
            out.push_back(outward({P(i, j), P(i + 1, j), P(i + 1, j + 1)}, o));
            out.push_back(outward({P(i, j), P(i + 1, j + 1), P(i, j + 1)}, o));
        }
    for (int j = 0; j < slices; ++j) out.push_back(outward({P(stacks - 1, j), P(stacks, 0), P(stacks - 1, j + 1)}, o));
    Rng r(seed);
    for (int b = 0; b < n_boxes; ++b) {
        const double th = std::acos(2.0 * r.uniform() - 1.0), ph = 2.0 * M_PI * r.uniform();
        V3 c{ax * std::sin(th) * std::cos(ph), ay * std::sin(th) * std::sin(ph), az * std::cos(th)};
        V3 h{0.5 * (bmin + (bmax - bmin) * r.uniform()), 0.5 * (bmin + (bmax - bmin) * r.uniform()),
             0.5 * (bmin + (bmax - bmin) * r.uniform())};
        box(out, c, h);
    }
    return out;
}

V3 rotate(const float q[4], V3 v) {  // R(q) v, q = (w, x, y, z) unit
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    V3 u{x, y, z};
    V3 t = mul(cross(u, v), 2.0);
    return add(add(v, mul(t, w)), cross(u, t));
}

}  // namespace

extern "C" {

int qpso_mesh_icosphere(int subdiv, float ax, float ay, float az, float* out_soa) {
    if (subdiv < 0 || subdiv > 7) return -1;
    int F = 20;
    for (int s = 0; s < subdiv; ++s) F *= 4;
    if (!out_soa) return F;
    auto tris = icosphere(subdiv, ax, ay, az);
    write_soa(tris, out_soa);
    return (int)tris.size();
}

int qpso_mesh_uv_ellipsoid(int slices, int stacks, float ax, float ay, float az, int n_boxes,
                           float box_min, float box_max, uint64_t seed, float* out_soa) {
    if (slices < 3 || stacks < 2 || n_boxes < 0) return -1;
    const int F = 2 * slices * (stacks - 1) + 12 * n_boxes;
    if (!out_soa) return F;
    auto tris = uv_ellipsoid(slices, stacks, ax, ay, az, n_boxes, box_min, box_max, seed);
    write_soa(tris, out_soa);
    return (int)tris.size();
}

void qpso_random_quat(uint64_t seed, float q[4]) {
    Rng r(seed ^ 0x5157A7ull);
    double a[4], n = 0;
    for (int i = 0; i < 4; ++i) { a[i] = r.normal(); n += a[i] * a[i]; }
    n = std::sqrt(n);
    if (a[0] < 0) n = -n;  // scalar part >= 0, like the solver's top-half cast
    for (int i = 0; i < 4; ++i) q[i] = (float)(a[i] / n);
}

int qpso_scene_render(const qpso_scene_object* objs, int n_objs, double fx, double fy, double cx,
                      double cy, float plane_z, float sigma, float hole_frac, uint64_t seed,
                      uint16_t* depth_out, int32_t* label_out) {
    if (!depth_out || n_objs < 0) return -1;
    std::vector<double> zbuf(kRows * kCols, plane_z > 0 ? (double)plane_z : INFINITY);
    std::vector<int32_t> lab(kRows * kCols, -1);
    for (int o = 0; o < n_objs; ++o) {
        const qpso_scene_object& ob = objs[o];
        const int F = ob.n_faces;
        const float* A = ob.mesh_soa;
        const float q[4] = {ob.pose[3], ob.pose[4], ob.pose[5], ob.pose[6]};
        const V3 tr{ob.pose[0], ob.pose[1], ob.pose[2]};
        for (int f = 0; f < F; ++f) {
            V3 v[3];
            bool ok = true;
            double umin = 1e30, umax = -1e30, vmin = 1e30, vmax = -1e30;
            for (int j = 0; j < 3; ++j) {
                V3 m{A[f + j * F], A[3 * F + f + j * F], A[6 * F + f + j * F]};
                v[j] = add(rotate(q, m), tr);
                if (v[j].z <= 1e-6) { ok = false; break; }
                const double u = cx + fx * v[j].x / v[j].z, w = cy + fy * v[j].y / v[j].z;
                umin = std::min(umin, u); umax = std::max(umax, u);
                vmin = std::min(vmin, w); vmax = std::max(vmax, w);
            }
            if (!ok) continue;
            const int x0 = std::max(0, (int)std::floor(umin)), x1 = std::min(kCols - 1, (int)std::ceil(umax));
            const int y0 = std::max(0, (int)std::floor(vmin)), y1 = std::min(kRows - 1, (int)std::ceil(vmax));
            const V3 e1 = sub(v[1], v[0]), e2 = sub(v[2], v[0]);
            for (int y = y0; y <= y1; ++y)
                for (int x = x0; x <= x1; ++x) {
                    const V3 d{(x - cx) / fx, (y - cy) / fy, 1.0};  // pixel-centre ray
                    const V3 pv = cross(d, e2);
                    const double det = dot(e1, pv);
                    if (std::fabs(det) < 1e-14) continue;
                    const double inv = 1.0 / det;
                    const V3 s = mul(v[0], -1.0);  // ray origin is the camera centre
                    const double b1 = dot(s, pv) * inv;
                    if (b1 < 0.0 || b1 > 1.0) continue;
                    const V3 qv = cross(s, e1);
                    const double b2 = dot(d, qv) * inv;
                    if (b2 < 0.0 || b1 + b2 > 1.0) continue;
                    const double t = dot(e2, qv) * inv;  // z of the hit, since d.z = 1
                    const int idx = y * kCols + x;
                    if (t > 0 && t < zbuf[idx]) { zbuf[idx] = t; lab[idx] = o; }
                }
        }
    }
    Rng r(seed);
    for (int i = 0; i < kRows * kCols; ++i) {
        double z = zbuf[i];
        uint16_t d = 0;
        if (std::isfinite(z)) {
            z += sigma * r.normal();
            d = (uint16_t)std::min(65535.0, std::max(0.0, std::nearbyint(z)));
        }
        if (r.uniform() < hole_frac) d = 0;
        depth_out[i] = d;
    }
    if (label_out) std::memcpy(label_out, lab.data(), sizeof(int32_t) * kRows * kCols);
    return 0;
}

int qpso_scene_cluster(const uint16_t* depth, const int32_t* label, int obj, double fx, double fy,
                       double cx, double cy, qpso_cluster_roi* roi, uint16_t* cluster_depth_out) {
    int minx = kCols, miny = kRows, maxx = -1, maxy = -1;
    uint32_t n = 0;
    qpso_roi::Centroid cen(qpso_roi::kinv(fx, fy, cx, cy));
    for (int y = 0; y < kRows; ++y)
        for (int x = 0; x < kCols; ++x) {
            const int i = y * kCols + x;
            const bool in = label[i] == obj;
            if (cluster_depth_out) cluster_depth_out[i] = in ? depth[i] : 0;
            if (!in) continue;
            ++n;
            minx = std::min(minx, x); maxx = std::max(maxx, x);
            miny = std::min(miny, y); maxy = std::max(maxy, y);
            cen.add(x, y, depth[i]);
        }
    if (n == 0 || !cen.result(roi->centroid_mm)) return -1;
    roi->rect[0] = minx; roi->rect[1] = miny; roi->rect[2] = maxx + 1; roi->rect[3] = maxy + 1;
    roi->num_points = n;
    return 0;
}

}  // extern "C"
