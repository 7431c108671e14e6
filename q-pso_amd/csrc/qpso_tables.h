// qpso_tables.h -- host-side tables a solve's device data is built from:
// the random table, the mesh's vertex/raster/corner/run tables, and the
// observed-depth crop with its per-row prefix counts.  Host only (no HIP), so
// the CPU tests build it with the sanitizers.
#ifndef QPSO_TABLES_H_
#define QPSO_TABLES_H_

#include <stdint.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "qpso_internal.h"

namespace qpso_host {

// ---- GSL mt19937 (main.cu:1340-1352): gsl_rng_default with seed 0 -> 4357,
//      2002 init_genrand seeding, gsl_rng_uniform = u32 / 2^32 --------------
struct Mt19937 {
    uint32_t mt[624];
    int mti;
    explicit Mt19937(uint32_t s) {
        if (s == 0) s = 4357u;
        mt[0] = s;
        for (int i = 1; i < 624; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
        mti = 624;
    }
    uint32_t next() {
        if (mti >= 624) {
            for (int k = 0; k < 624; ++k) {
                const uint32_t y = (mt[k] & 0x80000000u) | (mt[(k + 1) % 624] & 0x7fffffffu);
                mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
            }
            mti = 0;
        }
        uint32_t y = mt[mti++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        return y;
    }
};

inline int rand_stride(int iters) { return std::max(512, 12 * (iters + 1)); }
// GSL replaces seed 0 by its default 4357 (gsl_rng_set): both name one table
inline uint32_t rand_seed(uint32_t seed) { return seed == 0 ? 4357u : seed; }

// T[p + s N] for s < S (main.cu:1337-1352, 1377-1384): the generator's
// stream in order, each u32 divided by 2^32 in double and rounded to float.
inline void rand_table(uint32_t seed, size_t n, std::vector<float>& out) {
    out.resize(n);
    Mt19937 mt(seed);
    for (auto& v : out) v = (float)((double)mt.next() / 4294967296.0);
}

// Deduplicate the corners of the non-indexed SoA mesh (bitwise-equal
// coordinates become one vertex, numbered by first use in face order, so the
// vertices of faces [0, f) are the prefix [0, v_prefix[f-1])) and build the
// raster table in face order: neighbouring faces in one wave write
// neighbouring pixels, which keeps ds_min_i32 free of bank conflicts.
// Runs of consecutive faces whose corners span at most kRunVerts vertex
// indices (the smallest projected-vertex cache of the kernels), greedy in face
// order.  Vertices are numbered by first use, so neighbouring faces of an
// ordinary mesh share a short index range.
constexpr int32_t kRunVerts = 10112 / 6;
inline void build_runs(const std::vector<int32_t>& tri, uint32_t F, std::vector<int32_t>& runs) {
    runs.clear();
    uint32_t f = 0;
    while (f < F) {
        const uint32_t start = f;
        int32_t lo = INT32_MAX, hi = INT32_MIN;
        for (; f < F; ++f) {
            const int32_t* t = &tri[(size_t)f * 3];
            const int32_t nlo = std::min(lo, std::min(t[0], std::min(t[1], t[2])));
            const int32_t nhi = std::max(hi, std::max(t[0], std::max(t[1], t[2])));
            if (f > start && nhi - nlo + 1 > kRunVerts) break;
            lo = nlo;
            hi = nhi;
        }
        if (hi - lo + 1 > kRunVerts) {  // one face alone spans more: no runs for this mesh
            runs.clear();
            return;
        }
        runs.push_back((int32_t)start);
        runs.push_back(lo);
        runs.push_back(hi + 1);
    }
}

inline void build_mesh_tables(const float* A, uint32_t F, std::vector<float>& verts, std::vector<QpsoFaceRef>& order,
                       std::vector<uint32_t>& v_prefix, std::vector<int32_t>& runs) {
    struct Key {
        uint32_t x, y, z;
        bool operator==(const Key& o) const { return x == o.x && y == o.y && z == o.z; }
    };
    struct KeyHash {
        size_t operator()(const Key& k) const {
            uint64_t h = k.x * 0x9E3779B97F4A7C15ull;
            h ^= (h >> 29) + k.y * 0xBF58476D1CE4E5B9ull;
            h ^= (h >> 31) + k.z * 0x94D049BB133111EBull;
            return (size_t)(h ^ (h >> 32));
        }
    };
    std::unordered_map<Key, int32_t, KeyHash> ids;
    ids.reserve((size_t)F * 2);
    std::vector<int32_t> tri((size_t)F * 3);
    std::vector<float> xs, ys, zs;
    v_prefix.resize(F);
    for (uint32_t f = 0; f < F; ++f) {
        for (int j = 0; j < 3; ++j) {
            const float x = A[f + j * (size_t)F], y = A[3 * (size_t)F + f + j * (size_t)F],
                        z = A[6 * (size_t)F + f + j * (size_t)F];
            Key k;
            std::memcpy(&k.x, &x, 4); std::memcpy(&k.y, &y, 4); std::memcpy(&k.z, &z, 4);
            auto it = ids.find(k);
            int32_t id;
            if (it == ids.end()) {
                id = (int32_t)xs.size();
                ids.emplace(k, id);
                xs.push_back(x); ys.push_back(y); zs.push_back(z);
            } else {
                id = it->second;
            }
            tri[(size_t)f * 3 + j] = id;
        }
        v_prefix[f] = (uint32_t)xs.size();
    }
    const size_t V = xs.size();
    verts.resize(3 * V);
    std::copy(xs.begin(), xs.end(), verts.begin());
    std::copy(ys.begin(), ys.end(), verts.begin() + V);
    std::copy(zs.begin(), zs.end(), verts.begin() + 2 * V);
    order.clear();
    order.reserve(F);
    for (uint32_t f = 0; f < F; ++f)
        order.push_back(QpsoFaceRef{tri[f * 3], tri[f * 3 + 1], tri[f * 3 + 2], (int32_t)f});
    build_runs(tri, F, runs);
}

// Corner j of face f (raster order): its mesh coordinates (x, y, z) and its
// unique-vertex index as the bits of the fourth float.
inline void build_corner_table(const float* soa, uint32_t F, const std::vector<QpsoFaceRef>& order,
                               std::vector<float>& corners) {
    corners.assign((size_t)F * 12, 0.0f);
    for (size_t f = 0; f < F; ++f) {
        const int32_t ids[3] = {order[f].v0, order[f].v1, order[f].v2};
        for (int j = 0; j < 3; ++j) {
            for (int c = 0; c < 3; ++c) corners[f * 12 + j * 4 + c] = soa[(size_t)c * 3 * F + f + (size_t)j * F];
            std::memcpy(&corners[f * 12 + j * 4 + 3], &ids[j], sizeof(int32_t));
        }
    }
}

// The cluster depth's rect (minx, miny, maxx, maxy; max exclusive) as a crop
// of pitch w = maxx - minx, and per crop row the number of observed
// (non-zero) pixels left of each column, (w + 1) counts per row.
inline void build_crop(const uint16_t* depth, int cols, const int32_t rect[4], std::vector<uint16_t>& crop,
                       std::vector<uint16_t>& rows) {
    const int rw = rect[2] - rect[0], rh = rect[3] - rect[1];
    crop.resize((size_t)rw * rh);
    for (int y = 0; y < rh; ++y)
        std::memcpy(&crop[(size_t)y * rw], depth + (size_t)(rect[1] + y) * cols + rect[0], sizeof(uint16_t) * rw);
    rows.resize((size_t)(rw + 1) * rh);
    for (int y = 0; y < rh; ++y) {
        uint16_t* r = &rows[(size_t)y * (rw + 1)];
        r[0] = 0;
        for (int x = 0; x < rw; ++x) r[x + 1] = (uint16_t)(r[x] + (crop[(size_t)y * rw + x] != 0 ? 1 : 0));
    }
}

}  // namespace qpso_host

#endif
