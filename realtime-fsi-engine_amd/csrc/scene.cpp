// scene.cpp — host helpers of the C-ABI: lattice spawn, icosphere mesh and the
// FP64 signed-distance grid of a triangle mesh.
//
// The SDF build follows SDFBoundary::buildFromMesh (src/fluid/SDFBoundary.cpp:278-436):
//   - grid origin = bbox - 5 cells, dims = ceil(span / cell)             (:296-308)
//   - triangle list with face normals, angle-weighted vertex normals and
//     edge normals (normalised sum of the two incident face normals)     (:318-383)
//   - a median-split BVH over triangle AABBs, leaves of at most 4        (:78-265)
//   - per voxel: nearest-first branch-and-bound closest point (Ericson,
//     Real-Time Collision Detection 5.1.5) and the sign of (P - closest)
//     against the pseudo-normal of the closest feature (Baerentzen-Aanaes) (:400-419)
// The grid is written x-fastest, the order the reference uploads to its 3D
// texture (:580-598), so it can be handed to fsi_pbf_set_sdf directly.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fsi_pbf.h"

namespace {

// ---------------- MT19937 (Matsumoto & Nishimura, 1998) ----------------
struct Mt19937 {
    uint32_t mt[624];
    int idx = 624;
    explicit Mt19937(uint32_t seed) {
        mt[0] = seed;
        for (int i = 1; i < 624; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
    }
    uint32_t next() {
        if (idx >= 624) {
            for (int i = 0; i < 624; ++i) {
                const uint32_t y = (mt[i] & 0x80000000u) | (mt[(i + 1) % 624] & 0x7fffffffu);
                mt[i] = mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
            }
            idx = 0;
        }
        uint32_t y = mt[idx++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        return y;
    }
};

struct D3 {
    double x, y, z;
};
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline D3 operator*(double s, D3 a) { return a * s; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(D3 a) { return std::sqrt(dot(a, a)); }
inline D3 vmin(D3 a, D3 b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
inline D3 vmax(D3 a, D3 b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }
inline double comp(D3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
inline D3 normalized_or(D3 v, D3 fallback) {
    const double l = norm(v);
    return l > 1e-12 ? v * (1.0 / l) : fallback;
}

enum Feature { F_FACE, F_EDGE_AB, F_EDGE_AC, F_EDGE_BC, F_VERT_A, F_VERT_B, F_VERT_C };

struct Tri {
    D3 a, b, c;
    int ia, ib, ic;
    D3 fn, en_ab, en_ac, en_bc;
};

// Closest point on triangle abc to p and the feature it lies on (Ericson 5.1.5).
inline D3 closest_on_tri(D3 p, const Tri& t, Feature& f) {
    const D3 ab = t.b - t.a, ac = t.c - t.a, ap = p - t.a;
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) { f = F_VERT_A; return t.a; }
    const D3 bp = p - t.b;
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) { f = F_VERT_B; return t.b; }
    const D3 cp = p - t.c;
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) { f = F_VERT_C; return t.c; }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) { f = F_EDGE_AB; return t.a + (d1 / (d1 - d3)) * ab; }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) { f = F_EDGE_AC; return t.a + (d2 / (d2 - d6)) * ac; }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        f = F_EDGE_BC;
        return t.b + ((d4 - d3) / ((d4 - d3) + (d5 - d6))) * (t.c - t.b);
    }
    const double denom = 1.0 / (va + vb + vc);
    f = F_FACE;
    return t.a + (vb * denom) * ab + (vc * denom) * ac;
}

struct Bvh {
    struct Node {
        D3 lo, hi;
        int start = 0, count = 0, left = -1, right = -1;
    };
    std::vector<Node> nodes;
    std::vector<int> order;
    std::vector<D3> tlo, thi, cen;
    const std::vector<Tri>* tris = nullptr;

    void build(const std::vector<Tri>& t) {
        tris = &t;
        const int n = (int)t.size();
        order.resize(n);
        tlo.resize(n);
        thi.resize(n);
        cen.resize(n);
        for (int i = 0; i < n; ++i) {
            order[i] = i;
            tlo[i] = vmin(vmin(t[i].a, t[i].b), t[i].c);
            thi[i] = vmax(vmax(t[i].a, t[i].b), t[i].c);
            cen[i] = (t[i].a + t[i].b + t[i].c) * (1.0 / 3.0);
        }
        nodes.clear();
        nodes.reserve(2 * (size_t)std::max(1, n));
        if (n > 0) split(0, n);
    }

    int split(int start, int count) {
        const int id = (int)nodes.size();
        nodes.emplace_back();
        D3 lo{1e30, 1e30, 1e30}, hi{-1e30, -1e30, -1e30}, clo = lo, chi = hi;
        for (int i = start; i < start + count; ++i) {
            const int t = order[i];
            lo = vmin(lo, tlo[t]);
            hi = vmax(hi, thi[t]);
            clo = vmin(clo, cen[t]);
            chi = vmax(chi, cen[t]);
        }
        nodes[id].lo = lo;
        nodes[id].hi = hi;
        nodes[id].start = start;
        nodes[id].count = count;
        if (count <= 4) return id;
        const D3 ext = chi - clo;
        int axis = 0;
        if (ext.y > ext.x) axis = 1;
        if (ext.z > comp(ext, axis)) axis = 2;
        const int mid = count / 2;
        std::nth_element(order.begin() + start, order.begin() + start + mid, order.begin() + start + count,
                         [&](int a, int b) { return comp(cen[a], axis) < comp(cen[b], axis); });
        const int l = split(start, mid);
        const int r = split(start + mid, count - mid);
        nodes[id].left = l;
        nodes[id].right = r;
        return id;
    }

    static double box_d2(D3 p, D3 lo, D3 hi) {
        double d2 = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double v = comp(p, k), a = comp(lo, k), b = comp(hi, k);
            if (v < a) d2 += (a - v) * (a - v);
            else if (v > b) d2 += (v - b) * (v - b);
        }
        return d2;
    }

    struct Hit {
        double d2 = 1e30;
        D3 point{0, 0, 0};
        Feature feature = F_FACE;
        int tri = -1;
    };

    void query(int ni, D3 p, Hit& best) const {
        const Node& nd = nodes[ni];
        if (box_d2(p, nd.lo, nd.hi) >= best.d2) return;
        if (nd.left < 0) {
            for (int i = nd.start; i < nd.start + nd.count; ++i) {
                const int t = order[i];
                Feature f;
                const D3 q = closest_on_tri(p, (*tris)[t], f);
                const D3 dq = p - q;
                const double d2 = dot(dq, dq);
                if (d2 < best.d2) {
                    best.d2 = d2;
                    best.point = q;
                    best.feature = f;
                    best.tri = t;
                }
            }
            return;
        }
        const double dl = box_d2(p, nodes[nd.left].lo, nodes[nd.left].hi);
        const double dr = box_d2(p, nodes[nd.right].lo, nodes[nd.right].hi);
        if (dl <= dr) {
            if (dl < best.d2) query(nd.left, p, best);
            if (dr < best.d2) query(nd.right, p, best);
        } else {
            if (dr < best.d2) query(nd.right, p, best);
            if (dl < best.d2) query(nd.left, p, best);
        }
    }
};

thread_local std::string g_scene_error;

}  // namespace

extern "C" {

int fsi_scene_lattice(uint32_t nx, uint32_t ny, uint32_t nz, const float origin[3], float spacing, float jitter,
                      uint32_t seed, const float v0[3], fsi_particle* out) {
    if (!origin || !out) return FSI_ERR_INVALID_ARG;
    Mt19937 rng(seed);
    const float vx = v0 ? v0[0] : 0.0f, vy = v0 ? v0[1] : 0.0f, vz = v0 ? v0[2] : 0.0f;
    size_t id = 0;
    auto draw = [&]() {
        const float u = (float)(rng.next() >> 8) * (1.0f / 16777216.0f);
        return (2.0f * u - 1.0f) * jitter;
    };
    for (uint32_t k = 0; k < nz; ++k)
        for (uint32_t j = 0; j < ny; ++j)
            for (uint32_t i = 0; i < nx; ++i, ++id) {
                fsi_particle& p = out[id];
                std::memset(&p, 0, sizeof(p));
                const float jx = draw(), jy = draw(), jz = draw();
                p.pos[0] = origin[0] + ((float)i + 0.5f) * spacing + jx;
                p.pos[1] = origin[1] + ((float)j + 0.5f) * spacing + jy;
                p.pos[2] = origin[2] + ((float)k + 0.5f) * spacing + jz;
                for (int a = 0; a < 3; ++a) p.pred_pos[a] = p.pos[a];
                p.vel[0] = vx;
                p.vel[1] = vy;
                p.vel[2] = vz;
            }
    return FSI_OK;
}

int fsi_mesh_icosphere(uint32_t subdivisions, const double centre[3], double radius, double* verts,
                       int32_t* faces) {
    if (!centre || !verts || !faces || subdivisions > 9) return FSI_ERR_INVALID_ARG;
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<D3> v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                         {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (D3& p : v) p = p * (1.0 / norm(p));
    std::vector<std::array<int, 3>> f = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11},
                                         {1, 5, 9}, {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                         {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9},
                                         {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    for (uint32_t s = 0; s < subdivisions; ++s) {
        std::map<std::pair<int, int>, int> mid;
        auto midpoint = [&](int a, int b) {
            const std::pair<int, int> key(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            D3 m = (v[a] + v[b]) * 0.5;
            m = m * (1.0 / norm(m));
            v.push_back(m);
            const int idx = (int)v.size() - 1;
            mid.emplace(key, idx);
            return idx;
        };
        std::vector<std::array<int, 3>> nf;
        nf.reserve(f.size() * 4);
        for (const auto& tri : f) {
            const int a = midpoint(tri[0], tri[1]), b = midpoint(tri[1], tri[2]), c = midpoint(tri[2], tri[0]);
            nf.push_back({tri[0], a, c});
            nf.push_back({tri[1], b, a});
            nf.push_back({tri[2], c, b});
            nf.push_back({a, b, c});
        }
        f.swap(nf);
    }
    for (size_t i = 0; i < v.size(); ++i) {
        verts[3 * i + 0] = centre[0] + radius * v[i].x;
        verts[3 * i + 1] = centre[1] + radius * v[i].y;
        verts[3 * i + 2] = centre[2] + radius * v[i].z;
    }
    for (size_t i = 0; i < f.size(); ++i)
        for (int k = 0; k < 3; ++k) faces[3 * i + k] = f[i][k];
    return FSI_OK;
}

int fsi_sdf_build(const double* verts, uint32_t nverts, const int32_t* faces, uint32_t nfaces, double cell,
                  double origin[3], int32_t dims[3], float* grid) {
    if (!verts || !faces || !origin || !dims || nverts == 0 || nfaces == 0 || !(cell > 0.0))
        return FSI_ERR_INVALID_ARG;
    for (uint32_t i = 0; i < 3 * nfaces; ++i)
        if (faces[i] < 0 || (uint32_t)faces[i] >= nverts) return FSI_ERR_INVALID_ARG;
    D3 lo{1e30, 1e30, 1e30}, hi{-1e30, -1e30, -1e30};
    for (uint32_t i = 0; i < nverts; ++i) {
        const D3 p{verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
        lo = vmin(lo, p);
        hi = vmax(hi, p);
    }
    const double buffer = cell * 5.0;
    const D3 o{lo.x - buffer, lo.y - buffer, lo.z - buffer};
    const D3 span{(hi.x + buffer) - o.x, (hi.y + buffer) - o.y, (hi.z + buffer) - o.z};
    int32_t dm[3] = {(int32_t)std::ceil(span.x / cell), (int32_t)std::ceil(span.y / cell),
                     (int32_t)std::ceil(span.z / cell)};
    origin[0] = o.x;
    origin[1] = o.y;
    origin[2] = o.z;
    for (int k = 0; k < 3; ++k) dims[k] = dm[k];
    if (!grid) return FSI_OK;

    // triangles with face normals
    std::vector<Tri> tris(nfaces);
    for (uint32_t i = 0; i < nfaces; ++i) {
        Tri& t = tris[i];
        t.ia = faces[3 * i];
        t.ib = faces[3 * i + 1];
        t.ic = faces[3 * i + 2];
        t.a = {verts[3 * t.ia], verts[3 * t.ia + 1], verts[3 * t.ia + 2]};
        t.b = {verts[3 * t.ib], verts[3 * t.ib + 1], verts[3 * t.ib + 2]};
        t.c = {verts[3 * t.ic], verts[3 * t.ic + 1], verts[3 * t.ic + 2]};
        D3 nrm = cross(t.b - t.a, t.c - t.a);
        const double l = norm(nrm);
        if (l > 1e-12) nrm = nrm * (1.0 / l);
        t.fn = nrm;
        t.en_ab = t.en_ac = t.en_bc = nrm;
    }
    // angle-weighted vertex normals
    std::vector<D3> vn(nverts, D3{0, 0, 0});
    auto angle = [](D3 at, D3 p1, D3 p2) {
        const D3 u = normalized_or(p1 - at, D3{0, 0, 0}), w = normalized_or(p2 - at, D3{0, 0, 0});
        return std::acos(std::clamp(dot(u, w), -1.0, 1.0));
    };
    for (const Tri& t : tris) {
        vn[t.ia] = vn[t.ia] + angle(t.a, t.b, t.c) * t.fn;
        vn[t.ib] = vn[t.ib] + angle(t.b, t.a, t.c) * t.fn;
        vn[t.ic] = vn[t.ic] + angle(t.c, t.a, t.b) * t.fn;
    }
    for (D3& n : vn) {
        const double l = norm(n);
        if (l > 1e-12) n = n * (1.0 / l);
    }
    // edge normals: shared by exactly two faces
    std::map<std::pair<int, int>, std::vector<int>> edges;
    for (int t = 0; t < (int)tris.size(); ++t) {
        auto add_edge = [&](int a, int b) { edges[{std::min(a, b), std::max(a, b)}].push_back(t); };
        add_edge(tris[t].ia, tris[t].ib);
        add_edge(tris[t].ia, tris[t].ic);
        add_edge(tris[t].ib, tris[t].ic);
    }
    for (const auto& kv : edges) {
        if (kv.second.size() != 2) continue;
        const D3 s = tris[kv.second[0]].fn + tris[kv.second[1]].fn;
        const double l = norm(s);
        if (l < 1e-12) continue;
        const D3 en = s * (1.0 / l);
        for (int t : kv.second) {
            Tri& tr = tris[t];
            const int a = kv.first.first, b = kv.first.second;
            if (std::min(tr.ia, tr.ib) == a && std::max(tr.ia, tr.ib) == b) tr.en_ab = en;
            else if (std::min(tr.ia, tr.ic) == a && std::max(tr.ia, tr.ic) == b) tr.en_ac = en;
            else tr.en_bc = en;
        }
    }
    Bvh bvh;
    bvh.build(tris);

    auto signed_distance = [&](D3 p) {
        Bvh::Hit hit;
        bvh.query(0, p, hit);
        const Tri& t = tris[hit.tri];
        D3 pn;
        switch (hit.feature) {
            case F_FACE: pn = t.fn; break;
            case F_EDGE_AB: pn = t.en_ab; break;
            case F_EDGE_AC: pn = t.en_ac; break;
            case F_EDGE_BC: pn = t.en_bc; break;
            case F_VERT_A: pn = vn[t.ia]; break;
            case F_VERT_B: pn = vn[t.ib]; break;
            default: pn = vn[t.ic]; break;
        }
        const double sign = dot(p - hit.point, pn) >= 0.0 ? 1.0 : -1.0;
        return (float)(sign * std::sqrt(hit.d2));
    };

    const int W = dm[0], H = dm[1], D = dm[2];
    unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nthreads; ++w) {
        pool.emplace_back([&, w]() {
            for (int z = (int)w; z < D; z += (int)nthreads)
                for (int y = 0; y < H; ++y)
                    for (int x = 0; x < W; ++x) {
                        const D3 p{o.x + x * cell, o.y + y * cell, o.z + z * cell};
                        grid[(size_t)x + (size_t)W * ((size_t)y + (size_t)H * (size_t)z)] = signed_distance(p);
                    }
        });
    }
    for (std::thread& t : pool) t.join();
    return FSI_OK;
}

}  // extern "C"
