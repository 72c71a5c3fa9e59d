// mesh_bvh.h — the FP64 triangle mesh of SDFBoundary on the host: triangles
// with per-feature pseudo-normals, angle-weighted vertex normals, and the
// median-split bounding-volume hierarchy with its branch-and-bound closest
// point query (src/fluid/SDFBoundary.cpp:22-265, :318-383).  Shared by the
// host SDF build (fsi_sdf_build) and the device mesh (fsi_mesh_*), which
// uploads this structure: both see the same tree.
//
// The arithmetic follows the reference's Eigen expressions operation for
// operation (divisions where it divides, left-to-right sums), so the tree,
// the normals and every distance are those of SDFBoundary.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <utility>
#include <vector>

namespace fsi {
namespace mesh {

struct D3 {
    double x, y, z;
};
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline D3 operator*(double s, D3 a) { return a * s; }
inline D3 operator/(D3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(D3 a) { return std::sqrt(dot(a, a)); }
inline D3 vmin(D3 a, D3 b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
inline D3 vmax(D3 a, D3 b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }
inline double comp(D3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
// Eigen's normalized(): v / |v| when |v|^2 > 0, else v.
inline D3 normalized(D3 v) {
    const double z = dot(v, v);
    return z > 0.0 ? v / std::sqrt(z) : v;
}

// The closest feature of a triangle (SDFBoundary.cpp:14-20).
enum Feature : int32_t { F_FACE = 0, F_EDGE_AB, F_EDGE_AC, F_EDGE_BC, F_VERT_A, F_VERT_B, F_VERT_C };

struct Tri {
    D3 a, b, c;
    int ia, ib, ic;
    D3 fn, en_ab, en_ac, en_bc;
};

// Closest point on triangle abc to p and the feature it lies on
// (closestPointOnTriangle, SDFBoundary.cpp:22-54; Ericson, Real-Time
// Collision Detection 5.1.5).
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

// TriangleBVH (SDFBoundary.cpp:78-265).  Nodes are stored parent before
// children (children have higher indices), leaves hold at most 4 triangles.
struct Bvh {
    struct Node {
        D3 lo, hi;
        int start = 0, count = 0, left = -1, right = -1;
    };
    std::vector<Node> nodes;
    std::vector<int> order;  // triangle ids, permuted by the build
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
            cen[i] = (t[i].a + t[i].b + t[i].c) / 3.0;
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
            if (v < a) {
                const double t = a - v;
                d2 += t * t;
            } else if (v > b) {
                const double t = v - b;
                d2 += t * t;
            }
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

    // Depth of every node (root 0), for a level-by-level refit.
    std::vector<int> depths() const {
        std::vector<int> d(nodes.size(), 0);
        for (size_t i = 0; i < nodes.size(); ++i)
            if (nodes[i].left >= 0) {
                d[nodes[i].left] = d[i] + 1;
                d[nodes[i].right] = d[i] + 1;
            }
        return d;
    }
};

// The triangle list and normals of SDFBoundary::buildFromMesh
// (SDFBoundary.cpp:318-383) and its BVH (:388).
struct Mesh {
    std::vector<Tri> tris;
    std::vector<D3> vn;  // angle-weighted vertex normals
    Bvh bvh;

    void build(const double* verts, uint32_t nverts, const int32_t* faces, uint32_t nfaces) {
        tris.assign(nfaces, Tri{});
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
            if (l > 1e-12) nrm = nrm / l;
            t.fn = nrm;
            t.en_ab = t.en_ac = t.en_bc = nrm;
        }
        vn.assign(nverts, D3{0, 0, 0});
        auto angle = [](D3 at, D3 p1, D3 p2) {
            return std::acos(std::clamp(dot(normalized(p1 - at), normalized(p2 - at)), -1.0, 1.0));
        };
        for (const Tri& t : tris) {
            vn[t.ia] = vn[t.ia] + angle(t.a, t.b, t.c) * t.fn;
            vn[t.ib] = vn[t.ib] + angle(t.b, t.a, t.c) * t.fn;
            vn[t.ic] = vn[t.ic] + angle(t.c, t.a, t.b) * t.fn;
        }
        for (D3& n : vn) {
            const double l = norm(n);
            if (l > 1e-12) n = n / l;
        }
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
            const D3 en = s / l;
            for (int t : kv.second) {
                Tri& tr = tris[t];
                const int a = kv.first.first, b = kv.first.second;
                if (std::min(tr.ia, tr.ib) == a && std::max(tr.ia, tr.ib) == b) tr.en_ab = en;
                else if (std::min(tr.ia, tr.ic) == a && std::max(tr.ia, tr.ic) == b) tr.en_ac = en;
                else tr.en_bc = en;
            }
        }
        bvh.build(tris);
    }

    // SDFBoundary::signedDistanceAt (SDFBoundary.cpp:400-419).
    float signed_distance(D3 p) const {
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
    }
};

}  // namespace mesh
}  // namespace fsi
