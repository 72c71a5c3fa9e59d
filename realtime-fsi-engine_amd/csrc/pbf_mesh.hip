// pbf_mesh.hip — the triangle mesh of the FSI coupling on the device
// (fsi_mesh_* and fsi_pbf_contact_forces in include/fsi_pbf.h).
//
// The mesh is prepared once per topology on the host (mesh_bvh.h: triangles,
// pseudo-normals, the median-split BVH of SDFBoundary) and uploaded in FP64.
// On the device:
//   - the signed-distance grid (SDFBoundary::recomputeDistances,
//     src/fluid/SDFBoundary.cpp:421-436): one thread per voxel, a stack walk
//     of the BVH in the reference's nearest-child-first order, the closest
//     point of Ericson 5.1.5 and the pseudo-normal sign, all in FP64, written
//     x-fastest, the order the solver samples;
//   - the refit after a deformation (updateGeometryAndRefit / refitFromMesh,
//     :438-467): triangle corners and face normals from the new vertices,
//     node boxes bottom-up one tree level at a time; edge and vertex
//     pseudo-normals keep their build values, as in the reference;
//   - the dirty refit (refitDirty, :469-548): voxels within a band of every
//     vertex that moved (old and new position) are marked and only those are
//     recomputed;
//   - the contact forces (computeContactForces, src/main.cpp:604-710) over the
//     solver's contact candidates: closest triangle, push direction and
//     magnitude, barycentric split onto the three vertices.  The per-vertex
//     sums run in a fixed order: the three (vertex, corner) keys of every
//     candidate are radix-sorted by vertex (stable, so candidates stay in
//     particle-id order) and each vertex adds its terms in that order.  The
//     reference adds per-thread partial sums in a thread-dependent order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/fsi_pbf.h"
#include "mesh_bvh.h"
#include "pbf_core.h"
#include "pbf_mesh.h"

using namespace fsi;

namespace {

constexpr int kMB = 256;
constexpr int kStack = 64;  // BVH walk stack; meshes whose tree is deeper are refused
constexpr int kRefitBlock = 1024;

__device__ __forceinline__ uint32_t mgid() { return blockIdx.x * kMB + threadIdx.x; }
inline dim3 mgrid(size_t n) { return dim3((unsigned)((n + kMB - 1) / kMB)); }

struct DNode {
    double lo[3], hi[3];
    int32_t start, count, left, right;
};
struct DTri {  // BVH order: slot k holds triangle order[k]
    double a[3], b[3], c[3];
    int32_t id, ia, ib, ic;
};
struct DNorm {  // by triangle id: face, edge AB, AC, BC pseudo-normals
    double fn[3], ab[3], ac[3], bc[3];
};

// ---- FP64 geometry, the operations of mesh_bvh.h ---------------------------

struct Q3 {
    double x, y, z;
};
__device__ __forceinline__ Q3 q3(const double* p) { return Q3{p[0], p[1], p[2]}; }
__device__ __forceinline__ void put3(double* p, Q3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
__device__ __forceinline__ Q3 qadd(Q3 a, Q3 b) { return Q3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ Q3 qsub(Q3 a, Q3 b) { return Q3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ Q3 qscl(Q3 a, double s) { return Q3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ Q3 qdiv(Q3 a, double s) { return Q3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ double qdot(Q3 a, Q3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ Q3 qcross(Q3 a, Q3 b) {
    return Q3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ Q3 qmin(Q3 a, Q3 b) { return Q3{fmin(a.x, b.x), fmin(a.y, b.y), fmin(a.z, b.z)}; }
__device__ __forceinline__ Q3 qmax(Q3 a, Q3 b) { return Q3{fmax(a.x, b.x), fmax(a.y, b.y), fmax(a.z, b.z)}; }

__device__ __forceinline__ Q3 closest_on_tri(Q3 p, const DTri& t, int32_t& f) {
    const Q3 A = q3(t.a), B = q3(t.b), C = q3(t.c);
    const Q3 ab = qsub(B, A), ac = qsub(C, A), ap = qsub(p, A);
    const double d1 = qdot(ab, ap), d2 = qdot(ac, ap);
    if (d1 <= 0 && d2 <= 0) { f = mesh::F_VERT_A; return A; }
    const Q3 bp = qsub(p, B);
    const double d3 = qdot(ab, bp), d4 = qdot(ac, bp);
    if (d3 >= 0 && d4 <= d3) { f = mesh::F_VERT_B; return B; }
    const Q3 cp = qsub(p, C);
    const double d5 = qdot(ab, cp), d6 = qdot(ac, cp);
    if (d6 >= 0 && d5 <= d6) { f = mesh::F_VERT_C; return C; }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) { f = mesh::F_EDGE_AB; return qadd(A, qscl(ab, d1 / (d1 - d3))); }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) { f = mesh::F_EDGE_AC; return qadd(A, qscl(ac, d2 / (d2 - d6))); }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        f = mesh::F_EDGE_BC;
        return qadd(B, qscl(qsub(C, B), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
    }
    const double denom = 1.0 / (va + vb + vc);
    f = mesh::F_FACE;
    return qadd(qadd(A, qscl(ab, vb * denom)), qscl(ac, vc * denom));
}

__device__ __forceinline__ double box_d2(Q3 p, const DNode& nd) {
    const double v[3] = {p.x, p.y, p.z};
    double d2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (v[k] < nd.lo[k]) {
            const double t = nd.lo[k] - v[k];
            d2 += t * t;
        } else if (v[k] > nd.hi[k]) {
            const double t = v[k] - nd.hi[k];
            d2 += t * t;
        }
    }
    return d2;
}

struct DHit {
    double d2;
    Q3 p;
    int32_t feature, tri;
};

// TriangleBVH::closestPoint (SDFBoundary.cpp:148-153, 225-253) without
// recursion: the farther child is pushed under the nearer one, and every
// popped node is tested against the best hit as the recursive call tests it
// on entry, so nodes are visited and pruned in the same order.  The stack
// never holds more than depth + 1 entries.
__device__ DHit bvh_closest(const DNode* __restrict__ nodes, const DTri* __restrict__ tris, Q3 p) {
    DHit best{1e30, Q3{0.0, 0.0, 0.0}, mesh::F_FACE, -1};
    int32_t stack[kStack];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const DNode nd = nodes[stack[--sp]];
        if (box_d2(p, nd) >= best.d2) continue;
        if (nd.left < 0) {
            for (int32_t i = nd.start; i < nd.start + nd.count; ++i) {
                const DTri t = tris[i];
                int32_t f;
                const Q3 q = closest_on_tri(p, t, f);
                const Q3 dq = qsub(p, q);
                const double d2 = qdot(dq, dq);
                if (d2 < best.d2) {
                    best.d2 = d2;
                    best.p = q;
                    best.feature = f;
                    best.tri = t.id;
                }
            }
            continue;
        }
        const double dl = box_d2(p, nodes[nd.left]);
        const double dr = box_d2(p, nodes[nd.right]);
        if (dl <= dr) {
            stack[sp++] = nd.right;
            stack[sp++] = nd.left;
        } else {
            stack[sp++] = nd.left;
            stack[sp++] = nd.right;
        }
    }
    return best;
}

struct MeshView {
    const DNode* nodes;
    const DTri* tris;
    const DNorm* norms;
    const double* vn;      // vertex normals, 3 per vertex
    const int32_t* faces;  // 3 per triangle id
    double origin[3], cell;
    int32_t dims[3];
};

// The dirty-refit statistics: the number of recomputed voxels and their
// bounding box, each kept in kSpread slots (block b folds into slot
// b mod kSpread; atomics of many blocks on one address would queue at the
// cache).  Words: [0, kSpread) count, then per axis k the minima at
// (1 + k) * kSpread and the maxima at (4 + k) * kSpread.
constexpr uint32_t kSpread = 64, kDirtyWords = 7 * kSpread;

// SDFBoundary::signedDistanceAt (SDFBoundary.cpp:400-419) of every voxel or,
// with a mask, of the marked voxels (clearing the marks and adding them to
// the dirty statistics).  Voxels x-fastest.
__global__ __launch_bounds__(kMB) void k_sdf(MeshView m, float* __restrict__ grid, uint32_t* __restrict__ mask,
                                             uint32_t total, int32_t* __restrict__ dirty) {
    const uint32_t v = mgid();
    bool work = v < total;
    if (work && mask) {
        work = mask[v] != 0u;
        if (work) mask[v] = 0u;
    }
    const uint32_t W = (uint32_t)m.dims[0], H = (uint32_t)m.dims[1];
    const int32_t x = (int32_t)(v % W), y = (int32_t)((v / W) % H), z = (int32_t)(v / (W * H));
    if (work) {
        const Q3 p{m.origin[0] + x * m.cell, m.origin[1] + y * m.cell, m.origin[2] + z * m.cell};
        const DHit hit = bvh_closest(m.nodes, m.tris, p);
        if (hit.tri < 0) {  // only when every distance is NaN (non-finite vertices)
            grid[v] = __int_as_float(0x7fc00000);
        } else {
            const DNorm& nt = m.norms[hit.tri];
            const int32_t* f = m.faces + 3 * (size_t)hit.tri;
            Q3 pn;
            switch (hit.feature) {
                case mesh::F_FACE: pn = q3(nt.fn); break;
                case mesh::F_EDGE_AB: pn = q3(nt.ab); break;
                case mesh::F_EDGE_AC: pn = q3(nt.ac); break;
                case mesh::F_EDGE_BC: pn = q3(nt.bc); break;
                case mesh::F_VERT_A: pn = q3(m.vn + 3 * (size_t)f[0]); break;
                case mesh::F_VERT_B: pn = q3(m.vn + 3 * (size_t)f[1]); break;
                default: pn = q3(m.vn + 3 * (size_t)f[2]); break;
            }
            const double sign = qdot(qsub(p, hit.p), pn) >= 0.0 ? 1.0 : -1.0;
            grid[v] = (float)(sign * sqrt(hit.d2));
        }
    }
    if (!mask) return;  // uniform over the grid
    // statistics of the block's recomputed voxels
    __shared__ int32_t part[7][kMB / 64];
    int32_t st[7] = {work ? 1 : 0,          work ? x : INT32_MAX, work ? y : INT32_MAX, work ? z : INT32_MAX,
                     work ? x : INT32_MIN, work ? y : INT32_MIN, work ? z : INT32_MIN};
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        for (int o = 32; o > 0; o >>= 1) {
            const int32_t u = __shfl_xor(st[k], o);
            st[k] = k == 0 ? st[k] + u : (k < 4 ? min(st[k], u) : max(st[k], u));
        }
        if ((threadIdx.x & 63u) == 0u) part[k][threadIdx.x / 64u] = st[k];
    }
    __syncthreads();
    if (threadIdx.x < 7u) {
        const int k = (int)threadIdx.x;
        int32_t r = part[k][0];
        for (uint32_t w = 1; w < kMB / 64; ++w)
            r = k == 0 ? r + part[k][w] : (k < 4 ? min(r, part[k][w]) : max(r, part[k][w]));
        int32_t marked = 0;
        for (uint32_t w = 0; w < kMB / 64; ++w) marked += part[0][w];
        if (marked > 0) {
            int32_t* slot = dirty + (size_t)k * kSpread + blockIdx.x % kSpread;
            if (k == 0) atomicAdd(slot, r);
            else if (k < 4) atomicMin(slot, r);
            else atomicMax(slot, r);
        }
    }
}

// updateGeometryAndRefit, first half (SDFBoundary.cpp:438-452): corners from
// the moved vertices, face normal (B - A) x (C - A) / |.| when |.| > 1e-12.
__global__ __launch_bounds__(kMB) void k_update_tris(const double* __restrict__ verts, DTri* __restrict__ tris,
                                                     DNorm* __restrict__ norms, uint32_t nt) {
    const uint32_t k = mgid();
    if (k >= nt) return;
    DTri t = tris[k];
    const Q3 A = q3(verts + 3 * (size_t)t.ia), B = q3(verts + 3 * (size_t)t.ib), C = q3(verts + 3 * (size_t)t.ic);
    put3(t.a, A);
    put3(t.b, B);
    put3(t.c, C);
    tris[k] = t;
    Q3 n = qcross(qsub(B, A), qsub(C, A));
    const double len = sqrt(qdot(n, n));
    if (len > 1e-12) n = qdiv(n, len);
    put3(norms[t.id].fn, n);
}

// TriangleBVH::refit (SDFBoundary.cpp:108-141): node boxes bottom-up.  One
// workgroup walks the tree levels from the deepest (level_nodes holds node
// ids grouped by depth, level_off[d] the first of depth d).
__global__ __launch_bounds__(kRefitBlock) void k_refit(DNode* __restrict__ nodes, const DTri* __restrict__ tris,
                                                       const int32_t* __restrict__ level_nodes,
                                                       const int32_t* __restrict__ level_off, int32_t levels) {
    for (int32_t d = levels - 1; d >= 0; --d) {
        for (int32_t i = level_off[d] + (int32_t)threadIdx.x; i < level_off[d + 1]; i += kRefitBlock) {
            DNode& nd = nodes[level_nodes[i]];
            Q3 lo{1e30, 1e30, 1e30}, hi{-1e30, -1e30, -1e30};
            if (nd.left < 0) {
                for (int32_t k = nd.start; k < nd.start + nd.count; ++k) {
                    const Q3 A = q3(tris[k].a), B = q3(tris[k].b), C = q3(tris[k].c);
                    lo = qmin(lo, qmin(qmin(A, B), C));
                    hi = qmax(hi, qmax(qmax(A, B), C));
                }
            } else {
                const DNode& l = nodes[nd.left];
                const DNode& r = nodes[nd.right];
                lo = qmin(q3(l.lo), q3(r.lo));
                hi = qmax(q3(l.hi), q3(r.hi));
            }
            put3(nd.lo, lo);
            put3(nd.hi, hi);
        }
        __syncthreads();
    }
}

// refitDirty steps 1-2 (SDFBoundary.cpp:480-525): every vertex that moved by
// more than move_eps marks the voxels of a band around its old and its new
// position.  One wave per vertex; its lanes split the band's box.  Marks are
// plain stores of 1 (no value is read back); k_sdf counts and clears them.
__device__ void mark_band(const double* w, double band, const MeshView& m, uint32_t lane,
                          uint32_t* __restrict__ mask) {
    int32_t b0[3], b1[3];
    for (int k = 0; k < 3; ++k) {
        const double lo = ((w[k] - band) - m.origin[k]) / m.cell;
        const double hi = ((w[k] + band) - m.origin[k]) / m.cell;
        const double fl = fmin(fmax(floor(lo), -1.0), (double)m.dims[k]);
        const double cl = fmin(fmax(ceil(hi), -1.0), (double)m.dims[k]);
        b0[k] = min(max((int32_t)fl, 0), m.dims[k] - 1);
        b1[k] = min(max((int32_t)cl, 0), m.dims[k] - 1);
    }
    const uint32_t nx = (uint32_t)(b1[0] - b0[0] + 1), ny = (uint32_t)(b1[1] - b0[1] + 1);
    const uint32_t nz = (uint32_t)(b1[2] - b0[2] + 1);
    const size_t W = (size_t)m.dims[0], H = (size_t)m.dims[1];
    for (uint32_t t = lane; t < nx * ny * nz; t += 64u) {
        const size_t x = (size_t)b0[0] + t % nx, y = (size_t)b0[1] + (t / nx) % ny, z = (size_t)b0[2] + t / (nx * ny);
        __hip_atomic_store(&mask[x + W * (y + H * z)], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ __launch_bounds__(kMB) void k_mark(const double* __restrict__ old_v, const double* __restrict__ new_v,
                                              uint32_t nv, double band, double move_eps2, MeshView m,
                                              uint32_t* __restrict__ mask) {
    const uint32_t i = mgid() / 64u, lane = threadIdx.x & 63u;
    if (i >= nv) return;  // the wave's vertex
    const Q3 nq = q3(new_v + 3 * (size_t)i), oq = q3(old_v + 3 * (size_t)i);
    const Q3 d = qsub(nq, oq);
    if (qdot(d, d) <= move_eps2) return;
    mark_band(old_v + 3 * (size_t)i, band, m, lane, mask);
    mark_band(new_v + 3 * (size_t)i, band, m, lane, mask);
}

__global__ __launch_bounds__(kMB) void k_dirty_init(int32_t* __restrict__ dirty) {
    for (uint32_t w = threadIdx.x; w < kDirtyWords; w += kMB) {
        const uint32_t k = w / kSpread;
        dirty[w] = k == 0 ? 0 : (k < 4 ? INT32_MAX : INT32_MIN);
    }
}

__global__ __launch_bounds__(kMB) void k_zero32(uint32_t* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * kMB + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ---- contact forces ---------------------------------------------------------

// counter += the number of this wave's lanes with hit: one atomic per wave,
// by its first active lane, instead of one per lane on the same address.
// Every active lane of the wave calls it.
__device__ __forceinline__ void count_wave(uint32_t* counter, bool hit) {
    const unsigned long long b = __ballot(hit);
    const int first = __ffsll((long long)__ballot(1)) - 1;
    if (b != 0ull && (int)__lane_id() == first) atomicAdd(counter, (uint32_t)__popcll(b));
}

// barycentric (src/main.cpp:557-571), FP64.
__device__ void barycentric(Q3 pt, Q3 A, Q3 B, Q3 C, float& u, float& v, float& w) {
    const Q3 v0 = qsub(B, A), v1 = qsub(C, A), v2 = qsub(pt, A);
    const double d00 = qdot(v0, v0), d01 = qdot(v0, v1), d11 = qdot(v1, v1);
    const double d20 = qdot(v2, v0), d21 = qdot(v2, v1);
    const double denom = d00 * d11 - d01 * d01;
    if (fabs(denom) < 1e-20) {
        u = 1.0f;
        v = 0.0f;
        w = 0.0f;
        return;
    }
    const double vv = (d11 * d20 - d01 * d21) / denom;
    const double ww = (d00 * d21 - d01 * d20) / denom;
    double uu = 1.0 - vv - ww;
    uu = 0.0 < uu ? uu : 0.0;
    double vc = 0.0 < vv ? vv : 0.0, wc = 0.0 < ww ? ww : 0.0;
    const double s = uu + vc + wc;
    if (s > 1e-20) {
        uu /= s;
        vc /= s;
        wc /= s;
    }
    u = (float)uu;
    v = (float)vc;
    w = (float)wc;
}

// One candidate of computeContactForces (src/main.cpp:630-675): its force
// (dir * mag) and barycentric weights, and the keys (vertex, or nverts when
// it makes no contact) of its three corners.  The candidates are the first
// min(*count, cap) of cand (all cap when count is NULL).
__global__ __launch_bounds__(kMB) void k_contact(MeshView m, const double* __restrict__ verts,
                                                 const float4* __restrict__ cand, const uint32_t* __restrict__ count,
                                                 uint32_t cap, float band, int use_vel, uint32_t nv,
                                                 float4* __restrict__ cf, uint32_t* __restrict__ keys,
                                                 uint32_t* __restrict__ counters) {
    const uint32_t c = mgid();
    if (c >= cap) return;
    uint32_t key[3] = {nv, nv, nv};
    float4 f0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), f1 = f0;
    const uint32_t n = count ? min(*count, cap) : cap;
    bool contact = false;
    if (c < n) {
        const float4 p4 = cand[2 * (size_t)c], v4 = cand[2 * (size_t)c + 1];
        const DHit hit = bvh_closest(m.nodes, m.tris, Q3{(double)p4.x, (double)p4.y, (double)p4.z});
        bool ok = hit.tri >= 0;
        float tx = 0.0f, ty = 0.0f, tz = 0.0f, dist = 0.0f;
        float px = 0.0f, py = 0.0f, pz = 0.0f;
        if (ok) {
            px = (float)hit.p.x;
            py = (float)hit.p.y;
            pz = (float)hit.p.z;
            tx = px - p4.x;
            ty = py - p4.y;
            tz = pz - p4.z;
            dist = sqrtf(tx * tx + ty * ty + tz * tz);
            ok = !(dist > band);
        }
        const int32_t* f = m.faces + 3 * (size_t)(ok ? hit.tri : 0);
        float dx = 0.0f, dy = 0.0f, dz = 0.0f, mag = 1.0f;
        Q3 A{0, 0, 0}, B{0, 0, 0}, C{0, 0, 0};
        if (ok) {
            A = q3(verts + 3 * (size_t)f[0]);
            B = q3(verts + 3 * (size_t)f[1]);
            C = q3(verts + 3 * (size_t)f[2]);
            if (dist > 1e-6f) {
                const float inv = 1.0f / dist;
                dx = tx * inv;
                dy = ty * inv;
                dz = tz * inv;
            } else {
                Q3 nrm = qcross(qsub(B, A), qsub(C, A));
                const double z2 = qdot(nrm, nrm);
                if (z2 > 0.0) nrm = qdiv(nrm, sqrt(z2));
                dx = (float)(-nrm.x);
                dy = (float)(-nrm.y);
                dz = (float)(-nrm.z);
            }
            if (use_vel) {
                const float vd = v4.x * dx + v4.y * dy + v4.z * dz;
                mag = 0.0f < vd ? vd : 0.0f;
                ok = !(mag <= 0.0f);
            }
        }
        if (ok) {
            float u, v, w;
            barycentric(Q3{(double)px, (double)py, (double)pz}, A, B, C, u, v, w);
            f0 = make_float4(dx * mag, dy * mag, dz * mag, 0.0f);
            f1 = make_float4(u, v, w, 0.0f);
            key[0] = (uint32_t)f[0];
            key[1] = (uint32_t)f[1];
            key[2] = (uint32_t)f[2];
            contact = true;
        }
    }
    count_wave(&counters[0], contact);
    cf[2 * (size_t)c] = f0;
    cf[2 * (size_t)c + 1] = f1;
    keys[3 * (size_t)c] = key[0];
    keys[3 * (size_t)c + 1] = key[1];
    keys[3 * (size_t)c + 2] = key[2];
}

__global__ __launch_bounds__(kMB) void k_seg_clear(int2* __restrict__ seg, uint32_t n) {
    const uint32_t i = mgid();
    if (i < n) seg[i] = make_int2(0, 0);
}

__global__ __launch_bounds__(kMB) void k_seg_build(const uint32_t* __restrict__ skeys, uint32_t n,
                                                   int2* __restrict__ seg) {
    const uint32_t i = mgid();
    if (i >= n) return;
    const uint32_t k = skeys[i];
    if (i == 0 || skeys[i - 1] != k) seg[k].x = (int32_t)i;
    if (i == n - 1 || skeys[i + 1] != k) seg[k].y = (int32_t)(i + 1);
}

// Per vertex: F += force * weight over its (candidate, corner) terms in
// candidate order; then the control-point test of src/main.cpp:689-694.
__global__ __launch_bounds__(kMB) void k_vertex_sum(const int2* __restrict__ seg, const uint32_t* __restrict__ sids,
                                                    const float4* __restrict__ cf, uint32_t nv, float threshold,
                                                    float4* __restrict__ forces, uint32_t* __restrict__ counters) {
    const uint32_t v = mgid();
    if (v >= nv) return;
    const int2 se = seg[v];
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    for (int32_t s = se.x; s < se.y; ++s) {
        const uint32_t i = sids[s];
        const uint32_t c = i / 3u, k = i - 3u * c;
        const float4 f = cf[2 * (size_t)c];
        const float4 w4 = cf[2 * (size_t)c + 1];
        const float w = k == 0u ? w4.x : (k == 1u ? w4.y : w4.z);
        fx = fx + f.x * w;
        fy = fy + f.y * w;
        fz = fz + f.z * w;
    }
    forces[v] = make_float4(fx, fy, fz, 0.0f);
    const float mg = sqrtf(fx * fx + fy * fy + fz * fz);
    count_wave(&counters[1], !(mg < 1e-5f) && mg >= threshold);
}

// SDFBoundary::closestTriangle (src/fluid/SDFBoundary.cpp:600-610) of every
// query point: the BVH walk in FP64 from the float point, the triangle's index
// in the build's face list and the closest point rounded to float.
__global__ __launch_bounds__(kMB) void k_closest(MeshView m, const float* __restrict__ pts, uint32_t n,
                                                 int32_t* __restrict__ face, float* __restrict__ out) {
    const uint32_t i = mgid();
    if (i >= n) return;
    const Q3 p{(double)pts[3 * (size_t)i], (double)pts[3 * (size_t)i + 1], (double)pts[3 * (size_t)i + 2]};
    const DHit hit = bvh_closest(m.nodes, m.tris, p);
    face[i] = hit.tri;
    out[3 * (size_t)i] = (float)hit.p.x;
    out[3 * (size_t)i + 1] = (float)hit.p.y;
    out[3 * (size_t)i + 2] = (float)hit.p.z;
}

}  // namespace

struct fsi_mesh {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ready = nullptr;  // after the last grid write
    hipEvent_t ev = nullptr;
    uint32_t nv = 0, nt = 0, nn = 0;
    int32_t levels = 0;
    double origin[3] = {0, 0, 0}, cell = 0.0;
    int32_t dims[3] = {0, 0, 0};
    size_t voxels = 0;
    DNode* nodes = nullptr;
    DTri* tris = nullptr;
    DNorm* norms = nullptr;
    double* vn = nullptr;
    int32_t* faces = nullptr;
    double* verts = nullptr;        // current
    double* verts_build = nullptr;  // baseline of the dirty refit
    double* verts_in = nullptr;     // staging of new vertices
    int32_t* level_nodes = nullptr;
    int32_t* level_off = nullptr;
    float* grid = nullptr;
    uint32_t* mask = nullptr;
    int32_t* dirty = nullptr;
    // contact scratch, sized for `cap` candidates
    uint32_t cap = 0;
    int key_bits = 1;
    float4* cf = nullptr;
    uint32_t* keys = nullptr;
    SortScratch sort{};
    int2* seg = nullptr;
    float4* forces = nullptr;
    uint32_t* counters = nullptr;
    // closest-triangle queries, sized for `qcap` points
    uint32_t qcap = 0;
    float* qpts = nullptr;
    int32_t* qface = nullptr;
    float* qout = nullptr;
};

namespace {

MeshView view(const fsi_mesh* m) {
    MeshView v{};
    v.nodes = m->nodes;
    v.tris = m->tris;
    v.norms = m->norms;
    v.vn = m->vn;
    v.faces = m->faces;
    for (int k = 0; k < 3; ++k) {
        v.origin[k] = m->origin[k];
        v.dims[k] = m->dims[k];
    }
    v.cell = m->cell;
    return v;
}

void free_contact(fsi_mesh* m) {
    dev_free(&m->cf);
    dev_free(&m->keys);
    dev_free(&m->sort.keys[0]);
    dev_free(&m->sort.keys[1]);
    dev_free(&m->sort.ids[0]);
    dev_free(&m->sort.ids[1]);
    dev_free(&m->sort.hist);
    dev_free(&m->sort.partial);
    m->cap = 0;
}

void free_mesh(fsi_mesh* m) {
    free_contact(m);
    dev_free(&m->nodes);
    dev_free(&m->tris);
    dev_free(&m->norms);
    dev_free(&m->vn);
    dev_free(&m->faces);
    dev_free(&m->verts);
    dev_free(&m->verts_build);
    dev_free(&m->verts_in);
    dev_free(&m->level_nodes);
    dev_free(&m->level_off);
    dev_free(&m->grid);
    dev_free(&m->mask);
    dev_free(&m->dirty);
    dev_free(&m->seg);
    dev_free(&m->forces);
    dev_free(&m->counters);
    dev_free(&m->qpts);
    dev_free(&m->qface);
    dev_free(&m->qout);
    m->qcap = 0;
}

int bind(fsi_mesh* m) {
    if (!m) return fail(FSI_ERR_INVALID_ARG, "mesh handle is NULL");
    FSI_HIP_TRY(hipSetDevice(m->device));
    return FSI_OK;
}

hipError_t record_ready(fsi_mesh* m) { return hipEventRecord(m->ready, m->stream); }

hipError_t launch_sdf(fsi_mesh* m, bool masked) {
    hipLaunchKernelGGL(k_sdf, mgrid(m->voxels), dim3(kMB), 0, m->stream, view(m), m->grid,
                       masked ? m->mask : nullptr, (uint32_t)m->voxels, m->dirty);
    return hipGetLastError();
}

// Copies new vertices in and runs updateGeometryAndRefit.
hipError_t update_geometry(fsi_mesh* m, const double* verts) {
    hipError_t e = hipMemcpyAsync(m->verts_in, verts, sizeof(double) * 3 * m->nv, hipMemcpyHostToDevice, m->stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_update_tris, mgrid(m->nt), dim3(kMB), 0, m->stream, m->verts_in, m->tris, m->norms, m->nt);
    hipLaunchKernelGGL(k_refit, dim3(1), dim3(kRefitBlock), 0, m->stream, m->nodes, m->tris, m->level_nodes,
                       m->level_off, m->levels);
    return hipGetLastError();
}

hipError_t ensure_contact(fsi_mesh* m, uint32_t cap) {
    if (cap <= m->cap) return hipSuccess;
    free_contact(m);
    const size_t nk = 3 * (size_t)cap;
    hipError_t e = hipSuccess;
#define ALLOC(p, cnt) \
    if ((e = dev_alloc(&(p), (cnt))) != hipSuccess) return e;
    ALLOC(m->cf, 2 * (size_t)cap);
    ALLOC(m->keys, nk);
    ALLOC(m->sort.keys[0], nk);
    ALLOC(m->sort.keys[1], nk);
    ALLOC(m->sort.ids[0], nk);
    ALLOC(m->sort.ids[1], nk);
    const size_t hist = sort_hist_entries((uint32_t)nk, kSortMaxDigitBits);
    ALLOC(m->sort.hist, hist);
    ALLOC(m->sort.partial, std::max(scan_partials_needed(hist), scan_partials_needed(nk)));
#undef ALLOC
    m->cap = cap;
    return hipSuccess;
}

}  // namespace

namespace fsi {

int mesh_contact_forces(fsi_mesh* m, const float4* contacts, const uint32_t* count, uint32_t cap, hipStream_t st,
                        float band, int use_vel_magnitude, float control_threshold, float* forces_host,
                        uint32_t counts[2]) {
    int rc = bind(m);
    if (rc) return rc;
    if (cap == 0) return fail(FSI_ERR_INVALID_ARG, "no contact capacity");
    if ((size_t)cap * 3 > 0xffffffffull) return fail(FSI_ERR_INVALID_ARG, "too many candidates");
    // the solver's stream follows the mesh's last update, and the mesh's
    // stream follows this pass
    FSI_HIP_TRY(hipEventRecord(m->ev, m->stream));
    FSI_HIP_TRY(hipStreamWaitEvent(st, m->ev, 0));
    if (cap > m->cap) {  // the scratch may still be read by an earlier pass
        FSI_HIP_TRY(hipStreamSynchronize(st));
        FSI_HIP_TRY(ensure_contact(m, cap));
    }
    const uint32_t nk = 3 * cap;
    hipLaunchKernelGGL(k_zero32, dim3(1), dim3(kMB), 0, st, m->counters, (size_t)2);
    hipLaunchKernelGGL(k_contact, mgrid(cap), dim3(kMB), 0, st, view(m), m->verts, contacts, count, cap, band,
                       use_vel_magnitude ? 1 : 0, m->nv, m->cf, m->keys, m->counters);
    uint32_t *skeys = nullptr, *sids = nullptr;
    FSI_HIP_TRY(radix_sort_pairs(m->keys, nk, m->key_bits, kSortMaxDigitBits, m->sort, &skeys, &sids, st));
    const uint32_t nseg = 1u << m->key_bits;
    hipLaunchKernelGGL(k_seg_clear, mgrid(nseg), dim3(kMB), 0, st, m->seg, nseg);
    hipLaunchKernelGGL(k_seg_build, mgrid(nk), dim3(kMB), 0, st, skeys, nk, m->seg);
    hipLaunchKernelGGL(k_vertex_sum, mgrid(m->nv), dim3(kMB), 0, st, m->seg, sids, m->cf, m->nv, control_threshold,
                       m->forces, m->counters);
    FSI_HIP_TRY(hipGetLastError());
    FSI_HIP_TRY(hipEventRecord(m->ev, st));
    FSI_HIP_TRY(hipStreamWaitEvent(m->stream, m->ev, 0));
    if (forces_host) {
        std::vector<float4> tmp(m->nv);
        FSI_HIP_TRY(hipMemcpyAsync(tmp.data(), m->forces, sizeof(float4) * m->nv, hipMemcpyDeviceToHost, st));
        FSI_HIP_TRY(hipStreamSynchronize(st));
        for (uint32_t v = 0; v < m->nv; ++v) {
            forces_host[3 * (size_t)v] = tmp[v].x;
            forces_host[3 * (size_t)v + 1] = tmp[v].y;
            forces_host[3 * (size_t)v + 2] = tmp[v].z;
        }
    }
    if (counts) {
        FSI_HIP_TRY(hipMemcpyAsync(counts, m->counters, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        FSI_HIP_TRY(hipStreamSynchronize(st));
    }
    return FSI_OK;
}

int mesh_device(const fsi_mesh* m) { return m ? m->device : -1; }

int mesh_wait_for(fsi_mesh* m, hipStream_t st) {
    int rc = bind(m);
    if (rc) return rc;
    FSI_HIP_TRY(hipEventRecord(m->ev, st));
    FSI_HIP_TRY(hipStreamWaitEvent(m->stream, m->ev, 0));
    return FSI_OK;
}

int mesh_grid_for_solver(fsi_mesh* m, const float** grid, int32_t dims[3], double origin[3], double* cell,
                         hipEvent_t* ready) {
    int rc = bind(m);
    if (rc) return rc;
    *grid = m->grid;
    for (int k = 0; k < 3; ++k) {
        dims[k] = m->dims[k];
        origin[k] = m->origin[k];
    }
    *cell = m->cell;
    *ready = m->ready;
    return FSI_OK;
}

}  // namespace fsi

// ---- contact forces without allocation --------------------------------------
// fsi_mesh_contact_forces_noalloc runs the kernels of fsi_pbf_contact_forces
// on caller-given device candidates, with its scratch (contact records, sort
// buffers, histogram, vertex segments) in one fixed device array instead of
// buffers the mesh allocates.  Limits: kNaMaxContacts candidates per call and
// kNaMaxVerts mesh vertices (keys 0..nverts fit 16 bits).
namespace {

constexpr uint32_t kNaMaxContacts = 65536;
constexpr uint32_t kNaMaxVerts = 65535;
constexpr size_t kNaKeys = 3 * (size_t)kNaMaxContacts;
constexpr size_t kNaHist = ((size_t)1 << fsi::kSortMaxDigitBits) * ((kNaKeys + 2047) / 2048);
constexpr size_t kNaPartial = 1024;
constexpr size_t kNaSeg = (size_t)1 << 16;
constexpr size_t kNaOffCf = 0;
constexpr size_t kNaOffKeys = kNaOffCf + 2 * (size_t)kNaMaxContacts * sizeof(float4);
constexpr size_t kNaOffK0 = kNaOffKeys + kNaKeys * sizeof(uint32_t);
constexpr size_t kNaOffK1 = kNaOffK0 + kNaKeys * sizeof(uint32_t);
constexpr size_t kNaOffI0 = kNaOffK1 + kNaKeys * sizeof(uint32_t);
constexpr size_t kNaOffI1 = kNaOffI0 + kNaKeys * sizeof(uint32_t);
constexpr size_t kNaOffHist = kNaOffI1 + kNaKeys * sizeof(uint32_t);
constexpr size_t kNaOffPartial = kNaOffHist + kNaHist * sizeof(uint32_t);
constexpr size_t kNaOffSeg = kNaOffPartial + kNaPartial * sizeof(uint32_t);
constexpr size_t kNaBytes = kNaOffSeg + kNaSeg * sizeof(int2);  // 6.6 MiB
static_assert(kNaOffKeys % 256 == 0 && kNaOffK0 % 256 == 0 && kNaOffK1 % 256 == 0 && kNaOffI0 % 256 == 0 &&
                  kNaOffI1 % 256 == 0 && kNaOffHist % 256 == 0 && kNaOffPartial % 256 == 0 && kNaOffSeg % 256 == 0,
              "scratch sections start on 256-byte boundaries");

// The scratch: one copy per device, held until the process ends.  Host code
// never uses this name as an address (it is a host-side stand-in); it uses
// the device address noalloc_scratch returns.
__device__ __align__(256) unsigned char scratch[kNaBytes];

constexpr int kNaDevices = 64;
std::mutex g_na_mu;
unsigned char* g_na_base[kNaDevices] = {};

// The address of scratch on `device`, which must be current: looked up once
// per device (hipGetSymbolAddress) and kept.
int noalloc_scratch(int device, unsigned char** out) {
    if (device < 0 || device >= kNaDevices) return fsi::fail(FSI_ERR_INVALID_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_na_mu);
    if (!g_na_base[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return fsi::fail(FSI_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_na_base[device] = static_cast<unsigned char*>(p);
    }
    *out = g_na_base[device];
    return FSI_OK;
}

}  // namespace

extern "C" {

int fsi_mesh_create(const double* verts, uint32_t nverts, const int32_t* faces, uint32_t nfaces, double cell,
                    fsi_mesh** out) {
    if (!out) return fail(FSI_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (!verts || !faces || nverts == 0 || nfaces == 0) return fail(FSI_ERR_INVALID_ARG, "empty mesh");
    if (!(cell > 0.0)) return fail(FSI_ERR_INVALID_ARG, "cell must be positive");
    if (nverts >= (1u << 30) || nfaces >= (1u << 28)) return fail(FSI_ERR_INVALID_ARG, "mesh too large");
    for (uint32_t i = 0; i < 3 * nfaces; ++i)
        if (faces[i] < 0 || (uint32_t)faces[i] >= nverts) return fail(FSI_ERR_INVALID_ARG, "face index out of range");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(FSI_ERR_NO_DEVICE, "no HIP device visible");
    // grid placement of SDFBoundary::buildFromMesh (SDFBoundary.cpp:290-308)
    mesh::D3 lo{1e30, 1e30, 1e30}, hi{-1e30, -1e30, -1e30};
    for (uint32_t i = 0; i < nverts; ++i) {
        const mesh::D3 p{verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
        lo = mesh::vmin(lo, p);
        hi = mesh::vmax(hi, p);
    }
    const double buffer = cell * 5.0;
    const mesh::D3 o{lo.x - buffer, lo.y - buffer, lo.z - buffer};
    const mesh::D3 span{(hi.x + buffer) - o.x, (hi.y + buffer) - o.y, (hi.z + buffer) - o.z};
    const double dd[3] = {std::ceil(span.x / cell), std::ceil(span.y / cell), std::ceil(span.z / cell)};
    if (dd[0] * dd[1] * dd[2] > 1.0 * (1u << 30)) return fail(FSI_ERR_INVALID_ARG, "SDF grid above 2^30 voxels");

    mesh::Mesh hm;
    hm.build(verts, nverts, faces, nfaces);
    const std::vector<int> depth = hm.bvh.depths();
    int32_t levels = 0;
    for (int d : depth) levels = std::max(levels, d + 1);
    if (levels + 1 >= kStack) return fail(FSI_ERR_INVALID_ARG, "BVH deeper than the device walk allows");

    fsi_mesh* m = new (std::nothrow) fsi_mesh();
    if (!m) return fail(FSI_ERR_INVALID_ARG, "out of host memory");
    auto bail = [&](int code) {
        if (m->stream) (void)hipStreamSynchronize(m->stream);
        free_mesh(m);
        if (m->ready) (void)hipEventDestroy(m->ready);
        if (m->ev) (void)hipEventDestroy(m->ev);
        if (m->stream) (void)hipStreamDestroy(m->stream);
        delete m;
        return code;
    };
    if (hipGetDevice(&m->device) != hipSuccess) return bail(fail(FSI_ERR_HIP, "hipGetDevice"));
    m->nv = nverts;
    m->nt = nfaces;
    m->nn = (uint32_t)hm.bvh.nodes.size();
    m->levels = levels;
    m->cell = cell;
    m->origin[0] = o.x;
    m->origin[1] = o.y;
    m->origin[2] = o.z;
    for (int k = 0; k < 3; ++k) m->dims[k] = (int32_t)dd[k];
    m->voxels = (size_t)m->dims[0] * m->dims[1] * m->dims[2];
    int bits = 1;
    while ((1u << bits) <= nverts) ++bits;  // keys 0..nverts (nverts: no contact)
    m->key_bits = bits;

    // host images
    std::vector<DNode> nodes(m->nn);
    for (uint32_t i = 0; i < m->nn; ++i) {
        const auto& s = hm.bvh.nodes[i];
        DNode& d = nodes[i];
        d.lo[0] = s.lo.x; d.lo[1] = s.lo.y; d.lo[2] = s.lo.z;
        d.hi[0] = s.hi.x; d.hi[1] = s.hi.y; d.hi[2] = s.hi.z;
        d.start = s.start;
        d.count = s.count;
        d.left = s.left;
        d.right = s.right;
    }
    auto put = [](double* p, mesh::D3 v) {
        p[0] = v.x;
        p[1] = v.y;
        p[2] = v.z;
    };
    std::vector<DTri> tris(nfaces);
    std::vector<DNorm> norms(nfaces);
    for (uint32_t k = 0; k < nfaces; ++k) {
        const int id = hm.bvh.order[k];
        const mesh::Tri& t = hm.tris[id];
        put(tris[k].a, t.a);
        put(tris[k].b, t.b);
        put(tris[k].c, t.c);
        tris[k].id = id;
        tris[k].ia = t.ia;
        tris[k].ib = t.ib;
        tris[k].ic = t.ic;
    }
    for (uint32_t id = 0; id < nfaces; ++id) {
        const mesh::Tri& t = hm.tris[id];
        put(norms[id].fn, t.fn);
        put(norms[id].ab, t.en_ab);
        put(norms[id].ac, t.en_ac);
        put(norms[id].bc, t.en_bc);
    }
    std::vector<double> vn(3 * (size_t)nverts);
    for (uint32_t v = 0; v < nverts; ++v) put(&vn[3 * (size_t)v], hm.vn[v]);
    std::vector<int32_t> level_off(levels + 1, 0), level_nodes(m->nn);
    for (int d : depth) ++level_off[d + 1];
    for (int d = 0; d < levels; ++d) level_off[d + 1] += level_off[d];
    {
        std::vector<int32_t> fill(level_off.begin(), level_off.end() - 1);
        for (uint32_t i = 0; i < m->nn; ++i) level_nodes[fill[depth[i]]++] = (int32_t)i;
    }

    hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&m->ready, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&m->ev, hipEventDisableTiming);
#define ALLOC(p, cnt) \
    if (e == hipSuccess) e = dev_alloc(&(p), (cnt));
    ALLOC(m->nodes, m->nn);
    ALLOC(m->tris, nfaces);
    ALLOC(m->norms, nfaces);
    ALLOC(m->vn, 3 * (size_t)nverts);
    ALLOC(m->faces, 3 * (size_t)nfaces);
    ALLOC(m->verts, 3 * (size_t)nverts);
    ALLOC(m->verts_build, 3 * (size_t)nverts);
    ALLOC(m->verts_in, 3 * (size_t)nverts);
    ALLOC(m->level_nodes, m->nn);
    ALLOC(m->level_off, (size_t)levels + 1);
    ALLOC(m->grid, m->voxels);
    ALLOC(m->mask, m->voxels);
    ALLOC(m->dirty, kDirtyWords);
    ALLOC(m->seg, (size_t)1 << bits);
    ALLOC(m->forces, nverts);
    ALLOC(m->counters, 2);
#undef ALLOC
    hipStream_t st = m->stream;
    if (e == hipSuccess) e = hipMemcpyAsync(m->nodes, nodes.data(), sizeof(DNode) * m->nn, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(m->tris, tris.data(), sizeof(DTri) * nfaces, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(m->norms, norms.data(), sizeof(DNorm) * nfaces, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(m->vn, vn.data(), sizeof(double) * vn.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(m->faces, faces, sizeof(int32_t) * 3 * (size_t)nfaces, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(m->verts, verts, sizeof(double) * 3 * (size_t)nverts, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(m->verts_build, verts, sizeof(double) * 3 * (size_t)nverts, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(m->level_nodes, level_nodes.data(), sizeof(int32_t) * m->nn, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(m->level_off, level_off.data(), sizeof(int32_t) * (levels + 1), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_zero32, mgrid(m->voxels), dim3(kMB), 0, st, m->mask, m->voxels);
        e = launch_sdf(m, false);
    }
    if (e == hipSuccess) e = record_ready(m);
    if (e == hipSuccess) e = hipStreamSynchronize(st);  // the host images go out of scope
    if (e != hipSuccess) return bail(fail(FSI_ERR_HIP, std::string("mesh upload: ") + hipGetErrorString(e)));
    // set-up of fsi_mesh_contact_forces_noalloc on this device (no allocation)
    unsigned char* na = nullptr;
    const int rc = noalloc_scratch(m->device, &na);
    if (rc) return bail(rc);
    *out = m;
    return FSI_OK;
}

int fsi_mesh_destroy(fsi_mesh* m) {
    if (!m) return FSI_OK;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    free_mesh(m);
    if (m->ready) (void)hipEventDestroy(m->ready);
    if (m->ev) (void)hipEventDestroy(m->ev);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return FSI_OK;
}

int fsi_mesh_grid_info(fsi_mesh* m, double origin[3], double* cell, int32_t dims[3]) {
    if (!m) return fail(FSI_ERR_INVALID_ARG, "mesh handle is NULL");
    for (int k = 0; k < 3; ++k) {
        if (origin) origin[k] = m->origin[k];
        if (dims) dims[k] = m->dims[k];
    }
    if (cell) *cell = m->cell;
    return FSI_OK;
}

int fsi_mesh_read_sdf(fsi_mesh* m, float* grid) {
    int rc = bind(m);
    if (rc) return rc;
    if (!grid) return fail(FSI_ERR_INVALID_ARG, "grid is NULL");
    FSI_HIP_TRY(hipMemcpyAsync(grid, m->grid, sizeof(float) * m->voxels, hipMemcpyDeviceToHost, m->stream));
    FSI_HIP_TRY(hipStreamSynchronize(m->stream));
    return FSI_OK;
}

int fsi_mesh_read_forces(fsi_mesh* m, float* forces) {
    int rc = bind(m);
    if (rc) return rc;
    if (!forces) return fail(FSI_ERR_INVALID_ARG, "forces is NULL");
    std::vector<float4> tmp(m->nv);
    FSI_HIP_TRY(hipMemcpyAsync(tmp.data(), m->forces, sizeof(float4) * m->nv, hipMemcpyDeviceToHost, m->stream));
    FSI_HIP_TRY(hipStreamSynchronize(m->stream));
    for (uint32_t v = 0; v < m->nv; ++v) {
        forces[3 * (size_t)v] = tmp[v].x;
        forces[3 * (size_t)v + 1] = tmp[v].y;
        forces[3 * (size_t)v + 2] = tmp[v].z;
    }
    return FSI_OK;
}

int fsi_mesh_refit(fsi_mesh* m, const double* verts) {
    int rc = bind(m);
    if (rc) return rc;
    if (!verts) return fail(FSI_ERR_INVALID_ARG, "verts is NULL");
    FSI_HIP_TRY(update_geometry(m, verts));
    FSI_HIP_TRY(launch_sdf(m, false));
    FSI_HIP_TRY(hipMemcpyAsync(m->verts, m->verts_in, sizeof(double) * 3 * m->nv, hipMemcpyDeviceToDevice, m->stream));
    FSI_HIP_TRY(hipMemcpyAsync(m->verts_build, m->verts_in, sizeof(double) * 3 * m->nv, hipMemcpyDeviceToDevice,
                               m->stream));
    FSI_HIP_TRY(record_ready(m));
    FSI_HIP_TRY(hipStreamSynchronize(m->stream));  // verts is the caller's memory
    return FSI_OK;
}

int fsi_mesh_closest(fsi_mesh* m, const float* points, uint32_t n, int32_t* faces, float* closest) {
    int rc = bind(m);
    if (rc) return rc;
    if (n == 0) return FSI_OK;
    if (!points || !faces || !closest) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    if (n > (1u << 28)) return fail(FSI_ERR_INVALID_ARG, "more than 2^28 query points");
    hipStream_t st = m->stream;
    if (n > m->qcap) {
        FSI_HIP_TRY(hipStreamSynchronize(st));
        dev_free(&m->qpts);
        dev_free(&m->qface);
        dev_free(&m->qout);
        m->qcap = 0;
        FSI_HIP_TRY(dev_alloc(&m->qpts, 3 * (size_t)n));
        FSI_HIP_TRY(dev_alloc(&m->qface, (size_t)n));
        FSI_HIP_TRY(dev_alloc(&m->qout, 3 * (size_t)n));
        m->qcap = n;
    }
    FSI_HIP_TRY(hipMemcpyAsync(m->qpts, points, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_closest, mgrid(n), dim3(kMB), 0, st, view(m), m->qpts, n, m->qface, m->qout);
    FSI_HIP_TRY(hipGetLastError());
    FSI_HIP_TRY(hipMemcpyAsync(faces, m->qface, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, st));
    FSI_HIP_TRY(hipMemcpyAsync(closest, m->qout, sizeof(float) * 3 * (size_t)n, hipMemcpyDeviceToHost, st));
    FSI_HIP_TRY(hipStreamSynchronize(st));
    return FSI_OK;
}

int fsi_mesh_refit_dirty(fsi_mesh* m, const double* verts, double band_radius, double move_eps,
                         int32_t dirty_min[3], int32_t dirty_max[3], uint32_t* dirty_voxels) {
    int rc = bind(m);
    if (rc) return rc;
    if (!verts) return fail(FSI_ERR_INVALID_ARG, "verts is NULL");
    if (!(band_radius >= 0.0) || !(move_eps >= 0.0)) return fail(FSI_ERR_INVALID_ARG, "negative band or eps");
    hipStream_t st = m->stream;
    FSI_HIP_TRY(hipMemcpyAsync(m->verts_in, verts, sizeof(double) * 3 * m->nv, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_dirty_init, dim3(1), dim3(kMB), 0, st, m->dirty);
    hipLaunchKernelGGL(k_mark, mgrid(64 * (size_t)m->nv), dim3(kMB), 0, st, m->verts_build, m->verts_in, m->nv,
                       band_radius, move_eps * move_eps, view(m), m->mask);
    hipLaunchKernelGGL(k_update_tris, mgrid(m->nt), dim3(kMB), 0, st, m->verts_in, m->tris, m->norms, m->nt);
    hipLaunchKernelGGL(k_refit, dim3(1), dim3(kRefitBlock), 0, st, m->nodes, m->tris, m->level_nodes, m->level_off,
                       m->levels);
    FSI_HIP_TRY(hipGetLastError());
    FSI_HIP_TRY(launch_sdf(m, true));
    FSI_HIP_TRY(hipMemcpyAsync(m->verts, m->verts_in, sizeof(double) * 3 * m->nv, hipMemcpyDeviceToDevice, st));
    FSI_HIP_TRY(hipMemcpyAsync(m->verts_build, m->verts_in, sizeof(double) * 3 * m->nv, hipMemcpyDeviceToDevice, st));
    FSI_HIP_TRY(record_ready(m));
    std::vector<int32_t> w(kDirtyWords);
    FSI_HIP_TRY(hipMemcpyAsync(w.data(), m->dirty, sizeof(int32_t) * kDirtyWords, hipMemcpyDeviceToHost, st));
    FSI_HIP_TRY(hipStreamSynchronize(st));
    int32_t d[7] = {0, INT32_MAX, INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN, INT32_MIN};
    for (uint32_t q = 0; q < kSpread; ++q) {
        d[0] += w[q];
        for (int k = 1; k < 4; ++k) d[k] = std::min(d[k], w[k * kSpread + q]);
        for (int k = 4; k < 7; ++k) d[k] = std::max(d[k], w[k * kSpread + q]);
    }
    if (dirty_voxels) *dirty_voxels = (uint32_t)d[0];
    for (int k = 0; k < 3; ++k) {
        if (dirty_min) dirty_min[k] = d[0] ? d[1 + k] : 0;
        if (dirty_max) dirty_max[k] = d[0] ? d[4 + k] : -1;
    }
    return FSI_OK;
}

int fsi_mesh_contact_forces_noalloc(fsi_mesh* m, const fsi_contact_candidate* candidates, uint32_t count,
                                    float band, int use_vel_magnitude, float control_threshold, float* forces,
                                    uint32_t* counts, void* stream) {
    int rc = bind(m);
    if (rc) return rc;
    if (!forces || !counts) return fail(FSI_ERR_INVALID_ARG, "forces and counts (device memory) are required");
    if ((reinterpret_cast<uintptr_t>(forces) & 15u) != 0) return fail(FSI_ERR_INVALID_ARG, "forces must be 16-byte aligned");
    if (count > 0 && !candidates) return fail(FSI_ERR_INVALID_ARG, "candidates is NULL");
    if (count > kNaMaxContacts)
        return fail(FSI_ERR_INVALID_ARG, "fsi_mesh_contact_forces_noalloc: " + std::to_string(count) +
                                             " candidates, at most " + std::to_string(kNaMaxContacts));
    if (m->nv > kNaMaxVerts)
        return fail(FSI_ERR_INVALID_ARG, "fsi_mesh_contact_forces_noalloc: mesh of " + std::to_string(m->nv) +
                                             " vertices, at most " + std::to_string(kNaMaxVerts));
    if (sort_hist_entries((uint32_t)kNaKeys, kSortMaxDigitBits) > kNaHist ||
        scan_partials_needed(std::max(kNaHist, kNaKeys)) > kNaPartial)
        return fail(FSI_ERR_STATE, "no-allocation scratch is smaller than the sort needs");
    unsigned char* base = nullptr;
    if ((rc = noalloc_scratch(m->device, &base))) return rc;
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : m->stream;
    if (st != m->stream) {  // after the mesh's last update
        FSI_HIP_TRY(hipEventRecord(m->ev, m->stream));
        FSI_HIP_TRY(hipStreamWaitEvent(st, m->ev, 0));
    }
    float4* cf = reinterpret_cast<float4*>(base + kNaOffCf);
    uint32_t* keys = reinterpret_cast<uint32_t*>(base + kNaOffKeys);
    SortScratch ss{};
    ss.keys[0] = reinterpret_cast<uint32_t*>(base + kNaOffK0);
    ss.keys[1] = reinterpret_cast<uint32_t*>(base + kNaOffK1);
    ss.ids[0] = reinterpret_cast<uint32_t*>(base + kNaOffI0);
    ss.ids[1] = reinterpret_cast<uint32_t*>(base + kNaOffI1);
    ss.hist = reinterpret_cast<uint32_t*>(base + kNaOffHist);
    ss.partial = reinterpret_cast<uint32_t*>(base + kNaOffPartial);
    int2* seg = reinterpret_cast<int2*>(base + kNaOffSeg);
    float4* fout = reinterpret_cast<float4*>(forces);
    // every call starts from its own inputs: the counters and (for no
    // candidates) the forces are cleared here, the rest is overwritten
    hipLaunchKernelGGL(k_zero32, dim3(1), dim3(kMB), 0, st, counts, (size_t)2);
    if (count == 0) {
        hipLaunchKernelGGL(k_zero32, mgrid(4 * (size_t)m->nv), dim3(kMB), 0, st, reinterpret_cast<uint32_t*>(forces),
                           4 * (size_t)m->nv);
    } else {
        const uint32_t nk = 3 * count;
        hipLaunchKernelGGL(k_contact, mgrid(count), dim3(kMB), 0, st, view(m), m->verts,
                           reinterpret_cast<const float4*>(candidates), nullptr, count, band,
                           use_vel_magnitude ? 1 : 0, m->nv, cf, keys, counts);
        uint32_t *skeys = nullptr, *sids = nullptr;
        FSI_HIP_TRY(radix_sort_pairs(keys, nk, m->key_bits, kSortMaxDigitBits, ss, &skeys, &sids, st));
        const uint32_t nseg = 1u << m->key_bits;
        hipLaunchKernelGGL(k_seg_clear, mgrid(nseg), dim3(kMB), 0, st, seg, nseg);
        hipLaunchKernelGGL(k_seg_build, mgrid(nk), dim3(kMB), 0, st, skeys, nk, seg);
        hipLaunchKernelGGL(k_vertex_sum, mgrid(m->nv), dim3(kMB), 0, st, seg, sids, cf, m->nv, control_threshold, fout,
                           counts);
    }
    FSI_HIP_TRY(hipGetLastError());
    if (st != m->stream) {  // a later mesh update waits for this pass
        FSI_HIP_TRY(hipEventRecord(m->ev, st));
        FSI_HIP_TRY(hipStreamWaitEvent(m->stream, m->ev, 0));
    }
    return FSI_OK;
}

}  // extern "C"
