/*
 * mesh_oracle.c — CPU oracle of the FSI coupling (TEST INFRASTRUCTURE ONLY).
 *
 * Plain-C restatement, by brute force over every triangle (no BVH), of:
 *   - SDFBoundary's signed distance (src/fluid/SDFBoundary.cpp): closest
 *     point on a triangle (closestPointOnTriangle :22-54, Ericson 5.1.5),
 *     face / edge / angle-weighted vertex pseudo-normals of buildFromMesh
 *     (:318-383), the sign of signedDistanceAt (:400-419), and the normals a
 *     refit keeps (updateGeometryAndRefit :438-452: face normals refreshed,
 *     edge and vertex normals from the build);
 *   - the dirty-voxel marking of refitDirty (:480-525);
 *   - closestTriangle (:600-610);
 *   - computeContactForces (src/main.cpp:604-710) with barycentric
 *     (:557-571), per-vertex sums in candidate order.
 * Only tests/ and bench.py's CPU baseline use it.  Among equidistant
 * triangles it keeps the lowest index (the BVH keeps the first it meets);
 * distances agree, and the split of a force onto a shared edge or vertex
 * agrees up to rounding.
 */
#include "mesh_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double x, y, z;
} v3;

static v3 mk(double x, double y, double z) {
    v3 r = {x, y, z};
    return r;
}
static v3 ld(const double* p) { return mk(p[0], p[1], p[2]); }
static v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 scl(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static v3 dvd(v3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
static double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static v3 cross(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static double norm(v3 a) { return sqrt(dot(a, a)); }
/* Eigen's normalized(): v / |v| when |v|^2 > 0 */
static v3 normalized(v3 v) {
    const double z = dot(v, v);
    return z > 0.0 ? dvd(v, sqrt(z)) : v;
}

enum { FACE = 0, EDGE_AB, EDGE_AC, EDGE_BC, VERT_A, VERT_B, VERT_C };

/* SDFBoundary.cpp:22-54 */
static v3 closest_on_tri(v3 p, v3 a, v3 b, v3 c, int* f) {
    const v3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) { *f = VERT_A; return a; }
    const v3 bp = sub(p, b);
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) { *f = VERT_B; return b; }
    const v3 cp = sub(p, c);
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) { *f = VERT_C; return c; }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) { *f = EDGE_AB; return add(a, scl(ab, d1 / (d1 - d3))); }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) { *f = EDGE_AC; return add(a, scl(ac, d2 / (d2 - d6))); }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        *f = EDGE_BC;
        return add(b, scl(sub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
    }
    const double denom = 1.0 / (va + vb + vc);
    *f = FACE;
    return add(add(a, scl(ab, vb * denom)), scl(ac, vc * denom));
}

static v3 corner(const double* verts, const int32_t* faces, uint32_t t, int k) {
    return ld(verts + 3 * (size_t)faces[3 * (size_t)t + k]);
}

/* (B - A) x (C - A) / |.| when |.| > 1e-12 (SDFBoundary.cpp:331-334, :446-449) */
static v3 face_normal(const double* verts, const int32_t* faces, uint32_t t) {
    const v3 a = corner(verts, faces, t, 0), b = corner(verts, faces, t, 1), c = corner(verts, faces, t, 2);
    v3 n = cross(sub(b, a), sub(c, a));
    const double len = norm(n);
    if (len > 1e-12) n = dvd(n, len);
    return n;
}

typedef struct {
    int32_t a, b;  /* a < b */
    uint32_t tri;
    int slot;      /* 0 AB, 1 AC, 2 BC */
} edge_rec;

static int edge_cmp(const void* pa, const void* pb) {
    const edge_rec* x = (const edge_rec*)pa;
    const edge_rec* y = (const edge_rec*)pb;
    if (x->a != y->a) return x->a < y->a ? -1 : 1;
    if (x->b != y->b) return x->b < y->b ? -1 : 1;
    if (x->tri != y->tri) return x->tri < y->tri ? -1 : 1;
    return x->slot - y->slot;
}

/* Pseudo-normals of buildFromMesh (SDFBoundary.cpp:318-383) from `verts`:
 * en: 3 edge normals per triangle (AB, AC, BC; the face normal for an edge
 * not shared by exactly two faces), vn: 3 per vertex. */
static void build_normals(const double* verts, uint32_t nv, const int32_t* faces, uint32_t nf, double* en,
                          double* vn) {
    v3* fn = (v3*)malloc(sizeof(v3) * nf);
    for (uint32_t t = 0; t < nf; ++t) {
        fn[t] = face_normal(verts, faces, t);
        for (int s = 0; s < 3; ++s) {
            en[9 * (size_t)t + 3 * s] = fn[t].x;
            en[9 * (size_t)t + 3 * s + 1] = fn[t].y;
            en[9 * (size_t)t + 3 * s + 2] = fn[t].z;
        }
    }
    memset(vn, 0, sizeof(double) * 3 * nv);
    for (uint32_t t = 0; t < nf; ++t) {
        const int32_t* f = faces + 3 * (size_t)t;
        const v3 p[3] = {corner(verts, faces, t, 0), corner(verts, faces, t, 1), corner(verts, faces, t, 2)};
        /* accumulate(iA, A, B, C), (iB, B, A, C), (iC, C, A, B) */
        const int other[3][2] = {{1, 2}, {0, 2}, {0, 1}};
        for (int k = 0; k < 3; ++k) {
            const v3 u = normalized(sub(p[other[k][0]], p[k])), w = normalized(sub(p[other[k][1]], p[k]));
            double d = dot(u, w);
            d = d < -1.0 ? -1.0 : (d > 1.0 ? 1.0 : d);
            const double ang = acos(d);
            double* q = vn + 3 * (size_t)f[k];
            q[0] = q[0] + ang * fn[t].x;
            q[1] = q[1] + ang * fn[t].y;
            q[2] = q[2] + ang * fn[t].z;
        }
    }
    for (uint32_t v = 0; v < nv; ++v) {
        const v3 n = ld(vn + 3 * (size_t)v);
        const double len = norm(n);
        if (len > 1e-12) {
            const v3 m = dvd(n, len);
            vn[3 * (size_t)v] = m.x;
            vn[3 * (size_t)v + 1] = m.y;
            vn[3 * (size_t)v + 2] = m.z;
        }
    }
    edge_rec* e = (edge_rec*)malloc(sizeof(edge_rec) * 3 * (size_t)nf);
    for (uint32_t t = 0; t < nf; ++t) {
        const int32_t* f = faces + 3 * (size_t)t;
        const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        for (int s = 0; s < 3; ++s) {
            const int32_t a = f[pairs[s][0]], b = f[pairs[s][1]];
            edge_rec r = {a < b ? a : b, a < b ? b : a, t, s};
            e[3 * (size_t)t + s] = r;
        }
    }
    qsort(e, 3 * (size_t)nf, sizeof(edge_rec), edge_cmp);
    for (size_t i = 0; i < 3 * (size_t)nf;) {
        size_t j = i;
        while (j < 3 * (size_t)nf && e[j].a == e[i].a && e[j].b == e[i].b) ++j;
        if (j - i == 2) {
            const v3 s = add(fn[e[i].tri], fn[e[i + 1].tri]);
            const double len = norm(s);
            if (!(len < 1e-12)) {
                const v3 n = dvd(s, len);
                for (size_t k = i; k < j; ++k) {
                    double* q = en + 9 * (size_t)e[k].tri + 3 * e[k].slot;
                    q[0] = n.x;
                    q[1] = n.y;
                    q[2] = n.z;
                }
            }
        }
        i = j;
    }
    free(e);
    free(fn);
}

typedef struct {
    double d2;
    v3 p;
    int feature;
    int64_t tri;
} hit_t;

static hit_t brute_closest(const double* verts, const int32_t* faces, uint32_t nf, v3 p) {
    hit_t best = {1e30, {0, 0, 0}, FACE, -1};
    for (uint32_t t = 0; t < nf; ++t) {
        int f;
        const v3 q = closest_on_tri(p, corner(verts, faces, t, 0), corner(verts, faces, t, 1),
                                    corner(verts, faces, t, 2), &f);
        const v3 d = sub(p, q);
        const double d2 = dot(d, d);
        if (d2 < best.d2) {
            best.d2 = d2;
            best.p = q;
            best.feature = f;
            best.tri = t;
        }
    }
    return best;
}

/* SDFBoundary::closestTriangle (src/fluid/SDFBoundary.cpp:600-610): the
 * nearest triangle of each float point (promoted to double), by brute force;
 * the closest point is rounded to float as the reference returns it. */
void oracle_closest_triangles(const double* verts, const int32_t* faces, uint32_t nf, const float* pts, uint32_t n,
                              int32_t* face_out, float* point_out, double* d2_out) {
    for (uint32_t i = 0; i < n; ++i) {
        const hit_t h = brute_closest(verts, faces, nf, mk(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
        face_out[i] = (int32_t)h.tri;
        point_out[3 * i] = (float)h.p.x;
        point_out[3 * i + 1] = (float)h.p.y;
        point_out[3 * i + 2] = (float)h.p.z;
        d2_out[i] = h.d2;
    }
}

void oracle_mesh_sdf(const double* verts, const double* verts_build, uint32_t nv, const int32_t* faces, uint32_t nf,
                     const double origin[3], double cell, const int32_t dims[3], float* grid) {
    double* en = (double*)malloc(sizeof(double) * 9 * (size_t)nf);
    double* vn = (double*)malloc(sizeof(double) * 3 * (size_t)nv);
    build_normals(verts_build, nv, faces, nf, en, vn);
    const int32_t W = dims[0], H = dims[1], D = dims[2];
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t z = 0; z < D; ++z)
        for (int32_t y = 0; y < H; ++y)
            for (int32_t x = 0; x < W; ++x) {
                const v3 p = mk(origin[0] + x * cell, origin[1] + y * cell, origin[2] + z * cell);
                const hit_t h = brute_closest(verts, faces, nf, p);
                v3 pn;
                const int32_t* f = faces + 3 * (size_t)h.tri;
                switch (h.feature) {
                    case FACE: pn = face_normal(verts, faces, (uint32_t)h.tri); break;
                    case EDGE_AB: pn = ld(en + 9 * (size_t)h.tri); break;
                    case EDGE_AC: pn = ld(en + 9 * (size_t)h.tri + 3); break;
                    case EDGE_BC: pn = ld(en + 9 * (size_t)h.tri + 6); break;
                    case VERT_A: pn = ld(vn + 3 * (size_t)f[0]); break;
                    case VERT_B: pn = ld(vn + 3 * (size_t)f[1]); break;
                    default: pn = ld(vn + 3 * (size_t)f[2]); break;
                }
                const double sign = dot(sub(p, h.p), pn) >= 0.0 ? 1.0 : -1.0;
                grid[(size_t)x + (size_t)W * ((size_t)y + (size_t)H * (size_t)z)] = (float)(sign * sqrt(h.d2));
            }
    free(en);
    free(vn);
}

/* markBand of refitDirty (SDFBoundary.cpp:492-516) */
static void mark_band(const double* w, double band, const double origin[3], double cell, const int32_t dims[3],
                      uint8_t* mask, uint32_t* count, int32_t box[6]) {
    int32_t b0[3], b1[3];
    for (int k = 0; k < 3; ++k) {
        const double lo = ((w[k] - band) - origin[k]) / cell;
        const double hi = ((w[k] + band) - origin[k]) / cell;
        int32_t i0 = (int32_t)floor(lo), i1 = (int32_t)ceil(hi);
        i0 = i0 < 0 ? 0 : (i0 > dims[k] - 1 ? dims[k] - 1 : i0);
        i1 = i1 < 0 ? 0 : (i1 > dims[k] - 1 ? dims[k] - 1 : i1);
        b0[k] = i0;
        b1[k] = i1;
    }
    for (int32_t x = b0[0]; x <= b1[0]; ++x)
        for (int32_t y = b0[1]; y <= b1[1]; ++y)
            for (int32_t z = b0[2]; z <= b1[2]; ++z) {
                const size_t v = (size_t)x + (size_t)dims[0] * ((size_t)y + (size_t)dims[1] * (size_t)z);
                if (mask[v]) continue;
                mask[v] = 1;
                ++*count;
            }
    for (int k = 0; k < 3; ++k) {
        if (b0[k] < box[k]) box[k] = b0[k];
        if (b1[k] > box[3 + k]) box[3 + k] = b1[k];
    }
}

uint32_t oracle_dirty_mask(const double* old_verts, const double* new_verts, uint32_t nv, double band,
                           double move_eps, const double origin[3], double cell, const int32_t dims[3],
                           uint8_t* mask, int32_t box[6]) {
    const size_t total = (size_t)dims[0] * dims[1] * dims[2];
    memset(mask, 0, total);
    uint32_t count = 0;
    for (int k = 0; k < 3; ++k) {
        box[k] = 0x7fffffff;
        box[3 + k] = -0x7fffffff - 1;
    }
    const double eps2 = move_eps * move_eps;
    for (uint32_t i = 0; i < nv; ++i) {
        const v3 d = sub(ld(new_verts + 3 * (size_t)i), ld(old_verts + 3 * (size_t)i));
        if (dot(d, d) <= eps2) continue;
        mark_band(old_verts + 3 * (size_t)i, band, origin, cell, dims, mask, &count, box);
        mark_band(new_verts + 3 * (size_t)i, band, origin, cell, dims, mask, &count, box);
    }
    return count;
}

/* barycentric (src/main.cpp:557-571) */
static void barycentric(v3 pt, v3 a, v3 b, v3 c, float* u, float* v, float* w) {
    const v3 v0 = sub(b, a), v1 = sub(c, a), v2 = sub(pt, a);
    const double d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1);
    const double d20 = dot(v2, v0), d21 = dot(v2, v1);
    const double denom = d00 * d11 - d01 * d01;
    if (fabs(denom) < 1e-20) {
        *u = 1.0f;
        *v = 0.0f;
        *w = 0.0f;
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
    *u = (float)uu;
    *v = (float)vc;
    *w = (float)wc;
}

void oracle_contact_forces(const double* verts, uint32_t nv, const int32_t* faces, uint32_t nf, const float* cand,
                           uint32_t ncand, float band, int use_vel, float threshold, float* forces,
                           uint32_t counts[2]) {
    memset(forces, 0, sizeof(float) * 3 * (size_t)nv);
    uint32_t contacts = 0;
    for (uint32_t i = 0; i < ncand; ++i) {
        const float* r = cand + 8 * (size_t)i;
        const float px = r[0], py = r[1], pz = r[2];
        const float vx = r[4], vy = r[5], vz = r[6];
        const hit_t h = brute_closest(verts, faces, nf, mk(px, py, pz));
        if (h.tri < 0) continue;
        const float cx = (float)h.p.x, cy = (float)h.p.y, cz = (float)h.p.z;
        const float tx = cx - px, ty = cy - py, tz = cz - pz;
        const float dist = sqrtf(tx * tx + ty * ty + tz * tz);
        if (dist > band) continue;
        const int32_t* f = faces + 3 * (size_t)h.tri;
        const v3 A = ld(verts + 3 * (size_t)f[0]), B = ld(verts + 3 * (size_t)f[1]), C = ld(verts + 3 * (size_t)f[2]);
        float dx, dy, dz;
        if (dist > 1e-6f) {
            const float inv = 1.0f / dist;
            dx = tx * inv;
            dy = ty * inv;
            dz = tz * inv;
        } else {
            const v3 n = normalized(cross(sub(B, A), sub(C, A)));
            dx = (float)(-n.x);
            dy = (float)(-n.y);
            dz = (float)(-n.z);
        }
        float mag = 1.0f;
        if (use_vel) {
            const float vd = vx * dx + vy * dy + vz * dz;
            mag = 0.0f < vd ? vd : 0.0f;
            if (mag <= 0.0f) continue;
        }
        float u, v, w;
        barycentric(mk(cx, cy, cz), A, B, C, &u, &v, &w);
        const float fx = dx * mag, fy = dy * mag, fz = dz * mag;
        const float wt[3] = {u, v, w};
        for (int k = 0; k < 3; ++k) {
            float* q = forces + 3 * (size_t)f[k];
            q[0] = q[0] + fx * wt[k];
            q[1] = q[1] + fy * wt[k];
            q[2] = q[2] + fz * wt[k];
        }
        ++contacts;
    }
    uint32_t controls = 0;
    for (uint32_t v = 0; v < nv; ++v) {
        const float* q = forces + 3 * (size_t)v;
        const float m = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        if (m < 1e-5f) continue;
        if (m >= threshold) ++controls;
    }
    counts[0] = contacts;
    counts[1] = controls;
}
