/*
 * mesh_oracle.h — CPU oracle of the FSI coupling (TEST INFRASTRUCTURE ONLY):
 * brute-force signed distance, dirty-voxel marking and contact forces.
 * mesh_oracle.c names the reference lines each function follows.
 *
 * Parity pins (tests/test_mesh_oracle.py): the distance to an icosphere
 * against the analytic sphere within the chord error, the three feature
 * regions of a single triangle in closed form, the sign inside / outside a
 * closed mesh, the dirty box of a single moved vertex by hand, and the
 * contact forces of particles above a plane triangle in closed form (push
 * direction, velocity magnitude, barycentric split, sum of forces).
 */
#ifndef MESH_ORACLE_H
#define MESH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Signed distance of every voxel (x fastest) to the triangles at `verts`,
 * with the edge / vertex pseudo-normals of a build at `verts_build` and the
 * face normals of `verts` (a refit keeps the build's edge and vertex
 * normals).  verts_build == verts for a fresh build. */
void oracle_mesh_sdf(const double* verts, const double* verts_build, uint32_t nv, const int32_t* faces, uint32_t nf,
                     const double origin[3], double cell, const int32_t dims[3], float* grid);

/* Voxels (x fastest, mask[v] = 1) a dirty refit from old_verts to new_verts
 * recomputes; returns their number, box = min xyz, max xyz. */
uint32_t oracle_dirty_mask(const double* old_verts, const double* new_verts, uint32_t nv, double band,
                           double move_eps, const double origin[3], double cell, const int32_t dims[3],
                           uint8_t* mask, int32_t box[6]);

/* Nearest triangle of each point (3 floats): its index (lowest among
 * equidistant ones), the closest point as float and its squared distance. */
void oracle_closest_triangles(const double* verts, const int32_t* faces, uint32_t nf, const float* pts, uint32_t n,
                              int32_t* face_out, float* point_out, double* d2_out);

/* Per-vertex forces (3 floats each) of the candidates (8 floats each:
 * pos xyzw, vel xyzw) in candidate order; counts = contacts, control points. */
void oracle_contact_forces(const double* verts, uint32_t nv, const int32_t* faces, uint32_t nf, const float* cand,
                           uint32_t ncand, float band, int use_vel, float threshold, float* forces,
                           uint32_t counts[2]);

#ifdef __cplusplus
}
#endif

#endif
