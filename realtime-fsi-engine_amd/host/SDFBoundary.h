// SDFBoundary.h — host-side signed-distance boundary with the reference's
// interface (include/fluid/SDFBoundary.h:38-167), built on the C-ABI of
// include/fsi_pbf.h.  The triangles, their BVH and the distance grid live on
// the device (fsi_mesh_*): buildFromMesh builds the grid there (the same
// FP64 arithmetic as the reference's host build, the same grid), the refits
// and closestTriangle run there, and the host keeps a copy of the grid for
// distances().  "Uploading" hands the device grid to the solver (a
// device-to-device copy before its next step) instead of filling a GL 3D
// texture.
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <vector>

#include "Types.h"

struct fsi_mesh;

// Small stand-ins for the Eigen vectors the reference returns, with the same
// accessors the reference's callers use (o.x(), dim.x()).
struct SdfVec3d {
    double v[3] = {0, 0, 0};
    double x() const { return v[0]; }
    double y() const { return v[1]; }
    double z() const { return v[2]; }
    double operator[](int i) const { return v[i]; }
};
struct SdfVec3i {
    int v[3] = {0, 0, 0};
    int x() const { return v[0]; }
    int y() const { return v[1]; }
    int z() const { return v[2]; }
    int operator[](int i) const { return v[i]; }
};

class SDFBoundary {
public:
    // SDFBoundary::buildFromMesh (src/fluid/SDFBoundary.cpp:278-398): a new
    // device mesh (triangles, pseudo-normals, BVH) and its grid over
    // bbox - 5 cells, dims ceil(span / cell).  On failure the boundary is
    // left invalid and the reason printed.
    void buildFromMesh(const std::vector<std::array<double, 3>>& verts, const std::vector<std::array<int, 3>>& faces,
                       double resolution, bool verbose = true);

    // SDFBoundary::refitFromMesh (:454-467): moved vertices of the same
    // topology; corners, face normals and BVH boxes refreshed, every voxel
    // recomputed, the edge and vertex pseudo-normals kept from the build.
    // No-op on a boundary that was never built.  Call uploadToGPU() after.
    void refitFromMesh(const std::vector<std::array<double, 3>>& verts);

    // SDFBoundary::refitDirty (:469-548): only voxels within bandRadius of a
    // vertex that moved more than moveEps since the last build or refit
    // (around its old and its new position) are recomputed; their box is
    // recorded for uploadDirtyRegion().  Returns whether anything was dirty.
    // Without a built mesh, as the reference, it takes the full path and
    // reports the whole (empty) grid dirty.  A vertex array whose size
    // differs from the build's is refused (printed, returns false): the
    // device refit reads exactly the build's vertex count, where the
    // reference would read past a shorter array.
    bool refitDirty(const std::vector<std::array<double, 3>>& newVerts, double bandRadius, double moveEps);

    // After each refitDirty, recompute the full grid and print the largest
    // difference over the recomputed voxels (include/fluid/SDFBoundary.h:86-89).
    void setDebugVerifyDirty(bool on) { _debugVerifyDirty = on; }

    // SDFBoundary::uploadToGPU (:580-598): marks the whole grid for the
    // solvers that hold this boundary (they copy it before their next step).
    void uploadToGPU() {
        if (valid()) ++_version;
    }
    // SDFBoundary::uploadDirtyRegion (:550-578): the same after a refitDirty
    // that marked something; no-op otherwise.  The solver's copy is
    // device-to-device over the whole grid, which at these grid sizes costs
    // no more than a sub-box copy.
    void uploadDirtyRegion() {
        if (_dirtyValid) ++_version;
    }

    bool valid() const { return !_distances.empty(); }
    const SdfVec3d& origin() const { return _origin; }
    double cellSize() const { return _cellSize; }
    const SdfVec3i& dimensions() const { return _dimensions; }

    // CPU grid in the reference's order: index = x*Ny*Nz + y*Nz + z
    // (include/fluid/SDFBoundary.h:110-113).
    const std::vector<float>& distances() const { return _distances; }
    int index(int x, int y, int z) const { return x * _dimensions.y() * _dimensions.z() + y * _dimensions.z() + z; }

    // include/fluid/SDFBoundary.h:117-127.
    struct ClosestSurface {
        int face = -1;              // index into the faces passed to buildFromMesh (-1 if none)
        PVec3 point{0.f, 0.f, 0.f}; // closest point on the surface
    };
    // SDFBoundary::closestTriangle (:600-610): nearest triangle to p on the
    // current (refit) geometry; face = -1 if no mesh is built.  One device
    // round trip per call: callers with many points use closestTriangles.
    ClosestSurface closestTriangle(const PVec3& p) const;
    // closestTriangle of every point in one device call.
    std::vector<ClosestSurface> closestTriangles(const std::vector<PVec3>& points) const;

    // The grid x-fastest (the device / texture order), as last read back.
    const std::vector<float>& textureOrder() const { return _texture; }
    // The device mesh (NULL before a successful build) and the upload count
    // the solver compares against.
    fsi_mesh* deviceMesh() const { return _mesh.get(); }
    uint64_t version() const { return _version; }
    // Dirty box of the last refitDirty, voxel coordinates inclusive; false
    // when it marked nothing.
    bool dirtyRegion(int lo[3], int hi[3]) const;

private:
    bool readGrid();  // device grid -> _texture and _distances

    SdfVec3d _origin;
    double _cellSize = 0.0;
    SdfVec3i _dimensions;
    std::vector<float> _distances;
    std::vector<float> _texture;
    std::shared_ptr<fsi_mesh> _mesh;
    size_t _nverts = 0;
    std::vector<std::array<double, 3>> _buildVerts;  // vertices of the last build or refit
    int _dirtyMin[3] = {0, 0, 0}, _dirtyMax[3] = {-1, -1, -1};
    bool _dirtyValid = false;
    bool _debugVerifyDirty = false;
    uint64_t _version = 0;
};
