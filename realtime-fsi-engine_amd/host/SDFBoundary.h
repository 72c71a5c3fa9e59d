// SDFBoundary.h — host-side signed-distance boundary with the reference's
// interface (include/fluid/SDFBoundary.h:38-167), built on the C-ABI of
// include/fsi_pbf.h.  The grid is built in FP64 on the host
// (fsi_sdf_build) and handed to the solver as a device buffer instead of a
// GL 3D texture.
#pragma once

#include <array>
#include <cstdint>
#include <vector>

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
    // SDFBoundary::buildFromMesh (src/fluid/SDFBoundary.cpp:278-398).
    void buildFromMesh(const std::vector<std::array<double, 3>>& verts, const std::vector<std::array<int, 3>>& faces,
                       double resolution, bool verbose = true);

    // SDFBoundary::uploadToGPU (src/fluid/SDFBoundary.cpp:580-598): marks the
    // grid for upload; a solver that holds this boundary copies it to the
    // device before its next step.
    void uploadToGPU() { ++_version; }

    bool valid() const { return !_distances.empty(); }
    const SdfVec3d& origin() const { return _origin; }
    double cellSize() const { return _cellSize; }
    const SdfVec3i& dimensions() const { return _dimensions; }

    // CPU grid in the reference's order: index = x*Ny*Nz + y*Nz + z
    // (include/fluid/SDFBoundary.h:110-113).
    const std::vector<float>& distances() const { return _distances; }
    int index(int x, int y, int z) const { return x * _dimensions.y() * _dimensions.z() + y * _dimensions.z() + z; }

    // The same grid x-fastest (the device / texture order).
    const std::vector<float>& textureOrder() const { return _texture; }
    uint64_t version() const { return _version; }

private:
    SdfVec3d _origin;
    double _cellSize = 0.0;
    SdfVec3i _dimensions;
    std::vector<float> _distances;
    std::vector<float> _texture;
    uint64_t _version = 0;
};
