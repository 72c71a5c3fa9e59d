// SDFBoundary.cpp — host mirror of the reference's SDF boundary build
// (src/fluid/SDFBoundary.cpp:278-398) over fsi_sdf_build.
#include "SDFBoundary.h"

#include <iostream>

#include "fsi_pbf.h"

void SDFBoundary::buildFromMesh(const std::vector<std::array<double, 3>>& verts,
                                const std::vector<std::array<int, 3>>& faces, double resolution, bool verbose) {
    std::vector<double> v(3 * verts.size());
    std::vector<int32_t> f(3 * faces.size());
    for (size_t i = 0; i < verts.size(); ++i)
        for (int k = 0; k < 3; ++k) v[3 * i + k] = verts[i][k];
    for (size_t i = 0; i < faces.size(); ++i)
        for (int k = 0; k < 3; ++k) f[3 * i + k] = faces[i][k];
    double o[3];
    int32_t d[3];
    int rc = fsi_sdf_build(v.data(), (uint32_t)verts.size(), f.data(), (uint32_t)faces.size(), resolution, o, d,
                           nullptr);
    if (rc != FSI_OK) {
        std::cerr << "SDFBoundary::buildFromMesh: invalid mesh or resolution\n";
        _distances.clear();
        _texture.clear();
        return;
    }
    const size_t total = (size_t)d[0] * d[1] * d[2];
    _texture.assign(total, 0.0f);
    rc = fsi_sdf_build(v.data(), (uint32_t)verts.size(), f.data(), (uint32_t)faces.size(), resolution, o, d,
                       _texture.data());
    if (rc != FSI_OK) {
        std::cerr << "SDFBoundary::buildFromMesh: build failed\n";
        _distances.clear();
        _texture.clear();
        return;
    }
    for (int k = 0; k < 3; ++k) {
        _origin.v[k] = o[k];
        _dimensions.v[k] = d[k];
    }
    _cellSize = resolution;
    _distances.assign(total, 0.0f);
    for (int z = 0; z < d[2]; ++z)
        for (int y = 0; y < d[1]; ++y)
            for (int x = 0; x < d[0]; ++x)
                _distances[index(x, y, z)] = _texture[(size_t)x + (size_t)d[0] * ((size_t)y + (size_t)d[1] * z)];
    if (verbose)
        std::cout << "SDF computed: " << d[0] << "x" << d[1] << "x" << d[2] << " (" << total << " voxels, "
                  << faces.size() << " triangles, BVH)\n";
}
