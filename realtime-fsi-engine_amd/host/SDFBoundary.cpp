// SDFBoundary.cpp — host mirror of the reference's SDF boundary
// (src/fluid/SDFBoundary.cpp:278-610) over the device mesh of fsi_mesh_*.
#include "SDFBoundary.h"

#include <algorithm>
#include <cmath>
#include <iostream>

#include "fsi_pbf.h"

namespace {

std::vector<double> flat(const std::vector<std::array<double, 3>>& v) {
    std::vector<double> out(3 * v.size());
    for (size_t i = 0; i < v.size(); ++i)
        for (int k = 0; k < 3; ++k) out[3 * i + k] = v[i][k];
    return out;
}

void report(const char* where, int rc) { std::cerr << where << ": " << fsi_last_error() << " (status " << rc << ")\n"; }

}  // namespace

void SDFBoundary::buildFromMesh(const std::vector<std::array<double, 3>>& verts,
                                const std::vector<std::array<int, 3>>& faces, double resolution, bool verbose) {
    _mesh.reset();
    _distances.clear();
    _texture.clear();
    _dirtyValid = false;
    std::vector<double> v = flat(verts);
    std::vector<int32_t> f(3 * faces.size());
    for (size_t i = 0; i < faces.size(); ++i)
        for (int k = 0; k < 3; ++k) f[3 * i + k] = faces[i][k];
    fsi_mesh* m = nullptr;
    const int rc = fsi_mesh_create(v.data(), (uint32_t)verts.size(), f.data(), (uint32_t)faces.size(), resolution, &m);
    if (rc != FSI_OK) {
        report("SDFBoundary::buildFromMesh", rc);
        return;
    }
    _mesh = std::shared_ptr<fsi_mesh>(m, [](fsi_mesh* p) { fsi_mesh_destroy(p); });
    _nverts = verts.size();
    _buildVerts = verts;
    double o[3], cell = 0.0;
    int32_t d[3];
    fsi_mesh_grid_info(m, o, &cell, d);
    for (int k = 0; k < 3; ++k) {
        _origin.v[k] = o[k];
        _dimensions.v[k] = d[k];
    }
    _cellSize = cell;
    if (!readGrid()) return;
    if (verbose)
        std::cout << "SDF computed: " << d[0] << "x" << d[1] << "x" << d[2] << " (" << _distances.size()
                  << " voxels, " << faces.size() << " triangles, BVH)\n";
}

bool SDFBoundary::readGrid() {
    const int dx = _dimensions.x(), dy = _dimensions.y(), dz = _dimensions.z();
    const size_t total = (size_t)dx * dy * dz;
    _texture.resize(total);
    const int rc = fsi_mesh_read_sdf(_mesh.get(), _texture.data());
    if (rc != FSI_OK) {
        report("SDFBoundary: reading the grid", rc);
        _distances.clear();
        _texture.clear();
        return false;
    }
    _distances.resize(total);
    for (int z = 0; z < dz; ++z)
        for (int y = 0; y < dy; ++y)
            for (int x = 0; x < dx; ++x)
                _distances[index(x, y, z)] = _texture[(size_t)x + (size_t)dx * ((size_t)y + (size_t)dy * z)];
    return true;
}

void SDFBoundary::refitFromMesh(const std::vector<std::array<double, 3>>& verts) {
    if (!_mesh) return;
    if (verts.size() != _nverts) {
        std::cerr << "SDFBoundary::refitFromMesh: " << verts.size() << " vertices, the build had " << _nverts << "\n";
        return;
    }
    std::vector<double> v = flat(verts);
    const int rc = fsi_mesh_refit(_mesh.get(), v.data());
    if (rc != FSI_OK) {
        report("SDFBoundary::refitFromMesh", rc);
        return;
    }
    readGrid();
    _buildVerts = verts;
    _dirtyValid = false;  // a full pass supersedes any pending dirty region
}

bool SDFBoundary::refitDirty(const std::vector<std::array<double, 3>>& newVerts, double bandRadius,
                             double moveEps) {
    if (!_mesh) {
        // no baseline: the full path (a no-op here) with the whole grid dirty
        refitFromMesh(newVerts);
        for (int k = 0; k < 3; ++k) {
            _dirtyMin[k] = 0;
            _dirtyMax[k] = _dimensions[k] - 1;
        }
        _dirtyValid = true;
        return true;
    }
    if (newVerts.size() != _nverts) {
        std::cerr << "SDFBoundary::refitDirty: " << newVerts.size() << " vertices, the build had " << _nverts << "\n";
        return false;
    }
    // the voxels the device pass recomputes, for the debug comparison: the
    // band around the old and the new position of every moved vertex
    std::vector<size_t> marked;
    if (_debugVerifyDirty) {
        const int dims[3] = {_dimensions.x(), _dimensions.y(), _dimensions.z()};
        std::vector<uint8_t> seen(_distances.size(), 0);
        auto mark = [&](const std::array<double, 3>& w) {
            int lo[3], hi[3];
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::clamp((int)std::floor((w[k] - bandRadius - _origin[k]) / _cellSize), 0, dims[k] - 1);
                hi[k] = std::clamp((int)std::ceil((w[k] + bandRadius - _origin[k]) / _cellSize), 0, dims[k] - 1);
            }
            for (int x = lo[0]; x <= hi[0]; ++x)
                for (int y = lo[1]; y <= hi[1]; ++y)
                    for (int z = lo[2]; z <= hi[2]; ++z) {
                        const size_t idx = (size_t)index(x, y, z);
                        if (!seen[idx]) {
                            seen[idx] = 1;
                            marked.push_back(idx);
                        }
                    }
        };
        for (size_t i = 0; i < newVerts.size(); ++i) {
            double d2 = 0.0;
            for (int k = 0; k < 3; ++k) d2 += (newVerts[i][k] - _buildVerts[i][k]) * (newVerts[i][k] - _buildVerts[i][k]);
            if (d2 <= moveEps * moveEps) continue;
            mark(_buildVerts[i]);
            mark(newVerts[i]);
        }
    }
    std::vector<double> v = flat(newVerts);
    int32_t lo[3], hi[3];
    uint32_t count = 0;
    const int rc = fsi_mesh_refit_dirty(_mesh.get(), v.data(), bandRadius, moveEps, lo, hi, &count);
    if (rc != FSI_OK) {
        report("SDFBoundary::refitDirty", rc);
        return false;
    }
    _buildVerts = newVerts;
    _dirtyValid = count > 0;
    if (_dirtyValid) {
        for (int k = 0; k < 3; ++k) {
            _dirtyMin[k] = lo[k];
            _dirtyMax[k] = hi[k];
        }
        readGrid();
    }
    if (_debugVerifyDirty && _dirtyValid) {
        std::vector<float> dirtyVals(marked.size());
        for (size_t k = 0; k < marked.size(); ++k) dirtyVals[k] = _distances[marked[k]];
        if (fsi_mesh_refit(_mesh.get(), v.data()) == FSI_OK && readGrid()) {  // full, authoritative
            double maxDiff = 0.0;
            for (size_t k = 0; k < marked.size(); ++k)
                maxDiff = std::max(maxDiff, (double)std::fabs(dirtyVals[k] - _distances[marked[k]]));
            std::cout << "[SDF dirty verify] " << marked.size() << " voxels, max |diff| = " << maxDiff << "\n";
        }
    }
    return _dirtyValid;
}

bool SDFBoundary::dirtyRegion(int lo[3], int hi[3]) const {
    for (int k = 0; k < 3; ++k) {
        lo[k] = _dirtyMin[k];
        hi[k] = _dirtyMax[k];
    }
    return _dirtyValid;
}

SDFBoundary::ClosestSurface SDFBoundary::closestTriangle(const PVec3& p) const {
    const std::vector<ClosestSurface> out = closestTriangles(std::vector<PVec3>{p});
    return out.empty() ? ClosestSurface{} : out[0];
}

std::vector<SDFBoundary::ClosestSurface> SDFBoundary::closestTriangles(const std::vector<PVec3>& points) const {
    std::vector<ClosestSurface> out(points.size());
    if (!_mesh || points.empty()) return out;
    std::vector<float> pts(3 * points.size()), q(3 * points.size());
    std::vector<int32_t> face(points.size());
    for (size_t i = 0; i < points.size(); ++i) {
        pts[3 * i] = points[i].x;
        pts[3 * i + 1] = points[i].y;
        pts[3 * i + 2] = points[i].z;
    }
    const int rc = fsi_mesh_closest(_mesh.get(), pts.data(), (uint32_t)points.size(), face.data(), q.data());
    if (rc != FSI_OK) {
        report("SDFBoundary::closestTriangle", rc);
        return out;
    }
    for (size_t i = 0; i < points.size(); ++i) {
        out[i].face = face[i];
        out[i].point = make_pvec3(q[3 * i], q[3 * i + 1], q[3 * i + 2]);
    }
    return out;
}
