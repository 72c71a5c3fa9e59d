// LaplacianDeformation.cpp — see LaplacianDeformation.h.
#include "LaplacianDeformation.h"

#include <iostream>

LaplacianDeformation::LaplacianDeformation(const std::vector<Vec3>& verts, const std::vector<std::array<int, 3>>& faces)
    : _points(verts) {
    std::vector<double> v;
    v.reserve(3 * verts.size());
    for (const Vec3& p : verts) v.insert(v.end(), p.begin(), p.end());
    std::vector<int32_t> f;
    f.reserve(3 * faces.size());
    for (const auto& t : faces) f.insert(f.end(), t.begin(), t.end());
    ok(fsi_deform_create(v.data(), (uint32_t)verts.size(), f.data(), (uint32_t)faces.size(), &_h), "LaplacianDeformation");
}

LaplacianDeformation::~LaplacianDeformation() { fsi_deform_destroy(_h); }

bool LaplacianDeformation::ok(int rc, const char* what) {
    if (rc == FSI_OK) return true;
    std::cerr << "LaplacianDeformation::" << what << ": " << fsi_last_error() << " (status " << rc << ")\n";
    return false;
}

void LaplacianDeformation::applyMode() {
    if (_h) ok(fsi_deform_set_mode(_h, _useLocal ? 1 : 0, _localKRing, _useIncremental ? 1 : 0), "setMode");
}

void LaplacianDeformation::addAnchor(int vertexIndex) {
    if (!_h) return;
    for (int a : _anchorIndices)
        if (a == vertexIndex) return;
    if (ok(fsi_deform_add_anchor(_h, vertexIndex), "addAnchor")) _anchorIndices.push_back(vertexIndex);
}

void LaplacianDeformation::setControlPoints(const std::vector<int>& indices, const std::vector<Vec3>& targets) {
    if (!_h || indices.size() != targets.size()) return;
    std::vector<double> t;
    t.reserve(3 * targets.size());
    for (const Vec3& p : targets) t.insert(t.end(), p.begin(), p.end());
    if (ok(fsi_deform_set_control_points(_h, indices.data(), t.data(), (uint32_t)indices.size()), "setControlPoints"))
        _controlIndices = indices;
}

void LaplacianDeformation::clearControlPoints() {
    if (_h && ok(fsi_deform_set_control_points(_h, nullptr, nullptr, 0), "clearControlPoints")) _controlIndices.clear();
}

void LaplacianDeformation::precomputeSystem() {
    if (_h) ok(fsi_deform_precompute(_h), "precomputeSystem");
}

void LaplacianDeformation::solve() {
    if (_h) ok(fsi_deform_solve(_h), "solve");
}

void LaplacianDeformation::setUseLocal(bool on) {
    _useLocal = on;
    applyMode();
}

void LaplacianDeformation::setLocalKRing(int k) {
    _localKRing = k;
    applyMode();
}

void LaplacianDeformation::setUseIncremental(bool on) {
    _useIncremental = on;
    applyMode();
}

int LaplacianDeformation::getLocalPatchSize() const {
    uint32_t info[4] = {0, 0, 0, 0};
    if (!_h || fsi_deform_info(_h, info) != FSI_OK) return 0;
    return (int)info[0];
}

double LaplacianDeformation::verifyIncremental() {
    double d = -1.0;
    if (_h) ok(fsi_deform_verify_incremental(_h, &d), "verifyIncremental");
    return d;
}

int LaplacianDeformation::deformStep(const std::vector<float>& forces, float dt, const fsi_deform_params& params) {
    uint32_t handles = 0;
    int solved = 0;
    if (!_h || forces.size() != 3 * _points.size()) return 0;
    ok(fsi_deform_step(_h, forces.data(), dt, &params, &handles, &solved), "deformStep");
    return (int)handles;
}

const std::vector<LaplacianDeformation::Vec3>& LaplacianDeformation::points() {
    if (_h) {
        std::vector<double> v(3 * _points.size());
        if (ok(fsi_deform_read_vertices(_h, v.data()), "points"))
            for (size_t i = 0; i < _points.size(); ++i) _points[i] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
    }
    return _points;
}
