// LaplacianDeformation.h — the reference's LaplacianDeformation interface
// (include/deformation/LaplacianDeformation.h) over the fsi_deform_* C-ABI.
//
// The reference deforms an OpenMesh PolyMesh in place; this mirror holds the
// vertex positions itself (points()), built from the same vertex / triangle
// lists the reference's buildDeformMeshFromData (src/main.cpp:731-739) feeds
// to OpenMesh.  Method names and argument meanings are the reference's;
// failures print and leave the state unchanged, as the reference does.
#pragma once

#include <array>
#include <vector>

#include "fsi_pbf.h"

class LaplacianDeformation {
public:
    using Vec3 = std::array<double, 3>;
    LaplacianDeformation(const std::vector<Vec3>& verts, const std::vector<std::array<int, 3>>& faces);
    ~LaplacianDeformation();
    LaplacianDeformation(const LaplacianDeformation&) = delete;
    LaplacianDeformation& operator=(const LaplacianDeformation&) = delete;

    // The weights, L and delta are built at construction (the C-ABI's
    // create); initialize() is kept for the reference's call sequence.
    void initialize() {}

    void addAnchor(int vertexIndex);
    void setControlPoints(const std::vector<int>& indices, const std::vector<Vec3>& targets);
    void clearControlPoints();
    const std::vector<int>& getControlIndices() const { return _controlIndices; }
    const std::vector<int>& getAnchorIndices() const { return _anchorIndices; }

    void precomputeSystem();
    void solve();

    void setUseLocal(bool on);
    bool isUsingLocal() const { return _useLocal; }
    void setLocalKRing(int k);
    int getLocalPatchSize() const;
    void setUseIncremental(bool on);
    bool isUsingIncremental() const { return _useIncremental; }
    double verifyIncremental();

    // deformStep (src/main.cpp:862-944) for per-vertex contact forces
    // (3 floats each); returns the number of handles this frame.
    int deformStep(const std::vector<float>& forces, float dt, const fsi_deform_params& params);

    // The deformed positions (the reference reads them from its mesh).
    const std::vector<Vec3>& points();

    fsi_deform* handle() const { return _h; }

private:
    bool ok(int rc, const char* what);
    void applyMode();

    fsi_deform* _h = nullptr;
    std::vector<Vec3> _points;
    std::vector<int> _anchorIndices, _controlIndices;
    bool _useLocal = true, _useIncremental = false;
    int _localKRing = 3;
};
