// PBFluids.h — the reference's fluid solver interface (include/fluid/PBFluids.h:88-268,
// include/core/config.h:9-63, include/fluid/Particle.h:8-12,
// include/core/Types.h:43-61) re-implemented over the gfx950 C-ABI of
// include/fsi_pbf.h.  Same class, method names, argument meanings and error
// behaviour (no exceptions: failures are printed and the call becomes a
// no-op), so the reference's main loop compiles against it unchanged except
// for the GL-only members noted below.
#pragma once

#include <cstdint>
#include <vector>

#include "SDFBoundary.h"
#include "Types.h"

struct fsi_pbf;


// 48 bytes, std430-padded (include/fluid/Particle.h:8-12).
struct Particle {
    PVec3 pos;
    F32 _pad0 = 0.f;
    PVec3 predPos;
    F32 _pad1 = 0.f;
    PVec3 vel;
    F32 _pad2 = 0.f;
};
static_assert(sizeof(Particle) == 48, "Particle must stay 48 bytes");

// include/fluid/PBFluids.h:67-70
struct ContactCandidate {
    PVec4 pos;
    PVec4 vel;
};
static_assert(sizeof(ContactCandidate) == 32, "ContactCandidate must stay 32 bytes");

// include/core/config.h:9-63, same field order and defaults (scene presets
// aggregate-initialise it positionally).
struct FluidConfig {
    F32 dt = 1.0f / 60.0f;
    PVec3 gravity = {0.0f, -3.81f, 0.0f};
    F32 h = 0.15f;
    F32 rho0 = 1200.0f;
    U32 solverIterations = 3;
    U32 substepIterations = 2;
    F32 eps = 0.0001f;
    U32 particleCount = 1 << 12;
    PVec3 spawnMin = {2.f, 0.2f, 2.f};
    PVec3 spawnMax = {5.f, 5.f, 5.f};
    bool spawnRandom = false;
    F32 spacing = 0.1f;
    PVec3 initialVelocity = {0.0f, 0.0f, 0.0f};
    PVec3 boundsMin = {0.0f, 0.0f, 0.0f};
    PVec3 boundsMax = {6.0f, 6.0f, 6.0f};
    F32 boundDamping = 0.5f;
    U32 hashSize = 1 << 17;
    bool enableSCorr = true;
    F32 kCorr = 0.001f;
    F32 nCorr = 4.0f;
    F32 deltaQ = 0.3f;
    F32 cohesionStrength = 0.001f;
    bool enableViscosity = true;
    F32 viscosity = 0.1f;
    F32 interactionRadius = 1.5f;
    F32 interactionStrength = 25.0f;
    bool enableVorticity = true;
    F32 vorticityEpsilon = 0.05f;
    bool enableAPBF = false;
    U32 minLOD = 2;
    U32 maxLOD = 7;
    F32 lodMaxDist = 15.0f;
};

class PBFluids {
public:
    PBFluids() = default;
    explicit PBFluids(const FluidConfig& p);
    ~PBFluids();
    PBFluids(const PBFluids&) = delete;
    PBFluids& operator=(const PBFluids&) = delete;

    void setParams(const FluidConfig& p);
    const FluidConfig& params() const { return _params; }

    void setCameraPos(const PVec3& pos);

    void setParticles(const std::vector<Particle>& particles);
    // Host copy as given to setParticles; like the reference it is not
    // refreshed by step() (use readbackParticles).
    std::vector<Particle>& particles() { return _particles; }
    const std::vector<Particle>& particles() const { return _particles; }

    void setCollisionPadding(F32 p);
    void setBounds(const PVec3& minBound, const PVec3& maxBound, F32 damping = 0.0f);

    // Non-owning; the caller keeps the boundary alive and calls uploadToGPU()
    // after rebuilding it.  nullptr detaches.
    void setSDFBoundary(const SDFBoundary* sdf) { _sdf = sdf; }
    void setSDFPadding(F32 padding);
    void setSDFContainment(bool inside);

    void setInteraction(bool active, const PVec3& pos);

    void step();

    // The reference binds its particle SSBO (binding 0) for the render
    // shaders.  Without a buffer this finishes pending work; with one (device
    // memory of particles().size() * 48 bytes, e.g. a GL/Vulkan buffer mapped
    // for HIP) it writes the 48-byte records there on the device, and the
    // renderer's stream, if given, waits for them.
    void bindParticlesForRendering() const;
    void bindParticlesForRendering(void* deviceBuffer, void* stream = nullptr) const;

    void readbackParticles(std::vector<Particle>& out);
    void filterContactCandidates(float contactBand);
    void readbackContactCandidates(std::vector<ContactCandidate>& out);

    // Access to the underlying C-ABI handle (timing, taps).
    fsi_pbf* handle() const { return _h; }
    // Status of the last C-ABI call (0 = ok).
    int lastStatus() const { return _status; }

private:
    bool ok(int rc, const char* what);
    void syncSDF();

    FluidConfig _params{};
    std::vector<Particle> _particles;
    fsi_pbf* _h = nullptr;
    int _status = 0;
    const SDFBoundary* _sdf = nullptr;
    const SDFBoundary* _uploadedSdf = nullptr;
    uint64_t _uploadedVersion = 0;
};
