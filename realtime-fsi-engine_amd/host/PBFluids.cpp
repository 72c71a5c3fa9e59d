// PBFluids.cpp — the reference's PBFluids class (src/fluid/PBFluids.cpp) over
// the C-ABI of include/fsi_pbf.h.
#include "PBFluids.h"

#include <iostream>

#include "fsi_pbf.h"

namespace {

fsi_fluid_config to_c(const FluidConfig& p) {
    fsi_fluid_config c{};
    c.dt = p.dt;
    c.gravity[0] = p.gravity.x;
    c.gravity[1] = p.gravity.y;
    c.gravity[2] = p.gravity.z;
    c.h = p.h;
    c.rho0 = p.rho0;
    c.solver_iterations = p.solverIterations;
    c.substep_iterations = p.substepIterations;
    c.eps = p.eps;
    c.particle_count = p.particleCount;
    c.spawn_min[0] = p.spawnMin.x;
    c.spawn_min[1] = p.spawnMin.y;
    c.spawn_min[2] = p.spawnMin.z;
    c.spawn_max[0] = p.spawnMax.x;
    c.spawn_max[1] = p.spawnMax.y;
    c.spawn_max[2] = p.spawnMax.z;
    c.spawn_random = p.spawnRandom ? 1u : 0u;
    c.spacing = p.spacing;
    c.initial_velocity[0] = p.initialVelocity.x;
    c.initial_velocity[1] = p.initialVelocity.y;
    c.initial_velocity[2] = p.initialVelocity.z;
    c.bounds_min[0] = p.boundsMin.x;
    c.bounds_min[1] = p.boundsMin.y;
    c.bounds_min[2] = p.boundsMin.z;
    c.bounds_max[0] = p.boundsMax.x;
    c.bounds_max[1] = p.boundsMax.y;
    c.bounds_max[2] = p.boundsMax.z;
    c.bound_damping = p.boundDamping;
    c.hash_size = p.hashSize;
    c.enable_scorr = p.enableSCorr ? 1u : 0u;
    c.k_corr = p.kCorr;
    c.n_corr = p.nCorr;
    c.delta_q = p.deltaQ;
    c.cohesion_strength = p.cohesionStrength;
    c.enable_viscosity = p.enableViscosity ? 1u : 0u;
    c.viscosity = p.viscosity;
    c.interaction_radius = p.interactionRadius;
    c.interaction_strength = p.interactionStrength;
    c.enable_vorticity = p.enableVorticity ? 1u : 0u;
    c.vorticity_epsilon = p.vorticityEpsilon;
    c.enable_apbf = p.enableAPBF ? 1u : 0u;
    c.min_lod = p.minLOD;
    c.max_lod = p.maxLOD;
    c.lod_max_dist = p.lodMaxDist;
    return c;
}

}  // namespace

bool PBFluids::ok(int rc, const char* what) {
    _status = rc;
    if (rc == FSI_OK) return true;
    std::cerr << "PBFluids::" << what << ": " << fsi_last_error() << " (status " << rc << ")\n";
    return false;
}

PBFluids::PBFluids(const FluidConfig& p) : _params(p) {
    const fsi_fluid_config c = to_c(p);
    if (!ok(fsi_pbf_create(&c, &_h), "PBFluids")) _h = nullptr;
}

PBFluids::~PBFluids() {
    if (_h) fsi_pbf_destroy(_h);
}

// src/fluid/PBFluids.cpp:39-109
void PBFluids::setParams(const FluidConfig& p) {
    _params = p;
    if (!_h) return;
    const fsi_fluid_config c = to_c(p);
    ok(fsi_pbf_set_params(_h, &c), "setParams");
    _params.particleCount = (U32)_particles.size();
}

void PBFluids::setCameraPos(const PVec3& pos) {
    if (!_h) return;
    const float v[3] = {pos.x, pos.y, pos.z};
    ok(fsi_pbf_set_camera(_h, v), "setCameraPos");
}

// src/fluid/PBFluids.cpp:111-152
void PBFluids::setParticles(const std::vector<Particle>& particles) {
    _particles = particles;
    _params.particleCount = (U32)_particles.size();
    if (!_h) return;
    ok(fsi_pbf_set_particles(_h, reinterpret_cast<const fsi_particle*>(_particles.data()), (uint32_t)_particles.size()),
       "setParticles");
}

void PBFluids::setCollisionPadding(F32 p) {
    if (!_h) return;
    ok(fsi_pbf_set_collision_padding(_h, p), "setCollisionPadding");
}

// src/fluid/PBFluids.cpp:299-306
void PBFluids::setBounds(const PVec3& minBound, const PVec3& maxBound, F32 damping) {
    if (!_h) return;
    const float mn[3] = {minBound.x, minBound.y, minBound.z};
    const float mx[3] = {maxBound.x, maxBound.y, maxBound.z};
    ok(fsi_pbf_set_bounds(_h, mn, mx, damping), "setBounds");
}

void PBFluids::setSDFPadding(F32 padding) {
    if (!_h) return;
    ok(fsi_pbf_set_sdf_padding(_h, padding), "setSDFPadding");
}

void PBFluids::setSDFContainment(bool inside) {
    if (!_h) return;
    ok(fsi_pbf_set_sdf_containment(_h, inside ? 1 : 0), "setSDFContainment");
}

// src/fluid/PBFluids.cpp:312-317
void PBFluids::setInteraction(bool active, const PVec3& pos) {
    if (!_h) return;
    const float v[3] = {pos.x, pos.y, pos.z};
    ok(fsi_pbf_set_interaction(_h, active ? 1 : 0, v), "setInteraction");
}

void PBFluids::syncSDF() {
    if (_sdf == _uploadedSdf && (!_sdf || _sdf->version() == _uploadedVersion)) return;
    if (!_sdf || !_sdf->valid() || !_sdf->deviceMesh()) {
        ok(fsi_pbf_set_sdf(_h, nullptr, nullptr, nullptr, 0.0f), "setSDFBoundary");
    } else {
        // the boundary's device grid, copied device to device
        ok(fsi_pbf_set_sdf_from_mesh(_h, _sdf->deviceMesh()), "setSDFBoundary");
    }
    _uploadedSdf = _sdf;
    _uploadedVersion = _sdf ? _sdf->version() : 0;
}

// src/fluid/PBFluids.cpp:323-517
void PBFluids::step() {
    if (_particles.empty() || !_h) return;
    syncSDF();
    ok(fsi_pbf_step(_h), "step");
}

void PBFluids::bindParticlesForRendering() const {
    if (_h) fsi_pbf_sync(_h);
}

void PBFluids::bindParticlesForRendering(void* deviceBuffer, void* stream) const {
    if (!_h) return;
    const int rc = fsi_pbf_write_render_buffer(_h, deviceBuffer, stream);
    if (rc)
        std::cerr << "PBFluids::bindParticlesForRendering: " << fsi_last_error() << " (status " << rc << ")\n";
}

// src/fluid/PBFluids.cpp:154-191
void PBFluids::readbackParticles(std::vector<Particle>& out) {
    out.resize(_particles.size());
    if (!_h || out.empty()) return;
    uint32_t n = 0;
    ok(fsi_pbf_read_particles(_h, reinterpret_cast<fsi_particle*>(out.data()), (uint32_t)out.size(), &n),
       "readbackParticles");
}

// src/fluid/PBFluids.cpp:193-237
void PBFluids::filterContactCandidates(float contactBand) {
    if (!_h || !_sdf || !_sdf->valid() || contactBand <= 0.0f) return;
    syncSDF();
    ok(fsi_pbf_filter_contacts(_h, contactBand), "filterContactCandidates");
}

// src/fluid/PBFluids.cpp:239-284
void PBFluids::readbackContactCandidates(std::vector<ContactCandidate>& out) {
    out.clear();
    if (!_h) return;
    uint32_t count = 0;
    if (!ok(fsi_pbf_read_contacts(_h, nullptr, 0, &count), "readbackContactCandidates") || count == 0) return;
    out.resize(count);
    ok(fsi_pbf_read_contacts(_h, reinterpret_cast<fsi_contact_candidate*>(out.data()), count, &count),
       "readbackContactCandidates");
}
