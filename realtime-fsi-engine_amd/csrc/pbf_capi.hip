// pbf_capi.hip — single-scene solver handle and its C-ABI (include/fsi_pbf.h).
//
// The handle replaces class PBFluids (include/fluid/PBFluids.h:88-268): it owns
// the device buffers that were SSBOs and one HIP stream in place of the GL
// command stream; the frame sequence of PBFluids::step
// (src/fluid/PBFluids.cpp:323-517) lives in pbf_core.hip.  Stream order
// replaces the glMemoryBarrier after every dispatch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/fsi_pbf.h"
#include "pbf_core.h"
#include "pbf_mesh.h"

using namespace fsi;

struct fsi_pbf {
    Core core;
    fsi_fluid_config cfg{};
    fsi_pbf_constants k{};
    float collision_padding = 0.0f;
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0}, damping = 0.0f;
    // SDF boundary
    float* sdf_grid = nullptr;
    float2* sdf_brange = nullptr;  // block ranges of sdf_grid (Domain::sdf_brange)
    int32_t sdf_dims[3] = {0, 0, 0};
    float sdf_origin[3] = {0, 0, 0};
    float sdf_cell = 0.0f;
    bool sdf_valid = false;
    float sdf_padding = 0.0f;
    bool sdf_containment = false;
    // interaction / camera
    bool interacting = false;
    float ipos[3] = {0, 0, 0};
    float cam[3] = {0, 0, 0};
    std::vector<float> host_pred;  // predPos.xyz + pad1 per particle (never touched by the step)
    float4* pred = nullptr;        // the same on the device, for the render buffer
    // contacts
    float4* contacts = nullptr;
    uint32_t* contact_count = nullptr;
    bool contacts_valid = false;
};

namespace {

// PBFluids::setBounds (src/fluid/PBFluids.cpp:299-306)
void apply_bounds(fsi_pbf* h, const float mn[3], const float mx[3], float damping) {
    for (int a = 0; a < 3; ++a) {
        h->bmin[a] = mn[a] + h->collision_padding;
        h->bmax[a] = mx[a] - h->collision_padding;
    }
    h->damping = std::clamp(damping, 0.0f, 1.0f);
}

void repack(fsi_pbf* h) { pack_constants(h->cfg, h->bmin, h->bmax, h->damping, h->core.d.n, h->k); }

StepConst step_const(const fsi_pbf* h) {
    const fsi_pbf_constants& u = h->k;
    StepConst c = step_const_common(h->cfg, u, h->core.d.n, h->interacting, h->ipos, h->cam);
    Domain& d = c.dom;
    for (int a = 0; a < 3; ++a) {
        d.bmin[a] = u.bounds_min[a];
        d.bmax[a] = u.bounds_max[a];
        d.dims[a] = h->sdf_dims[a];
        d.sdf_origin[a] = h->sdf_origin[a];
    }
    d.hash_mask = u.hash_mask;
    d.hash_base = 0;
    d.enable_sdf = h->sdf_valid ? 1u : 0u;
    d.sdf_mode = h->sdf_containment ? 1u : 0u;
    d.sdf_cell = h->sdf_cell;
    d.sdf_pad = h->sdf_padding;
    d.sdf_grid = h->sdf_grid;
    d.sdf_brange = h->sdf_brange;
    for (int a = 0; a < 3; ++a) d.sdf_bdims[a] = sdf_blocks(h->sdf_dims[a]);
    set_binning(d, u.bounds_min, u.bounds_max, u.h, 7, 0, h->core.d.variant == 2);
    return c;
}

FramePlan frame_plan(const fsi_pbf* h) {
    FramePlan p;
    p.substeps = h->cfg.substep_iterations;
    p.max_iter = h->cfg.enable_apbf ? h->cfg.max_lod : h->cfg.solver_iterations;
    p.key_bits = log2u(h->cfg.hash_size);
    p.vorticity = h->cfg.enable_vorticity != 0;
    return p;
}

int bind(fsi_pbf* h) {
    if (!h) return fail(FSI_ERR_INVALID_ARG, "handle is NULL");
    FSI_HIP_TRY(hipSetDevice(h->core.device));
    return FSI_OK;
}

// Reorder a sorted-slot array of `width` floats per particle into id order.
void unpermute(const std::vector<uint32_t>& perm, const float* src, float* dst, int width) {
    for (size_t k = 0; k < perm.size(); ++k)
        std::memcpy(dst + (size_t)perm[k] * width, src + k * width, sizeof(float) * width);
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================

extern "C" {

const char* fsi_last_error(void) { return last_error(); }

int fsi_abi_version(void) { return FSI_PBF_ABI_VERSION; }

int fsi_fluid_config_default(fsi_fluid_config* o) {
    if (!o) return fail(FSI_ERR_INVALID_ARG, "out is NULL");
    std::memset(o, 0, sizeof(*o));
    // include/core/config.h:9-63
    o->dt = 1.0f / 60.0f;
    o->gravity[0] = 0.0f;
    o->gravity[1] = -3.81f;
    o->gravity[2] = 0.0f;
    o->h = 0.15f;
    o->rho0 = 1200.0f;
    o->solver_iterations = 3;
    o->substep_iterations = 2;
    o->eps = 0.0001f;
    o->particle_count = 1u << 12;
    o->spawn_min[0] = 2.f;
    o->spawn_min[1] = 0.2f;
    o->spawn_min[2] = 2.f;
    o->spawn_max[0] = 5.f;
    o->spawn_max[1] = 5.f;
    o->spawn_max[2] = 5.f;
    o->spawn_random = 0;
    o->spacing = 0.1f;
    o->bounds_max[0] = o->bounds_max[1] = o->bounds_max[2] = 6.0f;
    o->bound_damping = 0.5f;
    o->hash_size = 1u << 17;
    o->enable_scorr = 1;
    o->k_corr = 0.001f;
    o->n_corr = 4.0f;
    o->delta_q = 0.3f;
    o->cohesion_strength = 0.001f;
    o->enable_viscosity = 1;
    o->viscosity = 0.1f;
    o->interaction_radius = 1.5f;
    o->interaction_strength = 25.0f;
    o->enable_vorticity = 1;
    o->vorticity_epsilon = 0.05f;
    o->enable_apbf = 0;
    o->min_lod = 2;
    o->max_lod = 7;
    o->lod_max_dist = 15.0f;
    return FSI_OK;
}

int fsi_device_count(int* count) {
    if (!count) return fail(FSI_ERR_INVALID_ARG, "count is NULL");
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e == hipErrorNoDevice) {
        *count = 0;
        return FSI_OK;
    }
    if (e != hipSuccess) return fail(FSI_ERR_HIP, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
    return FSI_OK;
}

int fsi_pbf_create(const fsi_fluid_config* cfg, fsi_pbf** out) {
    if (!out) return fail(FSI_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    int rc = check_config(cfg);
    if (rc) return rc;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return fail(FSI_ERR_NO_DEVICE, "no HIP device visible");
    fsi_pbf* h = new (std::nothrow) fsi_pbf();
    if (!h) return fail(FSI_ERR_INVALID_ARG, "out of host memory");
    if ((rc = core_init(h->core))) {
        core_destroy(h->core);
        delete h;
        return rc;
    }
    h->core.d.relabel = true;  // storage in Morton order of the buckets (processing order 2)
    h->cfg = *cfg;
    apply_bounds(h, cfg->bounds_min, cfg->bounds_max, cfg->bound_damping);
    repack(h);
    *out = h;
    return FSI_OK;
}

int fsi_pbf_destroy(fsi_pbf* h) {
    if (!h) return FSI_OK;
    (void)hipSetDevice(h->core.device);
    if (h->core.stream) (void)hipStreamSynchronize(h->core.stream);
    dev_free(&h->contacts);
    dev_free(&h->contact_count);
    dev_free(&h->pred);
    dev_free(&h->sdf_grid);
    dev_free(&h->sdf_brange);
    core_destroy(h->core);
    delete h;
    return FSI_OK;
}

int fsi_pbf_set_params(fsi_pbf* h, const fsi_fluid_config* cfg) {
    int rc = bind(h);
    if (rc) return rc;
    if ((rc = check_config(cfg))) return rc;
    const uint32_t old_hash = h->cfg.hash_size;
    h->cfg = *cfg;
    h->cfg.particle_count = h->core.d.n;
    apply_bounds(h, cfg->bounds_min, cfg->bounds_max, cfg->bound_damping);
    repack(h);
    if (h->core.capacity > 0 && old_hash != h->cfg.hash_size) {
        FSI_HIP_TRY(hipStreamSynchronize(h->core.stream));
        FSI_HIP_TRY(core_alloc_cells(h->core, h->cfg.hash_size));
        h->core.stepped = false;
    }
    return FSI_OK;
}

int fsi_pbf_get_params(fsi_pbf* h, fsi_fluid_config* out) {
    if (!h || !out) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    *out = h->cfg;
    return FSI_OK;
}

int fsi_pbf_get_constants(fsi_pbf* h, fsi_pbf_constants* out) {
    if (!h || !out) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    *out = h->k;
    return FSI_OK;
}

int fsi_pbf_set_particles(fsi_pbf* h, const fsi_particle* ps, uint32_t n) {
    int rc = bind(h);
    if (rc) return rc;
    if (n > 0 && !ps) return fail(FSI_ERR_INVALID_ARG, "particles is NULL");
    if (n > (1u << 30)) return fail(FSI_ERR_INVALID_ARG, "more than 2^30 particles");
    Core& c = h->core;
    FSI_HIP_TRY(hipStreamSynchronize(c.stream));
    if (n != c.capacity || !c.d.pos) {
        FSI_HIP_TRY(core_alloc_particles(c, n));
        FSI_HIP_TRY(dev_alloc(&h->contacts, 2 * (size_t)n));
        FSI_HIP_TRY(dev_alloc(&h->contact_count, 1));
        FSI_HIP_TRY(dev_alloc(&h->pred, n));
    }
    c.d.n = n;
    h->cfg.particle_count = n;
    FSI_HIP_TRY(core_alloc_cells(c, h->cfg.hash_size));
    repack(h);
    c.stepped = false;
    h->contacts_valid = false;
    h->host_pred.assign(4 * (size_t)n, 0.0f);
    if (n == 0) return FSI_OK;
    std::vector<float> p4(4 * (size_t)n), v4(4 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
        const fsi_particle& q = ps[i];
        for (int a = 0; a < 3; ++a) {
            p4[4 * (size_t)i + a] = q.pos[a];
            v4[4 * (size_t)i + a] = q.vel[a];
            h->host_pred[4 * (size_t)i + a] = q.pred_pos[a];
        }
        p4[4 * (size_t)i + 3] = q.pad0;
        v4[4 * (size_t)i + 3] = q.pad2;
        h->host_pred[4 * (size_t)i + 3] = q.pad1;
    }
    FSI_HIP_TRY(hipMemcpyAsync(c.d.pos, p4.data(), p4.size() * sizeof(float), hipMemcpyHostToDevice, c.stream));
    FSI_HIP_TRY(hipMemcpyAsync(c.d.vel, v4.data(), v4.size() * sizeof(float), hipMemcpyHostToDevice, c.stream));
    FSI_HIP_TRY(hipMemcpyAsync(h->pred, h->host_pred.data(), h->host_pred.size() * sizeof(float),
                               hipMemcpyHostToDevice, c.stream));
    // SolverBuffer starts zeroed (src/fluid/PBFluids.cpp:123-124)
    FSI_HIP_TRY(hipMemsetAsync(c.d.P, 0, sizeof(float4) * n, c.stream));
    FSI_HIP_TRY(hipMemsetAsync(c.d.PL, 0, sizeof(float4) * n, c.stream));
    FSI_HIP_TRY(hipMemsetAsync(c.d.curl, 0, sizeof(float4) * n, c.stream));
    FSI_HIP_TRY(hipMemsetAsync(c.d.dp_tap, 0, sizeof(float4) * n, c.stream));
    FSI_HIP_TRY(hipMemsetAsync(c.d.rho, 0, sizeof(float) * n, c.stream));
    FSI_HIP_TRY(hipStreamSynchronize(c.stream));
    return FSI_OK;
}

int fsi_pbf_particle_count(fsi_pbf* h, uint32_t* n) {
    if (!h || !n) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    *n = h->core.d.n;
    return FSI_OK;
}

int fsi_pbf_read_particles(fsi_pbf* h, fsi_particle* out, uint32_t cap, uint32_t* n_out) {
    int rc = bind(h);
    if (rc) return rc;
    const uint32_t n = h->core.d.n;
    if (n_out) *n_out = n;
    if (n == 0) return FSI_OK;
    if (!out) return fail(FSI_ERR_INVALID_ARG, "out is NULL");
    if (cap < n) return fail(FSI_ERR_CAPACITY, "cap smaller than the particle count");
    std::vector<float> p4(4 * (size_t)n), v4(4 * (size_t)n);
    hipStream_t st = h->core.stream;
    FSI_HIP_TRY(hipMemcpyAsync(p4.data(), h->core.d.pos, p4.size() * sizeof(float), hipMemcpyDeviceToHost, st));
    FSI_HIP_TRY(hipMemcpyAsync(v4.data(), h->core.d.vel, v4.size() * sizeof(float), hipMemcpyDeviceToHost, st));
    FSI_HIP_TRY(hipStreamSynchronize(st));
    for (uint32_t i = 0; i < n; ++i) {
        fsi_particle& q = out[i];
        for (int a = 0; a < 3; ++a) {
            q.pos[a] = p4[4 * (size_t)i + a];
            q.vel[a] = v4[4 * (size_t)i + a];
            q.pred_pos[a] = h->host_pred[4 * (size_t)i + a];
        }
        q.pad0 = p4[4 * (size_t)i + 3];
        q.pad1 = h->host_pred[4 * (size_t)i + 3];
        q.pad2 = v4[4 * (size_t)i + 3];
    }
    return FSI_OK;
}

namespace {
// 48-byte Particle records {pos, predPos, vel} (three float4 each, id order).
__global__ __launch_bounds__(256) void k_pack_render(uint32_t n, const float4* __restrict__ pos,
                                                     const float4* __restrict__ pred,
                                                     const float4* __restrict__ vel, float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    out[3 * (size_t)i] = pos[i];
    out[3 * (size_t)i + 1] = pred[i];
    out[3 * (size_t)i + 2] = vel[i];
}
}  // namespace

int fsi_pbf_write_render_buffer(fsi_pbf* h, void* device_out, void* stream) {
    int rc = bind(h);
    if (rc) return rc;
    const uint32_t n = h->core.d.n;
    if (n == 0) return FSI_OK;
    if (!device_out) return fail(FSI_ERR_INVALID_ARG, "device_out is NULL");
    if ((uintptr_t)device_out % alignof(float4) != 0) return fail(FSI_ERR_INVALID_ARG, "device_out is not 16-byte aligned");
    hipStream_t st = h->core.stream;
    hipLaunchKernelGGL(k_pack_render, dim3((n + 255u) / 256u), dim3(256), 0, st, n, h->core.d.pos, h->pred,
                       h->core.d.vel, (float4*)device_out);
    FSI_HIP_TRY(hipGetLastError());
    if (stream) {
        // the caller's stream (e.g. the renderer's) waits for the pack
        hipEvent_t ev = nullptr;
        FSI_HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        hipError_t e = hipEventRecord(ev, st);
        if (e == hipSuccess) e = hipStreamWaitEvent((hipStream_t)stream, ev, 0);
        (void)hipEventDestroy(ev);
        FSI_HIP_TRY(e);
    }
    return FSI_OK;
}

int fsi_pbf_set_collision_padding(fsi_pbf* h, float padding) {
    if (!h) return fail(FSI_ERR_INVALID_ARG, "handle is NULL");
    h->collision_padding = padding;
    return FSI_OK;
}

int fsi_pbf_set_bounds(fsi_pbf* h, const float mn[3], const float mx[3], float damping) {
    if (!h || !mn || !mx) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    apply_bounds(h, mn, mx, damping);
    repack(h);
    return FSI_OK;
}

int fsi_pbf_set_sdf(fsi_pbf* h, const float* grid, const int32_t dims[3], const float origin[3], float cell) {
    int rc = bind(h);
    if (rc) return rc;
    FSI_HIP_TRY(hipStreamSynchronize(h->core.stream));
    if (!grid) {
        dev_free(&h->sdf_grid);
        dev_free(&h->sdf_brange);
        h->sdf_valid = false;
        return FSI_OK;
    }
    if (!dims || !origin) return fail(FSI_ERR_INVALID_ARG, "dims/origin is NULL");
    if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return fail(FSI_ERR_INVALID_ARG, "dims must be positive");
    if (!(cell > 0.0f)) return fail(FSI_ERR_INVALID_ARG, "cell must be positive");
    const size_t count = (size_t)dims[0] * (size_t)dims[1] * (size_t)dims[2];
    if (count > (size_t)1 << 31) return fail(FSI_ERR_INVALID_ARG, "SDF grid above 2^31 voxels");
    const bool same =
        h->sdf_grid && dims[0] == h->sdf_dims[0] && dims[1] == h->sdf_dims[1] && dims[2] == h->sdf_dims[2];
    if (!same) {
        FSI_HIP_TRY(dev_alloc(&h->sdf_grid, count));
        FSI_HIP_TRY(dev_alloc(&h->sdf_brange, sdf_block_count(dims)));
    }
    FSI_HIP_TRY(hipMemcpy(h->sdf_grid, grid, count * sizeof(float), hipMemcpyHostToDevice));
    FSI_HIP_TRY(launch_sdf_block_range(h->sdf_grid, dims, h->sdf_brange, h->core.stream));
    for (int a = 0; a < 3; ++a) {
        h->sdf_dims[a] = dims[a];
        h->sdf_origin[a] = origin[a];
    }
    h->sdf_cell = cell;
    h->sdf_valid = true;
    return FSI_OK;
}

int fsi_pbf_set_sdf_padding(fsi_pbf* h, float padding) {
    if (!h) return fail(FSI_ERR_INVALID_ARG, "handle is NULL");
    h->sdf_padding = padding;
    return FSI_OK;
}

int fsi_pbf_set_sdf_containment(fsi_pbf* h, int inside) {
    if (!h) return fail(FSI_ERR_INVALID_ARG, "handle is NULL");
    h->sdf_containment = inside != 0;
    return FSI_OK;
}

int fsi_pbf_set_interaction(fsi_pbf* h, int active, const float pos[3]) {
    if (!h) return fail(FSI_ERR_INVALID_ARG, "handle is NULL");
    h->interacting = active != 0;
    if (pos)
        for (int a = 0; a < 3; ++a) h->ipos[a] = pos[a];
    return FSI_OK;
}

int fsi_pbf_set_camera(fsi_pbf* h, const float pos[3]) {
    if (!h || !pos) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    for (int a = 0; a < 3; ++a) h->cam[a] = pos[a];
    return FSI_OK;
}

int fsi_pbf_step(fsi_pbf* h) {
    int rc = bind(h);
    if (rc) return rc;
    if (h->core.d.n == 0) return FSI_OK;  // PBFluids.cpp:325
    const StepConst c = step_const(h);
    h->core.d.batched = false;
    h->core.d.any_sdf = c.dom.enable_sdf == 1u;
    h->core.d.tiled = c.dom.dshift == kTileShift;
    rc = core_frame(h->core, c, frame_plan(h));
    h->contacts_valid = false;
    return rc;
}

int fsi_pbf_sync(fsi_pbf* h) {
    int rc = bind(h);
    if (rc) return rc;
    FSI_HIP_TRY(hipStreamSynchronize(h->core.stream));
    return FSI_OK;
}

int fsi_pbf_filter_contacts(fsi_pbf* h, float band) {
    int rc = bind(h);
    if (rc) return rc;
    // PBFluids.cpp:195-196: no-op without an SDF, with band <= 0 or before setParticles
    if (!h->sdf_valid || !(band > 0.0f) || h->core.d.n == 0) return FSI_OK;
    const StepConst c = step_const(h);
    h->core.d.batched = false;
    FSI_HIP_TRY(launch_contacts(c, h->core.d, band, h->contacts, h->contact_count, h->core.stream));
    h->contacts_valid = true;
    return FSI_OK;
}

int fsi_pbf_read_contacts(fsi_pbf* h, fsi_contact_candidate* out, uint32_t cap, uint32_t* count) {
    int rc = bind(h);
    if (rc) return rc;
    if (!count) return fail(FSI_ERR_INVALID_ARG, "count is NULL");
    *count = 0;
    if (!h->contacts_valid || h->core.d.n == 0) return FSI_OK;
    uint32_t c = 0;
    FSI_HIP_TRY(hipMemcpyAsync(&c, h->contact_count, sizeof(uint32_t), hipMemcpyDeviceToHost, h->core.stream));
    FSI_HIP_TRY(hipStreamSynchronize(h->core.stream));
    c = std::min(c, h->core.d.n);
    *count = c;
    const uint32_t m = std::min(c, cap);
    if (m > 0) {
        if (!out) return fail(FSI_ERR_INVALID_ARG, "out is NULL");
        FSI_HIP_TRY(hipMemcpy(out, h->contacts, sizeof(fsi_contact_candidate) * m, hipMemcpyDeviceToHost));
    }
    return FSI_OK;
}

int fsi_pbf_set_sdf_from_mesh(fsi_pbf* h, fsi_mesh* m) {
    int rc = bind(h);
    if (rc) return rc;
    if (!m) return fail(FSI_ERR_INVALID_ARG, "mesh handle is NULL");
    if (mesh_device(m) != h->core.device) return fail(FSI_ERR_INVALID_ARG, "mesh and solver on different devices");
    const float* grid = nullptr;
    int32_t dims[3];
    double origin[3], cell = 0.0;
    hipEvent_t ready = nullptr;
    if ((rc = mesh_grid_for_solver(m, &grid, dims, origin, &cell, &ready))) return rc;
    const size_t count = (size_t)dims[0] * (size_t)dims[1] * (size_t)dims[2];
    const bool same =
        h->sdf_grid && dims[0] == h->sdf_dims[0] && dims[1] == h->sdf_dims[1] && dims[2] == h->sdf_dims[2];
    if (!same) {
        FSI_HIP_TRY(hipStreamSynchronize(h->core.stream));
        FSI_HIP_TRY(dev_alloc(&h->sdf_grid, count));
        FSI_HIP_TRY(dev_alloc(&h->sdf_brange, sdf_block_count(dims)));
    }
    FSI_HIP_TRY(hipStreamWaitEvent(h->core.stream, ready, 0));
    FSI_HIP_TRY(hipMemcpyAsync(h->sdf_grid, grid, count * sizeof(float), hipMemcpyDeviceToDevice, h->core.stream));
    FSI_HIP_TRY(launch_sdf_block_range(h->sdf_grid, dims, h->sdf_brange, h->core.stream));
    if ((rc = mesh_wait_for(m, h->core.stream))) return rc;
    for (int a = 0; a < 3; ++a) {
        h->sdf_dims[a] = dims[a];
        h->sdf_origin[a] = (float)origin[a];
    }
    h->sdf_cell = (float)cell;
    h->sdf_valid = true;
    return FSI_OK;
}

int fsi_pbf_contact_forces(fsi_pbf* h, fsi_mesh* m, float band, int use_vel_magnitude, float control_threshold,
                           float* forces, uint32_t counts[2]) {
    int rc = bind(h);
    if (rc) return rc;
    if (!m) return fail(FSI_ERR_INVALID_ARG, "mesh handle is NULL");
    if (mesh_device(m) != h->core.device) return fail(FSI_ERR_INVALID_ARG, "mesh and solver on different devices");
    if (!h->contacts_valid || h->core.d.n == 0)
        return fail(FSI_ERR_STATE, "no contact candidates (fsi_pbf_filter_contacts first)");
    return mesh_contact_forces(m, h->contacts, h->contact_count, h->core.d.n, h->core.stream, band,
                               use_vel_magnitude, control_threshold, forces, counts);
}

int fsi_pbf_read_grid(fsi_pbf* h, uint32_t* keys, uint32_t* ids, int32_t* cells) {
    int rc = bind(h);
    if (rc) return rc;
    const Core& c = h->core;
    const uint32_t n = c.d.n;
    FSI_HIP_TRY(hipStreamSynchronize(c.stream));
    if (keys && n) {
        if (c.stepped) FSI_HIP_TRY(hipMemcpy(keys, c.d.skeys, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        else std::memset(keys, 0, sizeof(uint32_t) * n);
    }
    if (ids && n) {
        if (c.stepped) FSI_HIP_TRY(hipMemcpy(ids, c.d.perm, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        else std::memset(ids, 0, sizeof(uint32_t) * n);
    }
    if (cells && c.d.cells)
        FSI_HIP_TRY(hipMemcpy(cells, c.d.cells, sizeof(int2) * (size_t)c.d.hash_size, hipMemcpyDeviceToHost));
    return FSI_OK;
}

int fsi_pbf_read_solver(fsi_pbf* h, float* ppl, float* dpr) {
    int rc = bind(h);
    if (rc) return rc;
    const Core& c = h->core;
    const uint32_t n = c.d.n;
    if (n == 0) return FSI_OK;
    FSI_HIP_TRY(hipStreamSynchronize(c.stream));
    if (!c.stepped) {
        if (ppl) std::memset(ppl, 0, sizeof(float) * 4 * n);
        if (dpr) std::memset(dpr, 0, sizeof(float) * 4 * n);
        return FSI_OK;
    }
    std::vector<uint32_t> perm(n);
    // storage slot -> particle id of the last substep's arrays
    FSI_HIP_TRY(hipMemcpy(perm.data(), c.d.stored_relabel ? c.d.mperm : c.d.perm, sizeof(uint32_t) * n,
                          hipMemcpyDeviceToHost));
    std::vector<float> tmp(4 * (size_t)n);
    if (ppl) {
        FSI_HIP_TRY(hipMemcpy(tmp.data(), c.d.P, sizeof(float4) * n, hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < n; ++k) tmp[4 * (size_t)k + 3] = 0.0f;  // Integrate resets lambda
        unpermute(perm, tmp.data(), ppl, 4);
    }
    if (dpr) {
        const bool vort = h->cfg.enable_vorticity != 0;
        FSI_HIP_TRY(hipMemcpy(tmp.data(), vort ? c.d.curl : c.d.dp_tap, sizeof(float4) * n, hipMemcpyDeviceToHost));
        std::vector<float> rho(n);
        FSI_HIP_TRY(hipMemcpy(rho.data(), c.d.rho, sizeof(float) * n, hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < n; ++k) tmp[4 * (size_t)k + 3] = rho[k];
        unpermute(perm, tmp.data(), dpr, 4);
    }
    return FSI_OK;
}

int fsi_pbf_set_taps(fsi_pbf* h, int enable) {
    if (!h) return fail(FSI_ERR_INVALID_ARG, "handle is NULL");
    h->core.taps = enable != 0;
    return FSI_OK;
}

int fsi_pbf_read_counts(fsi_pbf* h, uint32_t* cand, uint32_t* nbr) {
    int rc = bind(h);
    if (rc) return rc;
    const Core& c = h->core;
    if (!c.taps) return fail(FSI_ERR_STATE, "taps are off (fsi_pbf_set_taps)");
    const uint32_t n = c.d.n;
    if (n == 0) return FSI_OK;
    FSI_HIP_TRY(hipStreamSynchronize(c.stream));
    if (!c.stepped) return fail(FSI_ERR_STATE, "no step has run");
    std::vector<uint32_t> perm(n), tmp(n);
    // storage slot -> particle id of the last substep's arrays
    FSI_HIP_TRY(hipMemcpy(perm.data(), c.d.stored_relabel ? c.d.mperm : c.d.perm, sizeof(uint32_t) * n,
                          hipMemcpyDeviceToHost));
    if (cand) {
        FSI_HIP_TRY(hipMemcpy(tmp.data(), c.d.cand, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < n; ++k) cand[perm[k]] = tmp[k];
    }
    if (nbr) {
        FSI_HIP_TRY(hipMemcpy(tmp.data(), c.d.nbr, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < n; ++k) nbr[perm[k]] = tmp[k];
    }
    return FSI_OK;
}

int fsi_pbf_set_timing(fsi_pbf* h, int enable) {
    if (!h) return fail(FSI_ERR_INVALID_ARG, "handle is NULL");
    h->core.timing = enable != 0;
    h->core.timing_mask = 0xffu;
    return FSI_OK;
}

int fsi_pbf_set_timing_passes(fsi_pbf* h, uint32_t mask) {
    if (!h) return fail(FSI_ERR_INVALID_ARG, "handle is NULL");
    h->core.timing = mask != 0;
    h->core.timing_mask = mask;
    return FSI_OK;
}

int fsi_pbf_timing_reset(fsi_pbf* h) {
    int rc = bind(h);
    if (rc) return rc;
    if ((rc = core_drain_timing(h->core))) return rc;
    std::memset(&h->core.acc, 0, sizeof(h->core.acc));
    return FSI_OK;
}

int fsi_pbf_read_timing(fsi_pbf* h, fsi_pbf_timing* out) {
    int rc = bind(h);
    if (rc) return rc;
    if (!out) return fail(FSI_ERR_INVALID_ARG, "out is NULL");
    if ((rc = core_drain_timing(h->core))) return rc;
    *out = h->core.acc;
    return FSI_OK;
}

int fsi_pbf_set_tuning(fsi_pbf* h, int knob, int value) {
    if (!h) return fail(FSI_ERR_INVALID_ARG, "handle is NULL");
    return core_set_tuning(h->core, knob, value);
}

int fsi_pbf_get_tuning(fsi_pbf* h, int knob, int* value) {
    if (!h || !value) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    return core_get_tuning(h->core, knob, value);
}

int fsi_pbf_autotune(fsi_pbf* h, uint32_t frames, int* gather, int* digit_bits) {
    int rc = bind(h);
    if (rc) return rc;
    h->core.d.batched = false;
    // the binning follows the gather form in step_const
    const Rebin rebin = [h](bool, StepConst& k) {
        k = step_const(h);
        h->core.d.any_sdf = k.dom.enable_sdf == 1u;
        h->core.d.tiled = k.dom.dshift == kTileShift;
        return (int)FSI_OK;
    };
    return core_autotune(h->core, rebin, frame_plan(h), frames, gather, digit_bits);
}

int fsi_pbf_read_autotune(fsi_pbf* h, fsi_tune_trial* out, uint32_t cap, uint32_t* n) {
    if (!h) return fail(FSI_ERR_INVALID_ARG, "handle is NULL");
    return core_read_autotune(h->core, out, cap, n);
}

int fsi_pbf_set_graph(fsi_pbf* h, int enable) {
    int rc = bind(h);
    if (rc) return rc;
    FSI_HIP_TRY(hipStreamSynchronize(h->core.stream));
    h->core.use_graph = enable != 0;
    if (!h->core.use_graph) core_invalidate_graph(h->core);
    return FSI_OK;
}

int fsi_pbf_get_stream(fsi_pbf* h, void** stream) {
    if (!h || !stream) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    *stream = (void*)h->core.stream;
    return FSI_OK;
}

int fsi_pbf_set_stream(fsi_pbf* h, void* stream) {
    int rc = bind(h);
    if (rc) return rc;
    return core_set_stream(h->core, (hipStream_t)stream);
}

}  // extern "C"
