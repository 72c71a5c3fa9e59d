// pbf_capi.hip — solver handle and C-ABI entry points (include/fsi_pbf.h).
//
// The handle replaces class PBFluids (include/fluid/PBFluids.h:88-268): it owns
// the device buffers that were SSBOs, one HIP stream in place of the GL command
// stream, and enqueues the substep kernels in the order of PBFluids::step
// (src/fluid/PBFluids.cpp:323-517).  Stream order replaces the
// glMemoryBarrier after every dispatch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/fsi_pbf.h"
#include "pbf_launch.h"

using namespace fsi;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return fail(FSI_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));       \
    } while (0)

bool is_pow2(uint32_t x) { return x != 0 && (x & (x - 1)) == 0; }

int log2u(uint32_t x) {
    int b = 0;
    while ((1u << b) < x) ++b;
    return b;
}

template <class T>
hipError_t dev_alloc(T** p, size_t count) {
    if (*p) {
        hipError_t e = hipFree(*p);
        *p = nullptr;
        if (e != hipSuccess) return e;
    }
    if (count == 0) return hipSuccess;
    return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * count);
}

template <class T>
void dev_free(T** p) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
}

struct TimedPair {
    int pass;
    hipEvent_t a, b;
};

}  // namespace

struct fsi_pbf {
    int device = 0;
    hipStream_t stream = nullptr;
    fsi_fluid_config cfg{};
    fsi_pbf_constants k{};
    float collision_padding = 0.0f;
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0}, damping = 0.0f;
    // SDF boundary
    float* sdf_grid = nullptr;
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
    // state
    DeviceState d{};
    uint32_t capacity = 0;
    uint32_t cells_capacity = 0;
    std::vector<float> host_pred;  // predPos.xyz + pad1 per particle (never touched on device)
    bool stepped = false;
    bool taps = false;
    int sort_digit_bits = kSortMaxDigitBits;
    // contacts
    float4* contacts = nullptr;
    uint32_t* contact_count = nullptr;
    bool contacts_valid = false;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> event_pool;
    std::vector<TimedPair> pending;
    fsi_pbf_timing acc{};
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

// calcPoly6Kernel of src/fluid/Kernels.cpp:9-27 (reached through
// calcSCorrKernel, :216-219) in FP32: the s_corr reference value wq.
float wq_of(float dq, float h) {
    const float pi = 3.14159265358979323846f;
    const float coeff = 315.0f / (64.0f * pi);
    const float h2 = h * h;
    const float r2 = dq * dq + 0.0f * 0.0f + 0.0f * 0.0f;
    if (r2 > h2) return 0.0f;
    const float h4 = h2 * h2;
    const float h9 = h4 * h4 * h;
    const float diff = h2 - r2;
    const float diff3 = diff * diff * diff;
    return (coeff / h9) * diff3;
}

// The constant block of PBFluids::setParams (src/fluid/PBFluids.cpp:44-101).
void pack_constants(fsi_pbf* h) {
    const fsi_fluid_config& p = h->cfg;
    fsi_pbf_constants& u = h->k;
    std::memset(&u, 0, sizeof(u));
    for (int a = 0; a < 3; ++a) {
        u.bounds_min[a] = h->bmin[a];
        u.bounds_max[a] = h->bmax[a];
        u.gravity_dt[a] = p.gravity[a];
    }
    u.gravity_dt[3] = p.dt / (float)p.substep_iterations;
    u.h = p.h;
    u.rho0 = p.rho0;
    u.eps = p.eps;
    u.wq = wq_of(p.delta_q * p.h, p.h);
    u.k_corr = p.k_corr;
    u.n_corr = p.n_corr;
    u.viscosity = p.viscosity;
    u.bound_damping = h->damping;
    u.hash_size = p.hash_size;
    u.particle_count = h->d.n;
    u.enable_scorr = p.enable_scorr ? 1u : 0u;
    u.enable_viscosity = p.enable_viscosity ? 1u : 0u;
    u.cohesion_strength = p.cohesion_strength;
    u.interaction_radius = p.interaction_radius;
    u.interaction_strength = p.interaction_strength;
    const float PI = 3.14159265358979323846f;
    const float hh = p.h;
    const float h2 = hh * hh;
    const float h3 = h2 * hh;
    const float h6 = h3 * h3;
    const float h9 = h6 * h3;
    u.w0_self = (315.0f / (64.0f * PI)) / h3;
    u.poly6_coeff = 315.0f / (64.0f * PI * h9);
    u.spiky_coeff = 45.0f / (PI * h6);
    u.inv_rho0 = 1.0f / p.rho0;
    u.hash_mask = p.hash_size - 1u;
    u.min_lod = p.enable_apbf ? p.min_lod : p.solver_iterations;
    u.max_lod = p.enable_apbf ? p.max_lod : p.solver_iterations;
    u.lod_max_dist = p.lod_max_dist;
    u.enable_apbf = p.enable_apbf ? 1u : 0u;
}

// Morton binning of one scene's processing order: cover its bounds (plus two
// cells) with at most 2^axis_bits bins per axis.
void set_binning(Domain& d, const float bmin[3], const float bmax[3], float h, uint32_t axis_bits,
                 uint32_t bin_base) {
    auto cell_clamped = [&](float x) {
        const double v = std::floor((double)x / (double)h);
        return (int64_t)std::max(-1.0e9, std::min(1.0e9, std::isfinite(v) ? v : 0.0));
    };
    int64_t span = 1;
    for (int a = 0; a < 3; ++a) {
        const int64_t lo = cell_clamped(bmin[a]) - 2;
        const int64_t hi = cell_clamped(bmax[a]) + 2;
        d.dcmin[a] = (int32_t)lo;
        span = std::max<int64_t>(span, hi - lo + 1);
    }
    uint32_t shift = 0;
    while ((span >> shift) > ((int64_t)1 << axis_bits) && shift < 31) ++shift;
    d.dshift = shift;
    d.axis_bits = axis_bits;
    d.bin_base = bin_base;
}

// The constants shared by every scene of a step (FluidConfigUBO minus the box).
StepConst step_const_common(const fsi_fluid_config& cfg, const fsi_pbf_constants& u, uint32_t n, bool interacting,
                            const float ipos[3], const float cam[3]) {
    StepConst c{};
    for (int a = 0; a < 3; ++a) {
        c.g[a] = u.gravity_dt[a];
        c.ipos[a] = ipos[a];
        c.cam[a] = cam[a];
    }
    c.sub_dt = u.gravity_dt[3];
    c.h = u.h;
    c.rho0 = u.rho0;
    c.eps = u.eps;
    c.wq = u.wq;
    c.k_corr = u.k_corr;
    c.n_corr = u.n_corr;
    c.viscosity = u.viscosity;
    c.damping = u.bound_damping;
    c.cohesion = u.cohesion_strength;
    c.irad = u.interaction_radius;
    c.istr = u.interaction_strength;
    c.w0_self = u.w0_self;
    c.poly6 = u.poly6_coeff;
    c.spiky = u.spiky_coeff;
    c.inv_rho0 = u.inv_rho0;
    c.lod_max_dist = u.lod_max_dist;
    c.n = n;
    c.enable_scorr = u.enable_scorr;
    c.enable_visc = u.enable_viscosity;
    c.min_lod = u.min_lod;
    c.max_lod = u.max_lod;
    c.enable_apbf = u.enable_apbf;
    c.enable_vort = cfg.enable_vorticity ? 1u : 0u;
    c.interacting = interacting ? 1u : 0u;
    c.vort_eps = cfg.vorticity_epsilon;
    return c;
}

StepConst step_const(const fsi_pbf* h) {
    const fsi_pbf_constants& u = h->k;
    StepConst c = step_const_common(h->cfg, u, h->d.n, h->interacting, h->ipos, h->cam);
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
    set_binning(d, u.bounds_min, u.bounds_max, u.h, 7, 0);
    return c;
}

int check_config(const fsi_fluid_config* cfg) {
    if (!cfg) return fail(FSI_ERR_INVALID_ARG, "config is NULL");
    if (!is_pow2(cfg->hash_size))
        return fail(FSI_ERR_INVALID_ARG, "hash_size must be a power of two (hashMask = hashSize - 1)");
    if (cfg->hash_size > (1u << 30)) return fail(FSI_ERR_INVALID_ARG, "hash_size above 2^30");
    if (cfg->substep_iterations == 0) return fail(FSI_ERR_INVALID_ARG, "substep_iterations must be >= 1");
    if (!(cfg->h > 0.0f) || !std::isfinite(cfg->h)) return fail(FSI_ERR_INVALID_ARG, "h must be positive");
    if (!(cfg->rho0 != 0.0f) || !std::isfinite(cfg->rho0)) return fail(FSI_ERR_INVALID_ARG, "rho0 must be nonzero");
    if (cfg->enable_apbf && cfg->max_lod > 1024u) return fail(FSI_ERR_INVALID_ARG, "max_lod above 1024");
    if (!cfg->enable_apbf && cfg->solver_iterations > 1024u)
        return fail(FSI_ERR_INVALID_ARG, "solver_iterations above 1024");
    return FSI_OK;
}

hipError_t alloc_cells(fsi_pbf* h) {
    if (h->cells_capacity == h->cfg.hash_size && h->d.cells) return hipSuccess;
    hipError_t e = dev_alloc(&h->d.cells, h->cfg.hash_size);
    if (e != hipSuccess) return e;
    h->cells_capacity = h->cfg.hash_size;
    h->d.hash_size = h->cfg.hash_size;
    std::vector<int32_t> neg(2 * (size_t)h->cfg.hash_size, -1);
    return hipMemcpyAsync(h->d.cells, neg.data(), neg.size() * sizeof(int32_t), hipMemcpyHostToDevice, h->stream);
}

hipError_t alloc_particles(fsi_pbf* h, uint32_t n) {
    DeviceState& d = h->d;
    hipError_t e = hipSuccess;
#define ALLOC(ptr, count)                          \
    if ((e = dev_alloc(&(ptr), (count))) != hipSuccess) return e;
    ALLOC(d.pos, n);
    ALLOC(d.vel, n);
    ALLOC(d.xstar, n);
    ALLOC(d.keys, n);
    ALLOC(d.P, n);
    ALLOC(d.PL, n);
    ALLOC(d.vs, n);
    ALLOC(d.curl, n);
    ALLOC(d.posS, n);
    ALLOC(d.rho, n);
    ALLOC(d.lod, n);
    ALLOC(d.dp_tap, n);
    ALLOC(d.cand, n);
    ALLOC(d.nbr, n);
    ALLOC(d.flags, n);
    ALLOC(d.slots, n);
    ALLOC(d.bins, kOrderBins);
    ALLOC(d.bin_of, n);
    ALLOC(d.bin_rank, n);
    ALLOC(d.order, n);
    ALLOC(d.list_count, n);
    ALLOC(d.list, (size_t)kNbrCap * n);
    ALLOC(d.sort.keys[0], n);
    ALLOC(d.sort.keys[1], n);
    ALLOC(d.sort.ids[0], n);
    ALLOC(d.sort.ids[1], n);
    const size_t hist = sort_hist_entries(n, kSortMaxDigitBits);
    ALLOC(d.sort.hist, hist);
    const size_t parts = std::max({scan_partials_needed(hist), scan_partials_needed(n),
                                   scan_partials_needed(kOrderBins)});
    ALLOC(d.sort.partial, parts);
    ALLOC(h->contacts, 2 * (size_t)n);
    ALLOC(h->contact_count, 1);
#undef ALLOC
    h->capacity = n;
    return hipSuccess;
}

void free_all(fsi_pbf* h) {
    DeviceState& d = h->d;
    dev_free(&d.pos);
    dev_free(&d.vel);
    dev_free(&d.xstar);
    dev_free(&d.keys);
    dev_free(&d.cells);
    dev_free(&d.P);
    dev_free(&d.PL);
    dev_free(&d.vs);
    dev_free(&d.curl);
    dev_free(&d.posS);
    dev_free(&d.rho);
    dev_free(&d.lod);
    dev_free(&d.dp_tap);
    dev_free(&d.cand);
    dev_free(&d.nbr);
    dev_free(&d.flags);
    dev_free(&d.slots);
    dev_free(&d.bins);
    dev_free(&d.bin_of);
    dev_free(&d.bin_rank);
    dev_free(&d.order);
    dev_free(&d.list_count);
    dev_free(&d.list);
    dev_free(&d.sort.keys[0]);
    dev_free(&d.sort.keys[1]);
    dev_free(&d.sort.ids[0]);
    dev_free(&d.sort.ids[1]);
    dev_free(&d.sort.hist);
    dev_free(&d.sort.partial);
    dev_free(&h->contacts);
    dev_free(&h->contact_count);
    dev_free(&h->sdf_grid);
}

// ---- timing ----

hipEvent_t take_event(fsi_pbf* h) {
    if (!h->event_pool.empty()) {
        hipEvent_t e = h->event_pool.back();
        h->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct PassTimer {
    fsi_pbf* h;
    int pass;
    hipEvent_t a = nullptr, b = nullptr;
    PassTimer(fsi_pbf* hh, int p) : h(hh), pass(p) {
        if (!h->timing) return;
        a = take_event(h);
        b = take_event(h);
        if (a && b) (void)hipEventRecord(a, h->stream);
    }
    ~PassTimer() {
        if (!h->timing || !a || !b) return;
        (void)hipEventRecord(b, h->stream);
        h->pending.push_back(TimedPair{pass, a, b});
    }
};

int drain_timing(fsi_pbf* h) {
    if (h->pending.empty()) return FSI_OK;
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (const TimedPair& t : h->pending) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, t.a, t.b));
        h->acc.ms[t.pass] += ms;
        h->acc.launches[t.pass] += 1;
        h->event_pool.push_back(t.a);
        h->event_pool.push_back(t.b);
    }
    h->pending.clear();
    return FSI_OK;
}

int bind(fsi_pbf* h) {
    if (!h) return fail(FSI_ERR_INVALID_ARG, "handle is NULL");
    HIP_TRY(hipSetDevice(h->device));
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

const char* fsi_last_error(void) { return g_last_error.c_str(); }

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
    if (hipGetDevice(&h->device) != hipSuccess) {
        delete h;
        return fail(FSI_ERR_NO_DEVICE, "hipGetDevice failed");
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) == hipSuccess) {
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
            std::string arch = prop.gcnArchName;
            delete h;
            return fail(FSI_ERR_NO_DEVICE, "device is " + arch + ", this library is built for gfx950");
        }
    }
    if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess) {
        delete h;
        return fail(FSI_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    h->cfg = *cfg;
    apply_bounds(h, cfg->bounds_min, cfg->bounds_max, cfg->bound_damping);
    pack_constants(h);
    *out = h;
    return FSI_OK;
}

int fsi_pbf_destroy(fsi_pbf* h) {
    if (!h) return FSI_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    free_all(h);
    for (const TimedPair& t : h->pending) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    for (hipEvent_t e : h->event_pool) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return FSI_OK;
}

int fsi_pbf_set_params(fsi_pbf* h, const fsi_fluid_config* cfg) {
    int rc = bind(h);
    if (rc) return rc;
    if ((rc = check_config(cfg))) return rc;
    const uint32_t old_hash = h->cfg.hash_size;
    h->cfg = *cfg;
    h->cfg.particle_count = h->d.n;
    apply_bounds(h, cfg->bounds_min, cfg->bounds_max, cfg->bound_damping);
    pack_constants(h);
    if (h->capacity > 0 && old_hash != h->cfg.hash_size) {
        HIP_TRY(hipStreamSynchronize(h->stream));
        HIP_TRY(alloc_cells(h));
        h->stepped = false;
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
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (n != h->capacity || !h->d.pos) HIP_TRY(alloc_particles(h, n));
    h->d.n = n;
    h->cfg.particle_count = n;
    HIP_TRY(alloc_cells(h));
    pack_constants(h);
    h->stepped = false;
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
    HIP_TRY(hipMemcpyAsync(h->d.pos, p4.data(), p4.size() * sizeof(float), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->d.vel, v4.data(), v4.size() * sizeof(float), hipMemcpyHostToDevice, h->stream));
    // SolverBuffer starts zeroed (src/fluid/PBFluids.cpp:123-124)
    HIP_TRY(hipMemsetAsync(h->d.P, 0, sizeof(float4) * n, h->stream));
    HIP_TRY(hipMemsetAsync(h->d.PL, 0, sizeof(float4) * n, h->stream));
    HIP_TRY(hipMemsetAsync(h->d.curl, 0, sizeof(float4) * n, h->stream));
    HIP_TRY(hipMemsetAsync(h->d.dp_tap, 0, sizeof(float4) * n, h->stream));
    HIP_TRY(hipMemsetAsync(h->d.rho, 0, sizeof(float) * n, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return FSI_OK;
}

int fsi_pbf_particle_count(fsi_pbf* h, uint32_t* n) {
    if (!h || !n) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    *n = h->d.n;
    return FSI_OK;
}

int fsi_pbf_read_particles(fsi_pbf* h, fsi_particle* out, uint32_t cap, uint32_t* n_out) {
    int rc = bind(h);
    if (rc) return rc;
    const uint32_t n = h->d.n;
    if (n_out) *n_out = n;
    if (n == 0) return FSI_OK;
    if (!out) return fail(FSI_ERR_INVALID_ARG, "out is NULL");
    if (cap < n) return fail(FSI_ERR_CAPACITY, "cap smaller than the particle count");
    std::vector<float> p4(4 * (size_t)n), v4(4 * (size_t)n);
    HIP_TRY(hipMemcpyAsync(p4.data(), h->d.pos, p4.size() * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(v4.data(), h->d.vel, v4.size() * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
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

int fsi_pbf_set_collision_padding(fsi_pbf* h, float padding) {
    if (!h) return fail(FSI_ERR_INVALID_ARG, "handle is NULL");
    h->collision_padding = padding;
    return FSI_OK;
}

int fsi_pbf_set_bounds(fsi_pbf* h, const float mn[3], const float mx[3], float damping) {
    if (!h || !mn || !mx) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    apply_bounds(h, mn, mx, damping);
    pack_constants(h);
    return FSI_OK;
}

int fsi_pbf_set_sdf(fsi_pbf* h, const float* grid, const int32_t dims[3], const float origin[3], float cell) {
    int rc = bind(h);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (!grid) {
        dev_free(&h->sdf_grid);
        h->sdf_valid = false;
        return FSI_OK;
    }
    if (!dims || !origin) return fail(FSI_ERR_INVALID_ARG, "dims/origin is NULL");
    if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return fail(FSI_ERR_INVALID_ARG, "dims must be positive");
    if (!(cell > 0.0f)) return fail(FSI_ERR_INVALID_ARG, "cell must be positive");
    const size_t count = (size_t)dims[0] * (size_t)dims[1] * (size_t)dims[2];
    if (count > (size_t)1 << 31) return fail(FSI_ERR_INVALID_ARG, "SDF grid above 2^31 voxels");
    const bool same = h->sdf_grid && dims[0] == h->sdf_dims[0] && dims[1] == h->sdf_dims[1] && dims[2] == h->sdf_dims[2];
    if (!same) HIP_TRY(dev_alloc(&h->sdf_grid, count));
    HIP_TRY(hipMemcpy(h->sdf_grid, grid, count * sizeof(float), hipMemcpyHostToDevice));
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
    if (h->d.n == 0) return FSI_OK;  // PBFluids.cpp:325
    const StepConst c = step_const(h);
    DeviceState& d = h->d;
    d.any_sdf = c.dom.enable_sdf == 1u;
    hipStream_t st = h->stream;
    const int key_bits = log2u(h->cfg.hash_size);
    const uint32_t max_iter = h->cfg.enable_apbf ? h->cfg.max_lod : h->cfg.solver_iterations;
    for (uint32_t s = 0; s < h->cfg.substep_iterations; ++s) {
        {
            PassTimer t(h, FSI_PASS_PREDICT);
            HIP_TRY(launch_predict(c, d, st));
        }
        {
            PassTimer t(h, FSI_PASS_SORT);
            HIP_TRY(radix_sort_pairs(d.keys, d.n, key_bits, h->sort_digit_bits, d.sort, &d.skeys, &d.perm, st));
        }
        {
            PassTimer t(h, FSI_PASS_TABLES);
            HIP_TRY(launch_cells(c, d, st));
            HIP_TRY(launch_reorder(c, d, st));
        }
        if (c.enable_apbf) {
            PassTimer t(h, FSI_PASS_LOD);
            HIP_TRY(launch_lod(c, d, st));
        }
        for (uint32_t it = 1; it <= max_iter; ++it) {
            {
                PassTimer t(h, FSI_PASS_LAMBDA);
                HIP_TRY(launch_lambda(c, d, it, h->taps, st));
            }
            {
                PassTimer t(h, FSI_PASS_DELTAP);
                HIP_TRY(launch_deltap(c, d, it, h->taps, st));
            }
        }
        {
            PassTimer t(h, FSI_PASS_INTEGRATE);
            HIP_TRY(launch_integrate(c, d, st));
        }
        h->acc.substeps++;
    }
    if (c.enable_vort) {
        PassTimer t(h, FSI_PASS_VORTICITY);
        HIP_TRY(launch_vorticity(c, d, st));
    }
    h->acc.frames++;
    h->stepped = true;
    h->contacts_valid = false;
    return FSI_OK;
}

int fsi_pbf_sync(fsi_pbf* h) {
    int rc = bind(h);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return FSI_OK;
}

int fsi_pbf_filter_contacts(fsi_pbf* h, float band) {
    int rc = bind(h);
    if (rc) return rc;
    // PBFluids.cpp:195-196: no-op without an SDF, with band <= 0 or before setParticles
    if (!h->sdf_valid || !(band > 0.0f) || h->d.n == 0) return FSI_OK;
    const StepConst c = step_const(h);
    HIP_TRY(launch_contacts(c, h->d, band, h->contacts, h->contact_count, h->stream));
    h->contacts_valid = true;
    return FSI_OK;
}

int fsi_pbf_read_contacts(fsi_pbf* h, fsi_contact_candidate* out, uint32_t cap, uint32_t* count) {
    int rc = bind(h);
    if (rc) return rc;
    if (!count) return fail(FSI_ERR_INVALID_ARG, "count is NULL");
    *count = 0;
    if (!h->contacts_valid || h->d.n == 0) return FSI_OK;
    uint32_t c = 0;
    HIP_TRY(hipMemcpyAsync(&c, h->contact_count, sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    c = std::min(c, h->d.n);
    *count = c;
    const uint32_t m = std::min(c, cap);
    if (m > 0) {
        if (!out) return fail(FSI_ERR_INVALID_ARG, "out is NULL");
        HIP_TRY(hipMemcpy(out, h->contacts, sizeof(fsi_contact_candidate) * m, hipMemcpyDeviceToHost));
    }
    return FSI_OK;
}

int fsi_pbf_read_grid(fsi_pbf* h, uint32_t* keys, uint32_t* ids, int32_t* cells) {
    int rc = bind(h);
    if (rc) return rc;
    const uint32_t n = h->d.n;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (keys && n) {
        if (h->stepped) HIP_TRY(hipMemcpy(keys, h->d.skeys, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        else std::memset(keys, 0, sizeof(uint32_t) * n);
    }
    if (ids && n) {
        if (h->stepped) HIP_TRY(hipMemcpy(ids, h->d.perm, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        else std::memset(ids, 0, sizeof(uint32_t) * n);
    }
    if (cells && h->d.cells)
        HIP_TRY(hipMemcpy(cells, h->d.cells, sizeof(int2) * (size_t)h->d.hash_size, hipMemcpyDeviceToHost));
    return FSI_OK;
}

int fsi_pbf_read_solver(fsi_pbf* h, float* ppl, float* dpr) {
    int rc = bind(h);
    if (rc) return rc;
    const uint32_t n = h->d.n;
    if (n == 0) return FSI_OK;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (!h->stepped) {
        if (ppl) std::memset(ppl, 0, sizeof(float) * 4 * n);
        if (dpr) std::memset(dpr, 0, sizeof(float) * 4 * n);
        return FSI_OK;
    }
    std::vector<uint32_t> perm(n);
    HIP_TRY(hipMemcpy(perm.data(), h->d.perm, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    std::vector<float> tmp(4 * (size_t)n);
    if (ppl) {
        HIP_TRY(hipMemcpy(tmp.data(), h->d.P, sizeof(float4) * n, hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < n; ++k) tmp[4 * (size_t)k + 3] = 0.0f;  // Integrate resets lambda
        unpermute(perm, tmp.data(), ppl, 4);
    }
    if (dpr) {
        const bool vort = h->cfg.enable_vorticity != 0;
        HIP_TRY(hipMemcpy(tmp.data(), vort ? h->d.curl : h->d.dp_tap, sizeof(float4) * n, hipMemcpyDeviceToHost));
        std::vector<float> rho(n);
        HIP_TRY(hipMemcpy(rho.data(), h->d.rho, sizeof(float) * n, hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < n; ++k) tmp[4 * (size_t)k + 3] = rho[k];
        unpermute(perm, tmp.data(), dpr, 4);
    }
    return FSI_OK;
}

int fsi_pbf_set_taps(fsi_pbf* h, int enable) {
    if (!h) return fail(FSI_ERR_INVALID_ARG, "handle is NULL");
    h->taps = enable != 0;
    return FSI_OK;
}

int fsi_pbf_read_counts(fsi_pbf* h, uint32_t* cand, uint32_t* nbr) {
    int rc = bind(h);
    if (rc) return rc;
    if (!h->taps) return fail(FSI_ERR_STATE, "taps are off (fsi_pbf_set_taps)");
    const uint32_t n = h->d.n;
    if (n == 0) return FSI_OK;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (!h->stepped) return fail(FSI_ERR_STATE, "no step has run");
    std::vector<uint32_t> perm(n), tmp(n);
    HIP_TRY(hipMemcpy(perm.data(), h->d.perm, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    if (cand) {
        HIP_TRY(hipMemcpy(tmp.data(), h->d.cand, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < n; ++k) cand[perm[k]] = tmp[k];
    }
    if (nbr) {
        HIP_TRY(hipMemcpy(tmp.data(), h->d.nbr, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < n; ++k) nbr[perm[k]] = tmp[k];
    }
    return FSI_OK;
}

int fsi_pbf_set_timing(fsi_pbf* h, int enable) {
    if (!h) return fail(FSI_ERR_INVALID_ARG, "handle is NULL");
    h->timing = enable != 0;
    return FSI_OK;
}

int fsi_pbf_timing_reset(fsi_pbf* h) {
    int rc = bind(h);
    if (rc) return rc;
    if ((rc = drain_timing(h))) return rc;
    std::memset(&h->acc, 0, sizeof(h->acc));
    return FSI_OK;
}

int fsi_pbf_read_timing(fsi_pbf* h, fsi_pbf_timing* out) {
    int rc = bind(h);
    if (rc) return rc;
    if (!out) return fail(FSI_ERR_INVALID_ARG, "out is NULL");
    if ((rc = drain_timing(h))) return rc;
    *out = h->acc;
    return FSI_OK;
}

int fsi_pbf_set_tuning(fsi_pbf* h, int knob, int value) {
    if (!h) return fail(FSI_ERR_INVALID_ARG, "handle is NULL");
    switch (knob) {
        case FSI_TUNE_GATHER:
            if (value < 0 || value > 1) return fail(FSI_ERR_INVALID_ARG, "gather form must be 0 or 1");
            h->d.variant = value;
            return FSI_OK;
        case FSI_TUNE_SORT_DIGIT_BITS:
            if (value < 1 || value > kSortMaxDigitBits) return fail(FSI_ERR_INVALID_ARG, "digit bits must be 1..10");
            h->sort_digit_bits = value;
            return FSI_OK;
        default:
            return fail(FSI_ERR_INVALID_ARG, "unknown knob");
    }
}

int fsi_pbf_get_tuning(fsi_pbf* h, int knob, int* value) {
    if (!h || !value) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    switch (knob) {
        case FSI_TUNE_GATHER: *value = h->d.variant; return FSI_OK;
        case FSI_TUNE_SORT_DIGIT_BITS: *value = h->sort_digit_bits; return FSI_OK;
        default: return fail(FSI_ERR_INVALID_ARG, "unknown knob");
    }
}

int fsi_pbf_get_stream(fsi_pbf* h, void** stream) {
    if (!h || !stream) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    *stream = (void*)h->stream;
    return FSI_OK;
}

}  // extern "C"
