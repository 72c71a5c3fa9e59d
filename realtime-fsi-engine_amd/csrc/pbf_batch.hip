// pbf_batch.hip — multi-scene solver handle (fsi_batch_* in include/fsi_pbf.h).
//
// Many independent scenes share one set of device buffers: scene s owns the
// particle ids [offset_s, offset_s + count_s) and the slice
// [hash_base_s, hash_base_s + hashSize_s) of the cell table.  Predict writes
// hash_base_s + hash(cell) as the key, so the one stable sort of the whole
// batch keeps every scene's particles in their own id range, each scene's
// order being exactly that of a single-scene handle, and one launch per pass
// (the kernels of pbf_kernels.hip in their multi-scene form) advances every
// scene.  Every scene follows PBFluids::step (src/fluid/PBFluids.cpp:323-517);
// the reference itself runs one scene per PBFluids.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/fsi_pbf.h"
#include "pbf_core.h"

using namespace fsi;

namespace {
struct SceneHost {
    uint32_t offset = 0, count = 0, hash_base = 0, hash_size = 0;
    float* sdf_grid = nullptr;  // device copy (owned)
    float2* sdf_brange = nullptr;  // its block ranges (owned)
};
}  // namespace

struct fsi_batch {
    Core core;
    fsi_fluid_config cfg{};
    fsi_pbf_constants k{};
    std::vector<SceneHost> scenes;
    std::vector<Domain> doms;  // host copy of the device table
    Domain* d_doms = nullptr;
    uint16_t* d_scene_of = nullptr;
    uint32_t total_hash = 0;
    uint32_t axis_bits = 1;  // Morton bits per axis of every scene's bins
    bool binned_tiles = false;  // doms are binned for the tiled gathers
    bool any_sdf = false;
    std::vector<float> host_pred;
};

namespace {

int bind(fsi_batch* b) {
    if (!b) return fail(FSI_ERR_INVALID_ARG, "handle is NULL");
    FSI_HIP_TRY(hipSetDevice(b->core.device));
    return FSI_OK;
}

void free_scenes(fsi_batch* b) {
    for (SceneHost& s : b->scenes) {
        dev_free(&s.sdf_grid);
        dev_free(&s.sdf_brange);
    }
    b->scenes.clear();
    b->doms.clear();
    dev_free(&b->d_doms);
    dev_free(&b->d_scene_of);
}

StepConst step_const(const fsi_batch* b) {
    const float zero[3] = {0, 0, 0};
    StepConst c = step_const_common(b->cfg, b->k, b->core.d.n, false, zero, zero);
    if (!b->doms.empty()) c.dom = b->doms[0];
    c.doms = b->d_doms;
    c.scene_of = b->d_scene_of;
    return c;
}

FramePlan frame_plan(const fsi_batch* b) {
    FramePlan p;
    p.substeps = b->cfg.substep_iterations;
    p.max_iter = b->cfg.enable_apbf ? b->cfg.max_lod : b->cfg.solver_iterations;
    p.key_bits = log2u(b->total_hash);
    p.vorticity = b->cfg.enable_vorticity != 0;
    return p;
}

void prepare(fsi_batch* b) {
    b->core.d.batched = true;
    b->core.d.any_sdf = b->any_sdf;
    bool tiled = true;
    for (const Domain& d : b->doms) tiled = tiled && d.dshift == kTileShift;
    b->core.d.tiled = tiled;
}

}  // namespace

extern "C" {

int fsi_batch_create(const fsi_fluid_config* cfg, fsi_batch** out) {
    if (!out) return fail(FSI_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    int rc = check_config(cfg);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(FSI_ERR_NO_DEVICE, "no HIP device visible");
    fsi_batch* b = new (std::nothrow) fsi_batch();
    if (!b) return fail(FSI_ERR_INVALID_ARG, "out of host memory");
    if ((rc = core_init(b->core))) {
        core_destroy(b->core);
        delete b;
        return rc;
    }
    b->core.d.relabel = true;  // storage in Morton order of the buckets (processing order 2)
    b->cfg = *cfg;
    const float zero[3] = {0, 0, 0};
    pack_constants(b->cfg, zero, zero, std::clamp(cfg->bound_damping, 0.0f, 1.0f), 0, b->k);
    *out = b;
    return FSI_OK;
}

int fsi_batch_destroy(fsi_batch* b) {
    if (!b) return FSI_OK;
    (void)hipSetDevice(b->core.device);
    if (b->core.stream) (void)hipStreamSynchronize(b->core.stream);
    free_scenes(b);
    core_destroy(b->core);
    delete b;
    return FSI_OK;
}

int fsi_batch_set_scenes(fsi_batch* b, const fsi_particle* ps, uint32_t total, const fsi_scene_desc* sc,
                         uint32_t nscenes) {
    int rc = bind(b);
    if (rc) return rc;
    if (nscenes == 0 || nscenes > 65535u) return fail(FSI_ERR_INVALID_ARG, "nscenes must be 1..65535");
    if (!sc || (total > 0 && !ps)) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    if (total > (1u << 30)) return fail(FSI_ERR_INVALID_ARG, "more than 2^30 particles");
    uint64_t off = 0, hash = 0;
    for (uint32_t s = 0; s < nscenes; ++s) {
        if (sc[s].particle_offset != off) return fail(FSI_ERR_INVALID_ARG, "scenes must tile the particles in order");
        if (!is_pow2(sc[s].hash_size)) return fail(FSI_ERR_INVALID_ARG, "scene hash_size must be a power of two");
        if (sc[s].sdf_grid) {
            if (sc[s].sdf_dims[0] <= 0 || sc[s].sdf_dims[1] <= 0 || sc[s].sdf_dims[2] <= 0 || !(sc[s].sdf_cell > 0.0f))
                return fail(FSI_ERR_INVALID_ARG, "scene SDF dims/cell must be positive");
        }
        off += sc[s].particle_count;
        hash += sc[s].hash_size;
    }
    if (off != total) return fail(FSI_ERR_INVALID_ARG, "scene counts do not sum to the particle total");
    if (hash > (1u << 30)) return fail(FSI_ERR_INVALID_ARG, "total hash size above 2^30");
    Core& c = b->core;
    FSI_HIP_TRY(hipStreamSynchronize(c.stream));
    free_scenes(b);
    // Morton bins: split the 2^21 bins evenly, 3 * axis_bits per scene
    b->binned_tiles = c.d.variant == 2;
    int sbits = 0;
    while ((1u << sbits) < nscenes) ++sbits;
    const uint32_t axis_bits = (uint32_t)std::max(1, std::min(7, (21 - sbits) / 3));
    b->axis_bits = axis_bits;
    b->scenes.resize(nscenes);
    b->doms.resize(nscenes);
    b->any_sdf = false;
    uint32_t hbase = 0;
    for (uint32_t s = 0; s < nscenes; ++s) {
        SceneHost& h = b->scenes[s];
        h.offset = sc[s].particle_offset;
        h.count = sc[s].particle_count;
        h.hash_base = hbase;
        h.hash_size = sc[s].hash_size;
        hbase += sc[s].hash_size;
        Domain& d = b->doms[s];
        std::memset(&d, 0, sizeof(d));
        for (int a = 0; a < 3; ++a) {
            d.bmin[a] = sc[s].bounds_min[a];
            d.bmax[a] = sc[s].bounds_max[a];
        }
        d.hash_mask = h.hash_size - 1u;
        d.hash_base = h.hash_base;
        set_binning(d, d.bmin, d.bmax, b->cfg.h, axis_bits, s << (3 * axis_bits), b->binned_tiles);
        if (sc[s].sdf_grid) {
            const size_t cnt = (size_t)sc[s].sdf_dims[0] * sc[s].sdf_dims[1] * sc[s].sdf_dims[2];
            FSI_HIP_TRY(dev_alloc(&h.sdf_grid, cnt));
            FSI_HIP_TRY(hipMemcpy(h.sdf_grid, sc[s].sdf_grid, cnt * sizeof(float), hipMemcpyHostToDevice));
            FSI_HIP_TRY(dev_alloc(&h.sdf_brange, sdf_block_count(sc[s].sdf_dims)));
            FSI_HIP_TRY(launch_sdf_block_range(h.sdf_grid, sc[s].sdf_dims, h.sdf_brange, b->core.stream));
            d.enable_sdf = 1u;
            d.sdf_mode = sc[s].sdf_containment ? 1u : 0u;
            for (int a = 0; a < 3; ++a) {
                d.dims[a] = sc[s].sdf_dims[a];
                d.sdf_origin[a] = sc[s].sdf_origin[a];
            }
            d.sdf_cell = sc[s].sdf_cell;
            d.sdf_pad = sc[s].sdf_padding;
            d.sdf_grid = h.sdf_grid;
            d.sdf_brange = h.sdf_brange;
            for (int a = 0; a < 3; ++a) d.sdf_bdims[a] = sdf_blocks(sc[s].sdf_dims[a]);
            b->any_sdf = true;
        }
    }
    b->total_hash = hbase;
    FSI_HIP_TRY(dev_alloc(&b->d_doms, nscenes));
    FSI_HIP_TRY(hipMemcpy(b->d_doms, b->doms.data(), sizeof(Domain) * nscenes, hipMemcpyHostToDevice));
    std::vector<uint16_t> scene_of(total);
    for (uint32_t s = 0; s < nscenes; ++s)
        std::fill(scene_of.begin() + sc[s].particle_offset, scene_of.begin() + sc[s].particle_offset + sc[s].particle_count,
                  (uint16_t)s);
    FSI_HIP_TRY(dev_alloc(&b->d_scene_of, std::max<uint32_t>(total, 1)));
    if (total) FSI_HIP_TRY(hipMemcpy(b->d_scene_of, scene_of.data(), sizeof(uint16_t) * total, hipMemcpyHostToDevice));

    if (total != c.capacity || !c.d.pos) FSI_HIP_TRY(core_alloc_particles(c, total));
    c.d.n = total;
    b->cfg.particle_count = total;
    FSI_HIP_TRY(core_alloc_cells(c, b->total_hash));
    c.stepped = false;
    core_invalidate_graph(c);
    b->host_pred.assign(4 * (size_t)total, 0.0f);
    if (total == 0) return FSI_OK;
    std::vector<float> p4(4 * (size_t)total), v4(4 * (size_t)total);
    for (uint32_t i = 0; i < total; ++i) {
        for (int a = 0; a < 3; ++a) {
            p4[4 * (size_t)i + a] = ps[i].pos[a];
            v4[4 * (size_t)i + a] = ps[i].vel[a];
            b->host_pred[4 * (size_t)i + a] = ps[i].pred_pos[a];
        }
        p4[4 * (size_t)i + 3] = ps[i].pad0;
        v4[4 * (size_t)i + 3] = ps[i].pad2;
        b->host_pred[4 * (size_t)i + 3] = ps[i].pad1;
    }
    FSI_HIP_TRY(hipMemcpyAsync(c.d.pos, p4.data(), p4.size() * sizeof(float), hipMemcpyHostToDevice, c.stream));
    FSI_HIP_TRY(hipMemcpyAsync(c.d.vel, v4.data(), v4.size() * sizeof(float), hipMemcpyHostToDevice, c.stream));
    FSI_HIP_TRY(hipMemsetAsync(c.d.P, 0, sizeof(float4) * total, c.stream));
    FSI_HIP_TRY(hipMemsetAsync(c.d.PL, 0, sizeof(float4) * total, c.stream));
    FSI_HIP_TRY(hipMemsetAsync(c.d.curl, 0, sizeof(float4) * total, c.stream));
    FSI_HIP_TRY(hipMemsetAsync(c.d.rho, 0, sizeof(float) * total, c.stream));
    FSI_HIP_TRY(hipStreamSynchronize(c.stream));
    return FSI_OK;
}

int fsi_batch_step(fsi_batch* b) {
    int rc = bind(b);
    if (rc) return rc;
    if (b->core.d.n == 0) return FSI_OK;
    prepare(b);
    return core_frame(b->core, step_const(b), frame_plan(b));
}

int fsi_batch_sync(fsi_batch* b) {
    int rc = bind(b);
    if (rc) return rc;
    FSI_HIP_TRY(hipStreamSynchronize(b->core.stream));
    return FSI_OK;
}

int fsi_batch_read_particles(fsi_batch* b, fsi_particle* out, uint32_t cap, uint32_t* n_out) {
    int rc = bind(b);
    if (rc) return rc;
    const uint32_t n = b->core.d.n;
    if (n_out) *n_out = n;
    if (n == 0) return FSI_OK;
    if (!out) return fail(FSI_ERR_INVALID_ARG, "out is NULL");
    if (cap < n) return fail(FSI_ERR_CAPACITY, "cap smaller than the particle count");
    std::vector<float> p4(4 * (size_t)n), v4(4 * (size_t)n);
    hipStream_t st = b->core.stream;
    FSI_HIP_TRY(hipMemcpyAsync(p4.data(), b->core.d.pos, p4.size() * sizeof(float), hipMemcpyDeviceToHost, st));
    FSI_HIP_TRY(hipMemcpyAsync(v4.data(), b->core.d.vel, v4.size() * sizeof(float), hipMemcpyDeviceToHost, st));
    FSI_HIP_TRY(hipStreamSynchronize(st));
    for (uint32_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            out[i].pos[a] = p4[4 * (size_t)i + a];
            out[i].vel[a] = v4[4 * (size_t)i + a];
            out[i].pred_pos[a] = b->host_pred[4 * (size_t)i + a];
        }
        out[i].pad0 = p4[4 * (size_t)i + 3];
        out[i].pad1 = b->host_pred[4 * (size_t)i + 3];
        out[i].pad2 = v4[4 * (size_t)i + 3];
    }
    return FSI_OK;
}

int fsi_batch_read_grid(fsi_batch* b, uint32_t scene, uint32_t* keys, uint32_t* ids, int32_t* cells) {
    int rc = bind(b);
    if (rc) return rc;
    if (scene >= b->scenes.size()) return fail(FSI_ERR_INVALID_ARG, "scene out of range");
    const SceneHost& s = b->scenes[scene];
    const Core& c = b->core;
    FSI_HIP_TRY(hipStreamSynchronize(c.stream));
    if (!c.stepped) return fail(FSI_ERR_STATE, "no step has run");
    if (keys && s.count) {
        FSI_HIP_TRY(hipMemcpy(keys, c.d.skeys + s.offset, sizeof(uint32_t) * s.count, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < s.count; ++i) keys[i] -= s.hash_base;
    }
    if (ids && s.count) {
        FSI_HIP_TRY(hipMemcpy(ids, c.d.perm + s.offset, sizeof(uint32_t) * s.count, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < s.count; ++i) ids[i] -= s.offset;
    }
    if (cells) {
        FSI_HIP_TRY(hipMemcpy(cells, c.d.cells + s.hash_base, sizeof(int2) * (size_t)s.hash_size,
                              hipMemcpyDeviceToHost));
        for (size_t i = 0; i < 2 * (size_t)s.hash_size; ++i)
            if (cells[i] != -1) cells[i] -= (int32_t)s.offset;
    }
    return FSI_OK;
}

int fsi_batch_set_graph(fsi_batch* b, int enable) {
    int rc = bind(b);
    if (rc) return rc;
    FSI_HIP_TRY(hipStreamSynchronize(b->core.stream));
    b->core.use_graph = enable != 0;
    if (!b->core.use_graph) core_invalidate_graph(b->core);
    return FSI_OK;
}

// Re-derives every scene's binning for the tiled gathers (bins of
// 2^kTileShift cells) or for the other forms, when it changes.
static int rebin_scenes(fsi_batch* b, bool tiles) {
    if (b->doms.empty() || b->binned_tiles == tiles) return FSI_OK;
    FSI_HIP_TRY(hipStreamSynchronize(b->core.stream));
    for (size_t s = 0; s < b->doms.size(); ++s) {
        Domain& d = b->doms[s];
        set_binning(d, d.bmin, d.bmax, b->cfg.h, b->axis_bits, (uint32_t)s << (3 * b->axis_bits), tiles);
    }
    FSI_HIP_TRY(hipMemcpy(b->d_doms, b->doms.data(), sizeof(Domain) * b->doms.size(), hipMemcpyHostToDevice));
    b->binned_tiles = tiles;
    core_invalidate_graph(b->core);
    return FSI_OK;
}

int fsi_batch_set_tuning(fsi_batch* b, int knob, int value) {
    int rc = bind(b);
    if (rc) return rc;
    if ((rc = core_set_tuning(b->core, knob, value))) return rc;
    return rebin_scenes(b, b->core.d.variant == 2);
}

int fsi_batch_get_tuning(fsi_batch* b, int knob, int* value) {
    if (!b || !value) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    return core_get_tuning(b->core, knob, value);
}

int fsi_batch_autotune(fsi_batch* b, uint32_t frames, int* gather, int* digit_bits) {
    int rc = bind(b);
    if (rc) return rc;
    const Rebin rebin = [b](bool tiles, StepConst& k) {
        const int rc = rebin_scenes(b, tiles);
        prepare(b);
        k = step_const(b);
        return rc;
    };
    return core_autotune(b->core, rebin, frame_plan(b), frames, gather, digit_bits);
}

int fsi_batch_read_autotune(fsi_batch* b, fsi_tune_trial* out, uint32_t cap, uint32_t* n) {
    if (!b) return fail(FSI_ERR_INVALID_ARG, "handle is NULL");
    return core_read_autotune(b->core, out, cap, n);
}

int fsi_batch_set_timing(fsi_batch* b, int enable) {
    if (!b) return fail(FSI_ERR_INVALID_ARG, "handle is NULL");
    b->core.timing = enable != 0;
    b->core.timing_mask = 0xffu;
    return FSI_OK;
}

int fsi_batch_set_timing_passes(fsi_batch* b, uint32_t mask) {
    if (!b) return fail(FSI_ERR_INVALID_ARG, "handle is NULL");
    b->core.timing = mask != 0;
    b->core.timing_mask = mask;
    return FSI_OK;
}

int fsi_batch_timing_reset(fsi_batch* b) {
    int rc = bind(b);
    if (rc) return rc;
    if ((rc = core_drain_timing(b->core))) return rc;
    std::memset(&b->core.acc, 0, sizeof(b->core.acc));
    return FSI_OK;
}

int fsi_batch_read_timing(fsi_batch* b, fsi_pbf_timing* out) {
    int rc = bind(b);
    if (rc) return rc;
    if (!out) return fail(FSI_ERR_INVALID_ARG, "out is NULL");
    if ((rc = core_drain_timing(b->core))) return rc;
    *out = b->core.acc;
    return FSI_OK;
}

}  // extern "C"
