// pbf_core.hip — buffers, frame sequence, timing, tuning and graph capture
// shared by the single-scene and multi-scene solver handles.
#include "pbf_core.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace fsi {

namespace {
thread_local std::string g_last_error;
}

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

const char* last_error() { return g_last_error.c_str(); }

bool is_pow2(uint32_t x) { return x != 0 && (x & (x - 1)) == 0; }

int log2u(uint32_t x) {
    int b = 0;
    while ((1ull << b) < (unsigned long long)x) ++b;
    return b;
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

// calcPoly6Kernel of src/fluid/Kernels.cpp:9-27 (reached through
// calcSCorrKernel, :216-219) in FP32: the s_corr reference value wq.
static float wq_of(float dq, float h) {
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

void pack_constants(const fsi_fluid_config& p, const float bmin[3], const float bmax[3], float damping, uint32_t n,
                    fsi_pbf_constants& u) {
    std::memset(&u, 0, sizeof(u));
    for (int a = 0; a < 3; ++a) {
        u.bounds_min[a] = bmin[a];
        u.bounds_max[a] = bmax[a];
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
    u.bound_damping = damping;
    u.hash_size = p.hash_size;
    u.particle_count = n;
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

void set_binning(Domain& d, const float bmin[3], const float bmax[3], float h, uint32_t axis_bits,
                 uint32_t bin_base, bool tiles) {
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
    d.bin_base = bin_base;
    if (tiles) {
        // bins of 2^kTileShift cells (the tiles of the tiled gathers) when
        // axis_bits bins per axis can cover the box with them
        uint32_t bits = 1;
        while (bits < axis_bits && span > ((int64_t)1 << (bits + kTileShift))) ++bits;
        if (span <= ((int64_t)1 << (bits + kTileShift))) {
            d.dshift = kTileShift;
            d.axis_bits = bits;
            return;
        }
    }
    uint32_t shift = 0;
    while ((span >> shift) > ((int64_t)1 << axis_bits) && shift < 31) ++shift;
    d.dshift = shift;
    d.axis_bits = axis_bits;
}

// Whether the operands of the pair divisions lie where div_rn_normal returns
// a / b (pbf_device.h): the spiky gradient's -spiky diff^2 / dist (dist in
// [1e-5, h], diff = h - dist at least half an ulp of h when positive) and
// s_corr's w / wq (w = poly6 (h^2 - r^2)^3, h^2 - r^2 at least half an ulp of
// h^2 when positive).  Numerators must lie in [2^-80, 2^60], denominators in
// [2^-40, 2^40] and quotients in [2^-100, 2^80]: inside the range on which
// the test library compares div_rn_normal with a / b, and far inside the one
// on which the two are the same instructions.
static bool pair_divisions_in_range(const StepConst& c) {
    const double alo = std::ldexp(1.0, -80), ahi = std::ldexp(1.0, 60);
    const double blo = std::ldexp(1.0, -40), bhi = std::ldexp(1.0, 40);
    const double qlo = std::ldexp(1.0, -100), qhi = std::ldexp(1.0, 80);
    auto in = [](double v, double lo, double hi) { return std::isfinite(v) && v >= lo && v <= hi; };
    const double h = c.h;
    if (!(h > 0.0) || !std::isfinite(h)) return false;
    const double uh = std::ldexp(1.0, std::ilogb(h) - 24), uh2 = std::ldexp(1.0, std::ilogb(h * h) - 24);
    const double K = std::fabs((double)c.spiky), amax = K * h * h, amin = K * uh * uh;
    bool ok = in(amin, alo, ahi) && in(amax, alo, ahi) && in(1e-5, blo, bhi) && in(h, blo, bhi) &&
              in(amin / h, qlo, qhi) && in(amax / 1e-5, qlo, qhi);
    if (c.enable_scorr != 0u) {
        const double P = std::fabs((double)c.poly6), wq = c.wq, h2 = h * h;
        const double wmax = P * h2 * h2 * h2, wmin = P * uh2 * uh2 * uh2;
        ok = ok && in(wq, blo, bhi) && in(wmin, alo, ahi) && in(wmax, alo, ahi) && in(wmin / wq, qlo, qhi) &&
             in(wmax / wq, qlo, qhi);
    }
    return ok;
}

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
    c.qscale = kQUnitsPerH / u.h;
    c.fast_div = pair_divisions_in_range(c) ? 1u : 0u;
    return c;
}

int core_init(Core& c) {
    FSI_HIP_TRY(hipGetDevice(&c.device));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c.device) == hipSuccess && std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(FSI_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950");
    FSI_HIP_TRY(hipStreamCreateWithFlags(&c.own_stream, hipStreamNonBlocking));
    c.stream = c.own_stream;
    return FSI_OK;
}

int core_set_stream(Core& c, hipStream_t s) {
    if (!s) s = c.own_stream;
    if (s == c.stream) return FSI_OK;
    hipDevice_t dev = -1;
    FSI_HIP_TRY(hipStreamGetDevice(s, &dev));
    if (dev != c.device) return fail(FSI_ERR_INVALID_ARG, "the stream belongs to another device");
    int rc = core_drain_timing(c);
    if (rc) return rc;
    FSI_HIP_TRY(hipStreamSynchronize(c.stream));
    c.stream = s;
    return FSI_OK;
}

void core_invalidate_graph(Core& c) {
    if (c.exec) (void)hipGraphExecDestroy(c.exec);
    if (c.graph) (void)hipGraphDestroy(c.graph);
    c.exec = nullptr;
    c.graph = nullptr;
    c.graph_valid = false;
    c.graph_key.clear();
}

void core_destroy(Core& c) {
    (void)hipSetDevice(c.device);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    core_invalidate_graph(c);
    DeviceState& d = c.d;
    dev_free(&d.pos);
    dev_free(&d.vel);
    dev_free(&d.xstar);
    dev_free(&d.keys);
    dev_free(&d.cells);
    dev_free(&d.mcells);
    dev_free(&d.P);
    dev_free(&d.Q);
    dev_free(&d.PL);
    dev_free(&d.vs);
    dev_free(&d.curl);
    dev_free(&d.posS);
    dev_free(&d.vdelta);
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
    dev_free(&d.bin_m);
    dev_free(&d.mperm);
    dev_free(&d.order);
    dev_free(&d.list_count);
    dev_free(&d.list);
    dev_free(&d.tflag);
    dev_free(&d.tidx);
    dev_free(&d.tiles);
    dev_free(&d.ntiles);
    dev_free(&d.defer);
    dev_free(&d.ndefer);
    dev_free(&d.sort.keys[0]);
    dev_free(&d.sort.keys[1]);
    dev_free(&d.sort.ids[0]);
    dev_free(&d.sort.ids[1]);
    dev_free(&d.sort.hist);
    dev_free(&d.sort.partial);
    for (const TimedPair& t : c.pending) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    c.pending.clear();
    for (hipEvent_t e : c.event_pool) (void)hipEventDestroy(e);
    c.event_pool.clear();
    if (c.own_stream) (void)hipStreamDestroy(c.own_stream);
    c.own_stream = nullptr;
    c.stream = nullptr;
}

hipError_t core_alloc_cells(Core& c, uint32_t entries) {
    if (c.cells_capacity != entries || !c.d.cells || !c.d.mcells) {
        hipError_t e = dev_alloc(&c.d.cells, entries);
        if (e == hipSuccess) e = dev_alloc(&c.d.mcells, entries);
        if (e != hipSuccess) return e;
        c.cells_capacity = entries;
        core_invalidate_graph(c);
    }
    c.d.hash_size = entries;
    std::vector<int32_t> neg(2 * (size_t)entries, -1);
    hipError_t e = hipMemcpyAsync(c.d.cells, neg.data(), neg.size() * sizeof(int32_t), hipMemcpyHostToDevice, c.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c.d.mcells, neg.data(), neg.size() * sizeof(int32_t), hipMemcpyHostToDevice, c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);  // neg is freed on return
    return e;
}

hipError_t core_alloc_particles(Core& c, uint32_t n) {
    DeviceState& d = c.d;
    hipError_t e = hipSuccess;
#define ALLOC(ptr, count) \
    if ((e = dev_alloc(&(ptr), (count))) != hipSuccess) return e;
    ALLOC(d.pos, n);
    ALLOC(d.vel, n);
    ALLOC(d.xstar, n);
    ALLOC(d.keys, n);
    ALLOC(d.P, n);
    ALLOC(d.Q, (size_t)n + 4);  // a group of four candidate loads may reach 3 entries past the last
    ALLOC(d.PL, n);
    ALLOC(d.vs, n);
    ALLOC(d.curl, n);
    ALLOC(d.posS, n);
    ALLOC(d.vdelta, n);
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
    ALLOC(d.bin_m, n);
    ALLOC(d.mperm, n);
    ALLOC(d.order, n);
    ALLOC(d.list_count, n);
    ALLOC(d.list, (size_t)kNbrCap * (((size_t)n + 63) & ~(size_t)63));  // whole 64-thread blocks
    ALLOC(d.tflag, n);
    ALLOC(d.tidx, n);
    ALLOC(d.tiles, (size_t)n + 1);
    ALLOC(d.ntiles, 1);
    ALLOC(d.defer, n);
    ALLOC(d.ndefer, kDeferWords);
    ALLOC(d.sort.keys[0], n);
    ALLOC(d.sort.keys[1], n);
    ALLOC(d.sort.ids[0], n);
    ALLOC(d.sort.ids[1], n);
    const size_t hist = sort_hist_entries(n, kSortMaxDigitBits);
    ALLOC(d.sort.hist, hist);
    const size_t parts =
        std::max({scan_partials_needed(hist), scan_partials_needed(n), scan_partials_needed(kOrderBins)});
    ALLOC(d.sort.partial, parts);
#undef ALLOC
    c.capacity = n;
    d.n = n;
    core_invalidate_graph(c);
    return hipSuccess;
}

namespace {

hipEvent_t take_event(Core& c) {
    if (!c.event_pool.empty()) {
        hipEvent_t e = c.event_pool.back();
        c.event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Records a HIP event pair around the launches of one pass (on the handle's
// stream) while timing is on and no graph is being captured.
struct PassTimer {
    Core& c;
    bool on;
    int pass;
    hipEvent_t a = nullptr, b = nullptr;
    PassTimer(Core& cc, bool enabled, int p) : c(cc), on(enabled && ((cc.timing_mask >> p) & 1u)), pass(p) {
        if (!on) return;
        a = take_event(c);
        b = take_event(c);
        if (a && b) (void)hipEventRecord(a, c.stream);
    }
    ~PassTimer() {
        if (!on || !a || !b) return;
        (void)hipEventRecord(b, c.stream);
        c.pending.push_back(TimedPair{pass, a, b});
    }
};

// One frame of PBFluids::step (src/fluid/PBFluids.cpp:344-505).
int enqueue_frame(Core& c, const StepConst& k, const FramePlan& plan, bool timed) {
    DeviceState& d = c.d;
    hipStream_t st = c.stream;
    for (uint32_t s = 0; s < plan.substeps; ++s) {
        {
            PassTimer t(c, timed, FSI_PASS_PREDICT);
            FSI_HIP_TRY(launch_predict(k, d, st));
        }
        {
            PassTimer t(c, timed, FSI_PASS_SORT);
            FSI_HIP_TRY(radix_sort_pairs(d.keys, d.n, plan.key_bits, c.sort_digit_bits, d.sort, &d.skeys, &d.perm, st));
            if (c.record_perm)
                FSI_HIP_TRY(hipMemcpyAsync(c.record_perm + (size_t)s * d.n, d.perm, sizeof(uint32_t) * d.n,
                                           hipMemcpyDeviceToDevice, st));
        }
        {
            PassTimer t(c, timed, FSI_PASS_TABLES);
            FSI_HIP_TRY(launch_cells(k, d, st));
            FSI_HIP_TRY(launch_reorder(k, d, st));
        }
        if (k.enable_apbf) {
            PassTimer t(c, timed, FSI_PASS_LOD);
            FSI_HIP_TRY(launch_lod(k, d, st));
        }
        for (uint32_t it = 1; it <= plan.max_iter; ++it) {
            {
                PassTimer t(c, timed, FSI_PASS_LAMBDA);
                FSI_HIP_TRY(launch_lambda(k, d, it, c.taps, st));
            }
            {
                PassTimer t(c, timed, FSI_PASS_DELTAP);
                FSI_HIP_TRY(launch_deltap(k, d, it, c.taps, st));
            }
        }
        {
            PassTimer t(c, timed, FSI_PASS_INTEGRATE);
            FSI_HIP_TRY(launch_integrate(k, d, st));
        }
    }
    if (plan.vorticity) {
        PassTimer t(c, timed, FSI_PASS_VORTICITY);
        FSI_HIP_TRY(launch_vorticity(k, d, st));
    }
    return FSI_OK;
}

}  // namespace

int core_frame(Core& c, const StepConst& k, const FramePlan& plan) {
    if (c.d.n == 0) return FSI_OK;
    int rc = FSI_OK;
    if (!c.use_graph) {
        rc = enqueue_frame(c, k, plan, c.timing);
    } else {
        // key: everything the launches of a frame depend on
        const int knobs[8] = {c.d.variant, c.sort_digit_bits, c.taps ? 1 : 0, c.d.any_sdf ? 1 : 0,
                              c.d.morton ? 1 : 0, c.d.relabel ? 1 : 0, c.d.block, c.d.tile_shape};
        std::vector<unsigned char> key(sizeof(StepConst) + sizeof(FramePlan) + sizeof(knobs) + sizeof(c.record_perm));
        std::memcpy(key.data(), &k, sizeof(StepConst));
        std::memcpy(key.data() + sizeof(StepConst), &plan, sizeof(FramePlan));
        std::memcpy(key.data() + sizeof(StepConst) + sizeof(FramePlan), knobs, sizeof(knobs));
        std::memcpy(key.data() + sizeof(StepConst) + sizeof(FramePlan) + sizeof(knobs), &c.record_perm,
                    sizeof(c.record_perm));
        if (!c.graph_valid || key != c.graph_key) {
            core_invalidate_graph(c);
            FSI_HIP_TRY(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue_frame(c, k, plan, false);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(c.stream, &g);
            if (rc != FSI_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return fail(FSI_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            c.graph = g;
            FSI_HIP_TRY(hipGraphInstantiate(&c.exec, c.graph, nullptr, nullptr, 0));
            c.graph_valid = true;
            c.graph_key = key;
        }
        FSI_HIP_TRY(hipGraphLaunch(c.exec, c.stream));
    }
    if (rc != FSI_OK) return rc;
    c.acc.substeps += plan.substeps;
    c.acc.frames++;
    c.stepped = true;
    return FSI_OK;
}

namespace {
// What core_autotune holds while it runs, released and restored on every exit.
struct TuneScope {
    Core& c;
    float4 *pos0 = nullptr, *vel0 = nullptr;
    hipEvent_t ea = nullptr, eb = nullptr;
    bool timing, graph;
    fsi_pbf_timing acc;
    explicit TuneScope(Core& core) : c(core), timing(core.timing), graph(core.use_graph), acc(core.acc) {
        c.timing = false;
        c.use_graph = false;
    }
    void restore_state() {
        const uint32_t n = c.d.n;
        if (pos0) (void)hipMemcpyAsync(c.d.pos, pos0, sizeof(float4) * n, hipMemcpyDeviceToDevice, c.stream);
        if (vel0) (void)hipMemcpyAsync(c.d.vel, vel0, sizeof(float4) * n, hipMemcpyDeviceToDevice, c.stream);
    }
    ~TuneScope() {
        restore_state();
        (void)hipStreamSynchronize(c.stream);
        if (ea) (void)hipEventDestroy(ea);
        if (eb) (void)hipEventDestroy(eb);
        dev_free(&pos0);
        dev_free(&vel0);
        c.timing = timing;
        c.use_graph = graph;
        c.acc = acc;
        core_invalidate_graph(c);
    }
};
}  // namespace

int core_autotune(Core& c, const Rebin& rebin, const FramePlan& plan, uint32_t frames, int* gather, int* digit_bits) {
    const uint32_t n = c.d.n;
    struct Knobs {
        int g, bits, block, tile, order;
    };
    const int order_now = c.d.relabel ? 2 : (c.d.morton ? 1 : 0);
    Knobs best{c.d.variant, c.sort_digit_bits, c.d.block, c.d.tile_shape, order_now};
    int rc = FSI_OK;
    if (n > 0 && frames > 0) {
        TuneScope scope(c);
        FSI_HIP_TRY(dev_alloc(&scope.pos0, n));
        FSI_HIP_TRY(dev_alloc(&scope.vel0, n));
        FSI_HIP_TRY(hipMemcpyAsync(scope.pos0, c.d.pos, sizeof(float4) * n, hipMemcpyDeviceToDevice, c.stream));
        FSI_HIP_TRY(hipMemcpyAsync(scope.vel0, c.d.vel, sizeof(float4) * n, hipMemcpyDeviceToDevice, c.stream));
        FSI_HIP_TRY(hipEventCreate(&scope.ea));
        FSI_HIP_TRY(hipEventCreate(&scope.eb));
        // the trials: forms 0 and 1 at both workgroup sizes, form 2 at every
        // tile shape when its binning gives tiles (else it would run as form
        // 1), each in Morton processing order over sorted-slot storage (1) and
        // over relabelled storage (2)
        std::vector<Knobs> trials;
        const int bits[] = {6, 8, 10};
        for (int o = 1; o <= 2; ++o)
            for (int b : bits) {
                for (int g = 0; g <= 1; ++g)
                    for (int bs : {128, 256}) trials.push_back({g, b, bs, c.d.tile_shape, o});
                for (int t = 0; t < kTileShapes; ++t) trials.push_back({2, b, c.d.block, t, o});
            }
        float best_ms = 1e30f;
        c.trials.clear();
        for (const Knobs& t : trials) {
            c.d.variant = t.g;
            c.sort_digit_bits = t.bits;
            c.d.block = t.block;
            c.d.tile_shape = t.tile;
            c.d.morton = t.order >= 1;
            c.d.relabel = t.order == 2;
            StepConst k{};
            if ((rc = rebin(t.g == 2, k)) != FSI_OK) break;
            if (t.g == 2 && !c.d.tiled) continue;
            if ((rc = enqueue_frame(c, k, plan, false)) != FSI_OK) break;  // warm-up
            (void)hipEventRecord(scope.ea, c.stream);
            for (uint32_t f = 0; f < frames && rc == FSI_OK; ++f) rc = enqueue_frame(c, k, plan, false);
            (void)hipEventRecord(scope.eb, c.stream);
            if (rc != FSI_OK) break;
            if (hipEventSynchronize(scope.eb) != hipSuccess) {
                rc = fail(FSI_ERR_HIP, "autotune: event wait failed");
                break;
            }
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, scope.ea, scope.eb);
            c.trials.push_back(fsi_tune_trial{t.g, t.bits, t.order, t.block, t.tile, ms});
            if (ms < best_ms) {
                best_ms = ms;
                best = t;
            }
            scope.restore_state();
        }
    }
    c.d.variant = best.g;
    c.sort_digit_bits = best.bits;
    c.d.block = best.block;
    c.d.tile_shape = best.tile;
    c.d.morton = best.order >= 1;
    c.d.relabel = best.order == 2;
    StepConst k{};
    const int rc2 = rebin(best.g == 2, k);
    if (rc == FSI_OK) rc = rc2;
    if (gather) *gather = best.g;
    if (digit_bits) *digit_bits = best.bits;
    return rc;
}

int core_read_autotune(const Core& c, fsi_tune_trial* out, uint32_t cap, uint32_t* n) {
    if (!n || (cap > 0 && !out)) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    *n = (uint32_t)c.trials.size();
    for (uint32_t i = 0; i < cap && i < *n; ++i) out[i] = c.trials[i];
    return FSI_OK;
}

int core_drain_timing(Core& c) {
    if (c.pending.empty()) return FSI_OK;
    FSI_HIP_TRY(hipStreamSynchronize(c.stream));
    for (const TimedPair& t : c.pending) {
        float ms = 0.0f;
        FSI_HIP_TRY(hipEventElapsedTime(&ms, t.a, t.b));
        c.acc.ms[t.pass] += ms;
        c.acc.launches[t.pass] += 1;
        c.event_pool.push_back(t.a);
        c.event_pool.push_back(t.b);
    }
    c.pending.clear();
    return FSI_OK;
}

int core_set_tuning(Core& c, int knob, int value) {
    switch (knob) {
        case FSI_TUNE_GATHER:
            if (value < 0 || value > 2) return fail(FSI_ERR_INVALID_ARG, "gather form must be 0, 1 or 2");
            c.d.variant = value;
            return FSI_OK;
        case FSI_TUNE_SORT_DIGIT_BITS:
            if (value < 1 || value > kSortMaxDigitBits) return fail(FSI_ERR_INVALID_ARG, "digit bits must be 1..10");
            c.sort_digit_bits = value;
            return FSI_OK;
        case FSI_TUNE_ORDER:
            if (value < 0 || value > 2) return fail(FSI_ERR_INVALID_ARG, "processing order must be 0, 1 or 2");
            c.d.morton = value >= 1;
            c.d.relabel = value == 2;
            return FSI_OK;
        case FSI_TUNE_BLOCK:
            if (value != 128 && value != 256) return fail(FSI_ERR_INVALID_ARG, "workgroup size must be 128 or 256");
            c.d.block = value;
            return FSI_OK;
        case FSI_TUNE_TILE:
            if (value < 0 || value >= kTileShapes) return fail(FSI_ERR_INVALID_ARG, "tile shape must be 0, 1 or 2");
            c.d.tile_shape = value;
            return FSI_OK;
        default:
            return fail(FSI_ERR_INVALID_ARG, "unknown knob");
    }
}

int core_get_tuning(const Core& c, int knob, int* value) {
    switch (knob) {
        case FSI_TUNE_GATHER: *value = c.d.variant; return FSI_OK;
        case FSI_TUNE_SORT_DIGIT_BITS: *value = c.sort_digit_bits; return FSI_OK;
        case FSI_TUNE_ORDER: *value = c.d.relabel ? 2 : (c.d.morton ? 1 : 0); return FSI_OK;
        case FSI_TUNE_BLOCK: *value = c.d.block; return FSI_OK;
        case FSI_TUNE_TILE: *value = c.d.tile_shape; return FSI_OK;
        default: return fail(FSI_ERR_INVALID_ARG, "unknown knob");
    }
}

}  // namespace fsi
