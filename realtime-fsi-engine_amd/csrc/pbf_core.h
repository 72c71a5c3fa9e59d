// pbf_core.h — the part of a solver handle shared by the single-scene handle
// (pbf_capi.hip) and the multi-scene handle (pbf_batch.hip): device buffers,
// the stream, the substep sequence of PBFluids::step
// (src/fluid/PBFluids.cpp:323-517), per-pass timing, tuning knobs and HIP
// graph capture of a frame.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <string>
#include <vector>

#include "../../include/fsi_pbf.h"
#include "pbf_launch.h"

namespace fsi {

// Sets the calling thread's last error text and returns code.
int fail(int code, const std::string& msg);
const char* last_error();

#define FSI_HIP_TRY(expr)                                                                       \
    do {                                                                                        \
        hipError_t _e = (expr);                                                                 \
        if (_e != hipSuccess)                                                                   \
            return ::fsi::fail(FSI_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

bool is_pow2(uint32_t x);
int log2u(uint32_t x);

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

// What one frame runs (PBFluids::step): substeps, solver iterations per
// substep, key bits of the sort, vorticity pass at the end.
struct FramePlan {
    uint32_t substeps = 1;
    uint32_t max_iter = 1;
    int key_bits = 1;
    bool vorticity = false;
};

struct Core {
    int device = 0;
    hipStream_t stream = nullptr;      // where every launch and copy of the handle goes
    hipStream_t own_stream = nullptr;  // the stream the handle created (stream unless the caller set one)
    DeviceState d{};
    uint32_t capacity = 0;        // particle buffers allocated for
    uint32_t cells_capacity = 0;  // cell-table entries allocated
    bool stepped = false;
    bool taps = false;
    int sort_digit_bits = kSortMaxDigitBits;
    // timing: HIP events around the launches of the passes in timing_mask
    bool timing = false;
    uint32_t timing_mask = 0xffu;  // bit FSI_PASS_* per pass
    std::vector<hipEvent_t> event_pool;
    std::vector<TimedPair> pending;
    fsi_pbf_timing acc{};
    // HIP graph of one frame
    bool use_graph = false;
    bool graph_valid = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned char> graph_key;  // StepConst + plan + knobs the graph was captured with
    std::vector<fsi_tune_trial> trials;    // the settings the last autotune timed
    // when set, every frame copies the sorted permutation of its substep s to
    // record_perm + s * n (the adjoint's forward keeps them for its replay)
    uint32_t* record_perm = nullptr;
};

// Binds the handle's device and creates its stream.
int core_init(Core& c);
// Enqueues on s from now on (nullptr: the handle's own stream) after the
// current stream has drained; s must belong to the handle's device.
int core_set_stream(Core& c, hipStream_t s);
void core_destroy(Core& c);
// (Re)allocates the per-particle buffers for n particles.
hipError_t core_alloc_particles(Core& c, uint32_t n);
// (Re)allocates and clears the cell table (total entries over all scenes).
hipError_t core_alloc_cells(Core& c, uint32_t entries);
// Enqueues one frame (direct launches, or the captured graph).
int core_frame(Core& c, const StepConst& k, const FramePlan& plan);
int core_drain_timing(Core& c);
int core_set_tuning(Core& c, int knob, int value);
int core_get_tuning(const Core& c, int knob, int* value);
void core_invalidate_graph(Core& c);
// The step constants of the frames to time, with the binning the tiled
// gathers need (tiles) or the one the other forms use; also sets the device
// state's `tiled`.
using Rebin = std::function<int(bool tiles, StepConst& out)>;
// Times `frames` frames under every combination of gather form, radix digit
// width, workgroup size (forms 0 and 1), tile shape (form 2, when the binning
// gives tiles) and storage order (Morton processing order over sorted slots,
// or relabelled), restoring the particle state (pos, vel) after each trial;
// keeps the fastest, leaves the binning for it, and returns its gather form
// and digit width (the rest through core_get_tuning).
int core_autotune(Core& c, const Rebin& rebin, const FramePlan& plan, uint32_t frames, int* gather, int* digit_bits);
// The trials of the last autotune (up to cap into out; *n = their count).
int core_read_autotune(const Core& c, fsi_tune_trial* out, uint32_t cap, uint32_t* n);

// Morton binning of one scene's processing order: cover its bounds (plus two
// cells) with at most 2^axis_bits bins per axis, numbered from bin_base, each
// bin as few cells as possible -- or, with `tiles`, 2^kTileShift cells per
// axis when that covers the box (the tiles of the tiled gathers).
void set_binning(Domain& d, const float bmin[3], const float bmax[3], float h, uint32_t axis_bits,
                 uint32_t bin_base, bool tiles);

// The step constants shared by every scene (FluidConfigUBO minus the box).
StepConst step_const_common(const fsi_fluid_config& cfg, const fsi_pbf_constants& u, uint32_t n, bool interacting,
                            const float ipos[3], const float cam[3]);

// The constant block of PBFluids::setParams (src/fluid/PBFluids.cpp:44-101)
// for the given (already padded) bounds and damping.
void pack_constants(const fsi_fluid_config& p, const float bmin[3], const float bmax[3], float damping, uint32_t n,
                    fsi_pbf_constants& u);

int check_config(const fsi_fluid_config* cfg);

}  // namespace fsi
