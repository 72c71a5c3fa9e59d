// pbf_launch.h — internal launch interface between the solver handle
// (pbf_capi.hip) and the kernels (pbf_kernels.hip, radix_sort.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "pbf_device.h"

namespace fsi {

// Keys per radix tile = 256 threads x kSortIPT.
constexpr int kSortIPT = 8;
// Neighbour-list capacity: rows of the lambda -> delta-p list in HBM.  A
// particle of a fluid at rest density has about 32 neighbours, more under
// compression and where two of its 27 cells hash to one bucket (visited
// twice); a longer row makes delta-p take the bucket walk instead.
constexpr int kNbrCap = 64;
// Widest radix digit the solver plans with (LDS: 44 KiB per workgroup at 10).
constexpr int kSortMaxDigitBits = 10;

struct SortScratch {
    uint32_t* keys[2];  // ping-pong key buffers, n each
    uint32_t* ids[2];   // ping-pong id buffers, n each
    uint32_t* hist;     // sort_hist_entries(n, kSortMaxDigitBits)
    uint32_t* partial;  // scan_partials_needed(hist entries)
};

size_t sort_hist_entries(uint32_t n, int max_digit_bits);
size_t scan_partials_needed(size_t n);
int sort_plan(int key_bits, int max_digit_bits, int digit_bits[8]);

// Stable sort of (keys_in[i], i) by the low key_bits bits of the key.  The
// result lands in one of the scratch ping-pong buffers (returned pointers).
hipError_t radix_sort_pairs(const uint32_t* keys_in, uint32_t n, int key_bits, int max_digit_bits,
                            SortScratch& s, uint32_t** keys_out, uint32_t** ids_out, hipStream_t st);

// out[i] = sum(in[0..i)), in == out allowed.
hipError_t scan_exclusive_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* partial, hipStream_t st);

// Device buffers of one solver handle (particle-id order unless "sorted").
constexpr uint32_t kDeferWords = 9;
constexpr int kTileShapes = 3;  // tiles of 1, 2 or 4 Morton bins of 4x4x4 cells

struct DeviceState {
    uint32_t n = 0, hash_size = 0;
    float4* pos = nullptr;     // pos.xyz, pad0
    float4* vel = nullptr;     // vel.xyz, pad2
    float4* xstar = nullptr;   // predicted position, id order
    uint32_t* keys = nullptr;  // unsorted keys, id order
    uint32_t* skeys = nullptr; // sorted keys (points into sort scratch)
    uint32_t* perm = nullptr;  // sorted slot -> particle id (points into sort scratch)
    int2* cells = nullptr;     // hash_size (start, end)
    float4* P = nullptr;       // sorted: x* being corrected
    uint2* Q = nullptr;        // sorted: quantised positions the next gather's candidate pre-test reads
    float4* PL = nullptr;      // sorted: x* and lambda
    float4* vs = nullptr;      // sorted: post-predict velocity (v*, pad2)
    float4* curl = nullptr;    // sorted: curl / rho0 of the last Integrate
    float4* posS = nullptr;    // sorted: committed positions of the last substep
    float4* vdelta = nullptr;  // sorted: XSPH velocity change of the last Integrate
    float* rho = nullptr;      // sorted: density of the last lambda pass
    uint32_t* lod = nullptr;   // sorted: APBF level of detail
    float4* dp_tap = nullptr;  // sorted: last delta-p (taps only)
    uint32_t* cand = nullptr;  // sorted: candidates of the last lambda pass (taps only)
    uint32_t* nbr = nullptr;   // sorted: neighbours of the last lambda pass (taps only)
    uint32_t* flags = nullptr; // n: contact flags
    uint32_t* slots = nullptr; // n: exclusive scan of the contact flags
    uint32_t* bins = nullptr;  // kOrderBins: Morton-bin counts, then their exclusive scan
    uint32_t* bin_of = nullptr;  // sorted slot -> Morton bin (relabelled: its bucket's, at the bucket's first slot)
    uint32_t* bin_rank = nullptr;  // sorted slot -> rank inside its bin (relabelled: the bucket's offset in it)
    uint32_t* bin_m = nullptr;   // relabelled: storage slot -> Morton bin of its bucket
    uint32_t* mperm = nullptr;   // relabelled: storage slot -> particle id
    int2* mcells = nullptr;      // relabelled: the cell table over storage slots
    uint32_t* order = nullptr; // processing order: thread t handles sorted slot order[t]
    uint32_t* list_count = nullptr;  // by thread: neighbours found by the last lambda pass
    uint32_t* list = nullptr;  // kNbrCap slots per thread, n rounded up to 64 (layout: wave_run in pbf_kernels.hip)
    uint32_t* tflag = nullptr; // n: 1 where a tile starts in the processing order
    uint32_t* tidx = nullptr;  // n: exclusive scan of tflag
    uint32_t* tiles = nullptr; // n + 1: processing-order position where tile i starts (then the end)
    uint32_t* ntiles = nullptr;  // 1: number of tiles
    uint32_t* defer = nullptr;   // n: slots the tiled pass left for the deferred pass
    uint32_t* ndefer = nullptr;  // kDeferWords: the deferred count, then the tile queue heads of the 8 XCDs
    SortScratch sort{};
    int variant = 1;           // gather form (tuning knob, see gather() in pbf_kernels.hip)
    int block = 128;           // workgroup size of the gather kernels, 128 or 256 (tuning knob)
    int tile_shape = 0;        // tiled gathers: tiles of 2^tile_shape bins, 0..2 (tuning knob)
    bool morton = true;        // processing order: Morton order of the cells (else sorted slots)
    bool relabel = false;      // storage order = processing order: whole buckets in Morton order of their cells
    bool stored_relabel = false;  // the storage order of the last enqueued reorder (for the host readbacks)
    bool tiled = false;        // every domain bins by 2^kTileShift cells (tiled gathers possible)
    bool batched = false;      // several scenes (StepConst::doms / scene_of)
    bool any_sdf = false;      // some domain has an SDF boundary
};

hipError_t launch_predict(const StepConst& c, DeviceState& d, hipStream_t st);
hipError_t launch_cells(const StepConst& c, DeviceState& d, hipStream_t st);
hipError_t launch_reorder(const StepConst& c, DeviceState& d, hipStream_t st);
hipError_t launch_lod(const StepConst& c, DeviceState& d, hipStream_t st);
hipError_t launch_lambda(const StepConst& c, DeviceState& d, uint32_t iter, bool taps, hipStream_t st);
hipError_t launch_deltap(const StepConst& c, DeviceState& d, uint32_t iter, bool taps, hipStream_t st);
hipError_t launch_integrate(const StepConst& c, DeviceState& d, hipStream_t st);
hipError_t launch_vorticity(const StepConst& c, DeviceState& d, hipStream_t st);
hipError_t launch_contacts(const StepConst& c, DeviceState& d, float band, float4* out, uint32_t* count,
                           hipStream_t st);
// Block ranges (Domain::sdf_brange) of an SDF grid of dims voxels (x fastest)
// into out, which holds sdf_block_count(dims) entries.
hipError_t launch_sdf_block_range(const float* grid, const int32_t dims[3], float2* out, hipStream_t st);
inline size_t sdf_block_count(const int32_t dims[3]) {
    return (size_t)sdf_blocks(dims[0]) * (size_t)sdf_blocks(dims[1]) * (size_t)sdf_blocks(dims[2]);
}

}  // namespace fsi
