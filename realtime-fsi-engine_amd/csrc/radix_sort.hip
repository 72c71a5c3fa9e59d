// radix_sort.hip — stable LSD radix sort of (key, id) pairs for gfx950.
//
// Replaces CS_RadixHistogram / CS_RadixPrefixSum / CS_RadixScatter
// (resources/shaders/compute/CS_Radix*.glsl, driven by
// src/fluid/PBFluids.cpp:355-404).  The reference sorts all 32 key bits with
// four 8-bit passes; the result is the stable order of the (key, id) pairs
// emitted in id order.  Keys here are masked to log2(hash_size) bits, so only
// that many bits are sorted (the upper passes of the reference are identity
// permutations); the digit width is a launch parameter (4..11 bits).
//
// Per pass:
//   hist    — one workgroup per tile of TILE keys builds an LDS histogram with
//             LDS atomics and writes it column-major: hist[digit * tiles + tile].
//   scan    — device-wide exclusive scan of the histogram (three kernels).
//   scatter — the tile is ranked stably in rounds of 256 keys: within a
//             64-lane wavefront the lanes holding the same digit are found
//             with one ballot per digit bit; per-wave digit counts in LDS give
//             the offsets across the four waves and the earlier rounds.  Keys
//             are first placed in LDS in digit order, then written out so that
//             consecutive lanes store consecutive addresses of a digit run.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbf_launch.h"

namespace fsi {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }

// ---------------- histogram ----------------

template <int BITS, int IPT>
__global__ __launch_bounds__(kThreads) void k_radix_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                         uint32_t shift, uint32_t tiles,
                                                         uint32_t* __restrict__ hist) {
    constexpr int RADIX = 1 << BITS;
    constexpr uint32_t TILE = kThreads * IPT;
    __shared__ uint32_t h[RADIX];
    for (int d = threadIdx.x; d < RADIX; d += kThreads) h[d] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * TILE;
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
        const uint32_t e = base + i * kThreads + threadIdx.x;
        if (e < n) atomicAdd(&h[(keys[e] >> shift) & (RADIX - 1)], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < RADIX; d += kThreads) hist[(size_t)d * tiles + blockIdx.x] = h[d];
}

// ---------------- scatter ----------------

template <int BITS, int IPT>
__global__ __launch_bounds__(kThreads) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                            const uint32_t* __restrict__ iin, uint32_t n,
                                                            uint32_t shift, uint32_t tiles,
                                                            const uint32_t* __restrict__ scanned,
                                                            uint32_t* __restrict__ kout,
                                                            uint32_t* __restrict__ iout) {
    constexpr int RADIX = 1 << BITS;
    constexpr uint32_t TILE = kThreads * IPT;
    __shared__ uint32_t gbase[RADIX];   // global start of this tile's run of digit d
    __shared__ uint32_t tstart[RADIX];  // start of digit d inside the LDS-sorted tile
    __shared__ uint32_t run[RADIX];     // keys of digit d placed so far
    __shared__ uint32_t wcnt[kWaves][RADIX];
    __shared__ uint32_t lkey[TILE];
    __shared__ uint32_t lid[TILE];
    __shared__ uint32_t wsum[kThreads / 64 + 1];

    const uint32_t tile = blockIdx.x;
    const uint32_t base = tile * TILE;
    const uint32_t count = (n - base) < TILE ? (n - base) : TILE;
    const size_t total = (size_t)RADIX * tiles;

    // Digit counts of this tile from the scanned histogram: consecutive entries
    // of the flattened exclusive scan differ by the count of the earlier entry.
    constexpr int DPT = (RADIX + kThreads - 1) / kThreads;
    uint32_t cnt[DPT];
    uint32_t local = 0;
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
        const int d = threadIdx.x * DPT + q;
        uint32_t c = 0;
        if (d < RADIX) {
            const size_t x = (size_t)d * tiles + tile;
            const uint32_t cur = scanned[x];
            const uint32_t nxt = (x + 1 < total) ? scanned[x + 1] : n;
            c = nxt - cur;
            gbase[d] = cur;
            run[d] = 0;
        }
        cnt[q] = c;
        local += c;
    }
    // Block-wide exclusive scan of the per-thread digit totals (digits are
    // assigned to threads in contiguous groups of DPT, so thread order is
    // digit order).
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = local;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t wave_off = 0;
    for (int w = 0; w < wave; ++w) wave_off += wsum[w];
    uint32_t acc = wave_off + incl - local;
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
        const int d = threadIdx.x * DPT + q;
        if (d < RADIX) tstart[d] = acc;
        acc += cnt[q];
    }
    for (int d = threadIdx.x; d < kWaves * RADIX; d += kThreads) (&wcnt[0][0])[d] = 0;
    __syncthreads();

    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int i = 0; i < IPT; ++i) {
        const uint32_t le = i * kThreads + threadIdx.x;  // element index inside the tile
        const bool valid = le < count;
        uint32_t key = 0, id = 0, d = 0;
        if (valid) {
            key = kin[base + le];
            id = iin ? iin[base + le] : (base + le);
            d = (key >> shift) & (RADIX - 1);
        }
        // lanes of this wave holding the same digit
        uint64_t peers = ballot64(valid);
#pragma unroll
        for (int b = 0; b < BITS; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = ballot64(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = __popcll(peers & lt_mask);
        const bool leader = valid && rank == 0;
        if (leader) wcnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t prefix = run[d];
            for (int w = 0; w < wave; ++w) prefix += wcnt[w][d];
            const uint32_t pos = tstart[d] + prefix + rank;
            lkey[pos] = key;
            lid[pos] = id;
        }
        __syncthreads();
        if (leader) {
            atomicAdd(&run[d], wcnt[wave][d]);
            wcnt[wave][d] = 0;
        }
        __syncthreads();
    }

    // Coalesced write-out: LDS slot e holds the e-th key of the tile in
    // (digit, input order); its global slot is gbase[d] + (e - tstart[d]).
    for (uint32_t e = threadIdx.x; e < count; e += kThreads) {
        const uint32_t key = lkey[e];
        const uint32_t d = (key >> shift) & (RADIX - 1);
        const uint32_t dst = gbase[d] + (e - tstart[d]);
        kout[dst] = key;
        iout[dst] = lid[e];
    }
}

// ---------------- device-wide exclusive scan (u32) ----------------

constexpr int kScanIPT = 16;
constexpr uint32_t kScanTile = kThreads * kScanIPT;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* wsum, uint32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t wave_off = 0;
    total = 0;
    for (int w = 0; w < kWaves; ++w) {
        if (w < wave) wave_off += wsum[w];
        total += wsum[w];
    }
    __syncthreads();
    return wave_off + incl - v;
}

__global__ __launch_bounds__(kThreads) void k_scan_reduce(const uint32_t* __restrict__ in, size_t n,
                                                          uint32_t* __restrict__ partial) {
    __shared__ uint32_t wsum[kWaves];
    const size_t base = (size_t)blockIdx.x * kScanTile;
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanIPT; ++i) {
        const size_t e = base + (size_t)i * kThreads + threadIdx.x;
        if (e < n) s += in[e];
    }
    uint32_t total;
    (void)block_exclusive_scan(s, wsum, total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// Single workgroup: exclusive scan of the per-tile partial sums, in place.
__global__ __launch_bounds__(kThreads) void k_scan_partials(uint32_t* __restrict__ partial, uint32_t m) {
    __shared__ uint32_t wsum[kWaves];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < m; base += kThreads) {
        const uint32_t e = base + threadIdx.x;
        const uint32_t v = e < m ? partial[e] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, wsum, total);
        if (e < m) partial[e] = carry + ex;
        carry += total;
    }
}

__global__ __launch_bounds__(kThreads) void k_scan_downsweep(const uint32_t* __restrict__ in, size_t n,
                                                             const uint32_t* __restrict__ partial,
                                                             uint32_t* __restrict__ out) {
    __shared__ uint32_t wsum[kWaves];
    // thread t owns kScanIPT consecutive elements (blocked arrangement)
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanIPT;
    uint32_t v[kScanIPT];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanIPT; ++i) {
        const size_t e = base + i;
        v[i] = e < n ? in[e] : 0u;
        s += v[i];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(s, wsum, total) + partial[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kScanIPT; ++i) {
        const size_t e = base + i;
        if (e < n) out[e] = run;
        run += v[i];
    }
}

template <int BITS, int IPT>
hipError_t sort_pass(const uint32_t* kin, const uint32_t* iin, uint32_t* kout, uint32_t* iout, uint32_t n,
                     uint32_t shift, SortScratch& s, hipStream_t st) {
    constexpr uint32_t TILE = kThreads * IPT;
    const uint32_t tiles = (n + TILE - 1) / TILE;
    const size_t hn = ((size_t)1 << BITS) * tiles;
    hipLaunchKernelGGL((k_radix_hist<BITS, IPT>), dim3(tiles), dim3(kThreads), 0, st, kin, n, shift, tiles, s.hist);
    hipError_t e = scan_exclusive_u32(s.hist, s.hist, hn, s.partial, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_radix_scatter<BITS, IPT>), dim3(tiles), dim3(kThreads), 0, st, kin, iin, n, shift, tiles,
                       s.hist, kout, iout);
    return hipGetLastError();
}

template <int IPT>
hipError_t sort_pass_bits(int bits, const uint32_t* kin, const uint32_t* iin, uint32_t* kout, uint32_t* iout,
                          uint32_t n, uint32_t shift, SortScratch& s, hipStream_t st) {
    switch (bits) {
        case 1: return sort_pass<1, IPT>(kin, iin, kout, iout, n, shift, s, st);
        case 2: return sort_pass<2, IPT>(kin, iin, kout, iout, n, shift, s, st);
        case 3: return sort_pass<3, IPT>(kin, iin, kout, iout, n, shift, s, st);
        case 4: return sort_pass<4, IPT>(kin, iin, kout, iout, n, shift, s, st);
        case 5: return sort_pass<5, IPT>(kin, iin, kout, iout, n, shift, s, st);
        case 6: return sort_pass<6, IPT>(kin, iin, kout, iout, n, shift, s, st);
        case 7: return sort_pass<7, IPT>(kin, iin, kout, iout, n, shift, s, st);
        case 8: return sort_pass<8, IPT>(kin, iin, kout, iout, n, shift, s, st);
        case 9: return sort_pass<9, IPT>(kin, iin, kout, iout, n, shift, s, st);
        case 10: return sort_pass<10, IPT>(kin, iin, kout, iout, n, shift, s, st);
        case 11: return sort_pass<11, IPT>(kin, iin, kout, iout, n, shift, s, st);
        default: return hipErrorInvalidValue;
    }
}

__global__ void k_iota(uint32_t* out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = i;
}

}  // namespace

hipError_t scan_exclusive_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* partial, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const uint32_t blocks = (uint32_t)((n + kScanTile - 1) / kScanTile);
    hipLaunchKernelGGL(k_scan_reduce, dim3(blocks), dim3(kThreads), 0, st, in, n, partial);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kThreads), 0, st, partial, blocks);
    hipLaunchKernelGGL(k_scan_downsweep, dim3(blocks), dim3(kThreads), 0, st, in, n, partial, out);
    return hipGetLastError();
}

size_t scan_partials_needed(size_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

size_t sort_hist_entries(uint32_t n, int max_digit_bits) {
    const uint32_t tile = kThreads * kSortIPT;
    const size_t tiles = (n + tile - 1) / tile;
    return ((size_t)1 << max_digit_bits) * (tiles ? tiles : 1);
}

int sort_plan(int key_bits, int max_digit_bits, int digit_bits[8]) {
    if (key_bits <= 0) return 0;
    const int passes = (key_bits + max_digit_bits - 1) / max_digit_bits;
    int left = key_bits;
    for (int p = 0; p < passes; ++p) {
        const int w = (left + (passes - p) - 1) / (passes - p);  // balanced digit widths
        digit_bits[p] = w;
        left -= w;
    }
    return passes;
}

hipError_t radix_sort_pairs(const uint32_t* keys_in, uint32_t n, int key_bits, int max_digit_bits,
                            SortScratch& s, uint32_t** keys_out, uint32_t** ids_out, hipStream_t st) {
    int digits[8];
    const int passes = sort_plan(key_bits, max_digit_bits, digits);
    if (passes == 0 || n == 0) {
        // every key is 0: the stable order is the id order
        if (n) {
            hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, st, s.ids[0], n);
            hipError_t e = hipMemcpyAsync(s.keys[0], keys_in, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st);
            if (e != hipSuccess) return e;
        }
        *keys_out = s.keys[0];
        *ids_out = s.ids[0];
        return hipGetLastError();
    }
    const uint32_t* kin = keys_in;
    const uint32_t* iin = nullptr;  // first pass: id = element index
    uint32_t shift = 0;
    for (int p = 0; p < passes; ++p) {
        uint32_t* ko = s.keys[p & 1];
        uint32_t* io = s.ids[p & 1];
        hipError_t e = sort_pass_bits<kSortIPT>(digits[p], kin, iin, ko, io, n, shift, s, st);
        if (e != hipSuccess) return e;
        kin = ko;
        iin = io;
        shift += digits[p];
    }
    *keys_out = const_cast<uint32_t*>(kin);
    *ids_out = const_cast<uint32_t*>(iin);
    return hipSuccess;
}

}  // namespace fsi
