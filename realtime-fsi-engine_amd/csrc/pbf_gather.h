// pbf_gather.h — the neighbour visit shared by the forward kernels
// (pbf_kernels.hip) and the recording / adjoint kernels (pbf_adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbf_device.h"

namespace fsi {

// Bucket numbers of the 27 cells around (cx, cy, cz) in the reference's visit
// order.  cell_hash multiplies each coordinate (wrapping int32, then XOR,
// CS_PredictAndHash.glsl:77-83); (c + d) * P = c * P + d * P mod 2^32, so
// the three products are formed once and the offsets added per bucket.
struct BucketWalk {
    uint32_t hx, hy, hz, mask, base;
    __device__ __forceinline__ BucketWalk(int32_t cx, int32_t cy, int32_t cz, const Domain& dm)
        : hx((uint32_t)(cx - 1) * 73856093u),
          hy((uint32_t)(cy - 1) * 19349663u),
          hz((uint32_t)(cz - 1) * 83492791u),
          mask(dm.hash_mask),
          base(dm.hash_base) {}
    __device__ __forceinline__ BucketWalk(int32_t cx, int32_t cy, int32_t cz, uint32_t hash_mask, uint32_t hash_base)
        : hx((uint32_t)(cx - 1) * 73856093u),
          hy((uint32_t)(cy - 1) * 19349663u),
          hz((uint32_t)(cz - 1) * 83492791u),
          mask(hash_mask),
          base(hash_base) {}
    // q = 9 dz + 3 dy + dx over offsets 0..2 (dx fastest)
    __device__ __forceinline__ uint32_t bucket(int q) const {
        const uint32_t ix = (uint32_t)(q % 3), iy = (uint32_t)((q / 3) % 3), iz = (uint32_t)(q / 9);
        return base + (((hx + ix * 73856093u) ^ (hy + iy * 19349663u) ^ (hz + iz * 83492791u)) & mask);
    }
};

// Visit the candidates of the 27 buckets around x in the reference order
// (CS_ComputeLambdas.glsl:126-150: bucket offsets dz, dy, dx with dx fastest,
// slots in sorted order, the particle's own slot skipped, an aliased bucket
// visited once per occurrence).  f(k, xk, r, r2) sees every candidate, with
// r = x - xk.  The next bucket's range is loaded while the current bucket is
// scanned, and candidates are loaded four at a time (indexes clamped to the
// bucket, so every load stays inside it): up to five loads are in flight per
// lane instead of one exposed latency per candidate.
template <class F>
__device__ __forceinline__ void scan_buckets(V3 x, float h, const Domain& dm, const int2* __restrict__ cells,
                                             const float4* __restrict__ X, uint32_t self, F&& f) {
    int32_t cx, cy, cz;
    cell_of(x, h, cx, cy, cz);
    const BucketWalk bw(cx, cy, cz, dm);
    int2 se = cells[bw.bucket(0)];
    auto visit = [&](uint32_t k, float4 xk) {
        if (k == self) return;
        const V3 r = sub(x, xyz(xk));
        f(k, xk, r, dot(r, r));
    };
#pragma unroll 1
    for (int q = 0; q < 27; ++q) {
        int2 nxt = make_int2(-1, -1);
        if (q < 26) nxt = cells[bw.bucket(q + 1)];
        // an empty bucket is (-1, -1): no iterations
        const uint32_t end = (uint32_t)se.y, last = end - 1u;
        for (uint32_t k = (uint32_t)se.x; k < end; k += 4) {
            const uint32_t k1 = min(k + 1, last), k2 = min(k + 2, last), k3 = min(k + 3, last);
            const float4 a0 = X[k], a1 = X[k1], a2 = X[k2], a3 = X[k3];
            visit(k, a0);
            if (k + 1 <= last) visit(k + 1, a1);
            if (k + 2 <= last) visit(k + 2, a2);
            if (k + 3 <= last) visit(k + 3, a3);
        }
        se = nxt;
    }
}

}  // namespace fsi
