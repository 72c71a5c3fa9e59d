// pbf_gather.h — the neighbour visit shared by the forward kernels
// (pbf_kernels.hip) and the recording / adjoint kernels (pbf_adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbf_device.h"

namespace fsi {

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
    int2 se = cells[dm.hash_base + cell_hash(cx - 1, cy - 1, cz - 1, dm.hash_mask)];
    auto visit = [&](int32_t k, float4 xk) {
        if ((uint32_t)k == self) return;
        const V3 r = sub(x, xyz(xk));
        f((uint32_t)k, xk, r, dot(r, r));
    };
#pragma unroll 1
    for (int q = 0; q < 27; ++q) {
        int2 nxt = make_int2(-1, -1);
        if (q < 26) {
            const int qn = q + 1;
            nxt = cells[dm.hash_base + cell_hash(cx + qn % 3 - 1, cy + (qn / 3) % 3 - 1, cz + qn / 9 - 1, dm.hash_mask)];
        }
        // an empty bucket is (-1, -1): no iterations
        const int32_t last = se.y - 1;
        for (int32_t k = se.x; k < se.y; k += 4) {
            const int32_t k1 = min(k + 1, last), k2 = min(k + 2, last), k3 = min(k + 3, last);
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
