// pbf_adjoint.hip — reverse-mode derivative of whole frames of the PBF step
// (fsi_adj_* in include/fsi_pbf.h).
//
// Forward: the frames run on the ordinary kernels; the state (pos, vel) at the
// start of every frame is kept as a checkpoint.
//
// Backward, frame by frame from the last: the frame is run again from its
// checkpoint by recording kernels that walk explicit neighbour lists instead
// of the hash buckets.  The lists hold the pairs each pass sums over in the
// forward's visit order (CS_ComputeLambdas.glsl:126-150: buckets z, y, x,
// sorted slots, aliased buckets repeated, pairs outside the support dropped),
// so the replay repeats the forward's FP32 operations exactly and the tape it
// writes (positions per solver iteration, lambda, density, constraint
// gradient, raw delta-p, XSPH sum, vorticity direction) is the one the
// forward saw.  The reverse kernels then run the passes backwards.  A pass
// that sums over the pairs (i, j) of slot i has a partial derivative for slot
// j too; each reverse pass writes those per pair, and a second kernel sums
// them into slot j through the transposed list (the pairs that reach j).  The
// rows of that list fill in no fixed order, so their sums are formed in 64-bit
// integer units of the row's largest term (RowSum), which do not depend on
// the order: the gradient is reproducible bit for bit.
//
// Scope: the substep of PBFluids::step (src/fluid/PBFluids.cpp:323-517) with
// s_corr, cohesion floor, XSPH viscosity, curl, AABB clamp with damping,
// speed clamp and vorticity confinement.  APBF levels of detail, the SDF
// boundary and the mouse interaction are not differentiated (a handle cannot
// enable them).  Differentiable inputs: initial positions and velocities,
// the XSPH coefficient and the relaxation epsilon.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/fsi_pbf.h"
#include "pbf_core.h"
#include "pbf_gather.h"

using namespace fsi;

namespace {

constexpr int kAB = 256;
// Pairs per slot in one list, and pairs reaching one slot in the transposed
// list.  A fluid at rest density has about 32 neighbours per particle; a
// bucket that two of the 27 cells hash to is visited twice (the reference
// does not deduplicate), which can double a count.
constexpr uint32_t kEdgeCap = 96;
constexpr uint32_t kInCap = 128;

__device__ __forceinline__ uint32_t agid() { return blockIdx.x * kAB + threadIdx.x; }
inline dim3 agrid(uint32_t n) { return dim3((n + kAB - 1) / kAB); }

// Quads: at 65,536 particles one thread per slot is one wave per SIMD, and
// the FP64 reverse passes stall on their loads and long dependency chains.
// The reverse kernels give each slot four adjacent lanes (lane = thread & 3)
// that take every fourth pair; per-slot sums are reduced across the quad.
constexpr uint32_t kQ = 4;
inline dim3 qgrid(uint32_t n) { return dim3((uint32_t)(((size_t)n * kQ + kAB - 1) / kAB)); }
__device__ __forceinline__ double qsum(double v) {
    v += __shfl_xor(v, 1, (int)kQ);
    v += __shfl_xor(v, 2, (int)kQ);
    return v;
}

// Order-independent sums over a transposed row.  The atomic appends of the
// list builders (quad_list) fill a row in whatever order they run, so its
// terms are added in a way that does not depend on the order: every term is
// rounded to a multiple of 2^(E - 55), E the binary exponent of the row's
// largest |term|, and the multiples are added as 64-bit integers (exact,
// associative).  A row holds at most kInCap = 128 terms of at most 2^55
// units: no overflow.  The rounding error, 2^-56 of the largest term per
// term, is below that of a sequential FP64 sum.  The same row always gives
// the same bits.
struct RowSum {
    int e = 0;  // exponent of the row's largest |term| (0 when all are zero)
    __device__ __forceinline__ void set(double maxabs) {
        e = 0;
        if (maxabs > 0.0) (void)frexp(maxabs, &e);
    }
    __device__ __forceinline__ long long units(double v) const { return (long long)rint(ldexp(v, 55 - e)); }
    __device__ __forceinline__ double value(long long u) const { return ldexp((double)u, e - 55); }
};

// Lanes per transposed row in the kernels that sum rows, and the terms each
// lane keeps in registers between the two passes of row_sum (rows of up to
// kR * kRKeep terms are read once).
constexpr uint32_t kR = 8, kRKeep = 4;
inline dim3 rgrid(uint32_t n) { return dim3((uint32_t)(((size_t)n * kR + kAB - 1) / kAB)); }
__device__ __forceinline__ double rmax(double v) {
#pragma unroll
    for (int o = 1; o < (int)kR; o <<= 1) v = fmax(v, __shfl_xor(v, o, (int)kR));
    return v;
}
__device__ __forceinline__ long long rsum_i64(long long v) {
#pragma unroll
    for (int o = 1; o < (int)kR; o <<= 1) v += __shfl_xor(v, o, (int)kR);
    return v;
}
__device__ __forceinline__ int ror(int v) {
#pragma unroll
    for (int o = 1; o < (int)kR; o <<= 1) v |= __shfl_xor(v, o, (int)kR);
    return v;
}

// The C order-independent sums of transposed row j (m terms) over the kR
// lanes of its group: get(e, v) gives the C terms of pair e.  Lane l takes
// the row's entries l, l + kR, ...; the first kRKeep of them stay in
// registers for the second pass, later ones are read again.  A non-finite
// term makes the sums NaN.
template <int C, class Get>
__device__ __forceinline__ void row_sum(const uint32_t* __restrict__ tin, uint32_t n, uint32_t j, uint32_t m,
                                        uint32_t lane, Get&& get, double out[C]) {
    double keep[kRKeep][C];
    double mx[C];
    int bad = 0;
#pragma unroll
    for (int c = 0; c < C; ++c) mx[c] = 0.0;
    auto scan = [&](const double* v) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const double a = fabs(v[c]);
            bad |= (a <= 1.7976931348623157e308) ? 0 : 1;
            mx[c] = fmax(mx[c], a);
        }
    };
#pragma unroll
    for (uint32_t t = 0; t < kRKeep; ++t) {
        const uint32_t q = lane + t * kR;
        if (q < m) {
            get(tin[(size_t)q * n + j], keep[t]);
        } else {
#pragma unroll
            for (int c = 0; c < C; ++c) keep[t][c] = 0.0;
        }
        scan(keep[t]);
    }
    for (uint32_t q = lane + kRKeep * kR; q < m; q += kR) {
        double v[C];
        get(tin[(size_t)q * n + j], v);
        scan(v);
    }
    bad = ror(bad);
    RowSum rs[C];
    long long u[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        rs[c].set(rmax(mx[c]));
        u[c] = 0;
    }
    if (!bad) {
#pragma unroll
        for (uint32_t t = 0; t < kRKeep; ++t)
#pragma unroll
            for (int c = 0; c < C; ++c) u[c] += rs[c].units(keep[t][c]);
        for (uint32_t q = lane + kRKeep * kR; q < m; q += kR) {
            double v[C];
            get(tin[(size_t)q * n + j], v);
#pragma unroll
            for (int c = 0; c < C; ++c) u[c] += rs[c].units(v[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
        u[c] = rsum_i64(u[c]);
        out[c] = bad ? __longlong_as_double(0x7ff8000000000000ll) : rs[c].value(u[c]);
    }
}

__device__ __forceinline__ V3 zero3() { return mk(0.0f, 0.0f, 0.0f); }
__device__ __forceinline__ float4 f4(V3 v, float w = 0.0f) { return make_float4(v.x, v.y, v.z, w); }

// The sorted keys of a substep from its permutation: skeys[k] = keys[perm[k]].
__global__ __launch_bounds__(kAB) void k_sorted_keys(uint32_t n, const uint32_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ perm, uint32_t* __restrict__ skeys) {
    const uint32_t k = blockIdx.x * kAB + threadIdx.x;
    if (k < n) skeys[k] = keys[perm[k]];
}

__global__ __launch_bounds__(kAB) void k_zero_u32(uint32_t* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * kAB + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Statistics kept as maxima over kSpread slots: block b folds its maximum
// into slot b mod kSpread and the reader takes the maximum over the slots.
// Atomics from many waves on one address queue at the cache; spread over
// kSpread addresses, with one atomic per block, they do not.
constexpr uint32_t kSpread = 64;
// flag words: 0 overflow bits, 1 replay mismatches, then the slots of the
// largest out-degree and of the largest in-degree of the lists
constexpr uint32_t kOutMax = 2, kInMax = 2 + kSpread, kFlagWords = 2 + 2 * kSpread;

// The block's maxima of v[0..C) folded into slots[c] + (block mod kSpread).
// Every thread of the block calls it (it holds a barrier).  Values are
// compared as unsigned integers (the bits of a non-negative float order the
// same way).
template <int C>
__device__ __forceinline__ void block_max(const uint32_t (&v)[C], uint32_t* const (&slots)[C]) {
    __shared__ uint32_t part[C][kAB / 64];
    uint32_t m[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        m[c] = v[c];
        for (int o = 32; o > 0; o >>= 1) m[c] = max(m[c], (uint32_t)__shfl_xor((int)m[c], o));
        if (__lane_id() == 0u) part[c][threadIdx.x / 64u] = m[c];
    }
    __syncthreads();
    if (threadIdx.x < (uint32_t)C) {
        uint32_t b = 0;
        for (uint32_t w = 0; w < kAB / 64; ++w) b = max(b, part[threadIdx.x][w]);
        if (b > 0u) atomicMax(slots[threadIdx.x] + blockIdx.x % kSpread, b);
    }
}

// ---- neighbour lists ------------------------------------------------------
// Pair (i, r) -- the r-th neighbour of slot i -- is stored at r * n + i; its
// number in the transposed list is that same index.

// The neighbour list of slot i, built by the four lanes of a quad in the
// reference's visit order (CS_ComputeLambdas.glsl:126-150): the lanes walk the
// same 27 buckets and test every fourth candidate (lane l takes the bucket's
// slots first + l, first + 4 + l, ...), and a ballot over the quad gives each
// neighbour its rank in visit order.  Four lanes per slot keep four times as
// many waves in flight as one lane per slot, and the candidate loads of a
// quad are contiguous.  inside(xk) is the support test r^2 <= h^2.  Returns
// the slot's neighbour count (entries past kEdgeCap are not stored).
//
// The transposed list is filled in the same pass: every stored pair e =
// r * n + i is appended to row j of tin (an atomic on the row's counter
// tcnt[j], which the frame cleared before its lists were built); a row longer
// than kInCap is refused (flag bit 2).  The rows fill in no fixed order, which
// the sums over them do not depend on (RowSum).  indeg is raised to the
// largest row length this lane saw.
template <class T4, class Inside>
__device__ __forceinline__ uint32_t quad_list(const T4* __restrict__ X, const int2* __restrict__ cells,
                                              const BucketWalk& bw, uint32_t n, uint32_t i, uint32_t lane,
                                              uint32_t* __restrict__ nbr, uint32_t* __restrict__ tin,
                                              uint32_t* __restrict__ tcnt, uint32_t* __restrict__ overflow,
                                              uint32_t& indeg, Inside&& inside) {
    const uint32_t qbase = __lane_id() & ~(kQ - 1u), below = (1u << lane) - 1u;
    uint32_t m = 0;
    auto place = [&](bool in, uint32_t s) {
        const uint32_t qm = (uint32_t)(__ballot(in) >> qbase) & 0xfu;
        if (in) {
            const uint32_t r = m + (uint32_t)__popc(qm & below);
            if (r < kEdgeCap) {
                const uint32_t e = r * n + i;
                nbr[e] = s;
                const uint32_t q = atomicAdd(&tcnt[s], 1u);
                if (q < kInCap) tin[(size_t)q * n + s] = e;
                else atomicOr(overflow, 2u);  // rare: the list is refused
                indeg = max(indeg, q + 1u);
            }
        }
        m += (uint32_t)__popc(qm);
    };
    // the cell-table entries of nine buckets are loaded together (one bucket
    // ahead left a load's latency exposed per bucket: a bucket is mostly one
    // iteration of the quad)
#pragma unroll 1
    for (int q0 = 0; q0 < 27; q0 += 9) {
        int2 se[9];
#pragma unroll
        for (int j = 0; j < 9; ++j) se[j] = cells[bw.bucket(q0 + j)];
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            // an empty bucket is (-1, -1): no iterations
            const uint32_t end = (uint32_t)se[j].y, last = end - 1u;
            for (uint32_t b = (uint32_t)se[j].x; b < end; b += 2u * kQ) {
                const uint32_t sa = b + lane, sb = b + kQ + lane;
                const T4 xa = X[min(sa, last)], xb = X[min(sb, last)];
                place(sa <= last && sa != i && inside(xa), sa);
                place(sb <= last && sb != i && inside(xb), sb);
            }
        }
    }
    return m;
}

// Stores the count of a quad-built list (lane 0), flags a list longer than
// kEdgeCap and keeps the largest count and the largest transposed row
// (indeg).  Every thread of the block calls it; m and indeg are 0 for lanes
// past the last slot.
__device__ __forceinline__ void store_count(uint32_t m, uint32_t indeg, bool live, uint32_t i, uint32_t lane,
                                            uint32_t* __restrict__ cnt, uint32_t* __restrict__ overflow) {
    block_max<2>({m, indeg}, {overflow + kOutMax, overflow + kInMax});
    if (live && lane == 0u) {
        if (m > kEdgeCap) {
            atomicOr(overflow, 1u);
            m = kEdgeCap;
        }
        cnt[i] = m;
    }
}

__global__ __launch_bounds__(kAB) void k_edges(StepConst c, const float4* __restrict__ X,
                                               const int2* __restrict__ cells, uint32_t* __restrict__ nbr,
                                               uint32_t* __restrict__ cnt, uint32_t* __restrict__ tin,
                                               uint32_t* __restrict__ tcnt, uint32_t* __restrict__ overflow) {
    const uint32_t g = agid(), i = g / kQ, lane = g % kQ;
    const bool live = i < c.n;
    uint32_t m = 0, indeg = 0;
    if (live) {
        const V3 x = xyz(X[i]);
        const float h2 = c.h * c.h;
        int32_t cx, cy, cz;
        cell_of(x, c.h, cx, cy, cz);
        m = quad_list(X, cells, BucketWalk(cx, cy, cz, c.dom), c.n, i, lane, nbr, tin, tcnt, overflow, indeg,
                      [&](const float4& xk) {
                          const V3 r = sub(x, xyz(xk));
                          return dot(r, r) <= h2;
                      });
    }
    store_count(m, indeg, live, i, lane, cnt, overflow);
}

// ---- recording forward ----------------------------------------------------

// f(j, X[j]) for the m entries of slot i's list (pair r at nbr[r * n + i]), in
// list order: four entries, then their four records, are loaded before any
// is used (one entry per iteration left each load's latency exposed).
// Entries past m repeat the last one and are not used.
template <class T4, class F>
__device__ __forceinline__ void walk4(const uint32_t* __restrict__ nbr, uint32_t n, uint32_t i, uint32_t m,
                                      const T4* __restrict__ X, F&& f) {
#pragma unroll 1
    for (uint32_t r = 0; r < m; r += 4) {
        uint32_t j[4];
        T4 x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) j[u] = nbr[(size_t)min(r + (uint32_t)u, m - 1u) * n + i];
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = X[j[u]];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (r + (uint32_t)u < m) f(j[u], x[u]);
    }
}

// CS_ComputeLambdas.glsl:108-162 over the list; tape: (lambda, rho), (g, D).
__global__ __launch_bounds__(kAB) void k_rec_lambda(StepConst c, const float4* __restrict__ P,
                                                    const uint32_t* __restrict__ nbr, const uint32_t* __restrict__ cnt,
                                                    float4* __restrict__ PL, float2* __restrict__ lr,
                                                    float4* __restrict__ aux) {
    const uint32_t i = agid();
    if (i >= c.n) return;
    const V3 xi = xyz(P[i]);
    const float inv = c.inv_rho0;
    float rho = c.w0_self;
    float sum_g2 = 0.0f;
    V3 gi = zero3();
    walk4(nbr, c.n, i, cnt[i], P, [&](uint32_t, const float4& xj) {
        const V3 d = sub(xi, xyz(xj));
        const float r2 = dot(d, d);
        rho = rho + poly6_r2(c, r2);
        const V3 gw = spiky(c, d, r2);
        const float gj2 = dot(gw, gw);
        sum_g2 = sum_g2 + gj2 * (inv * inv);
        gi = add(gi, gw);
    });
    gi = scl(gi, inv);
    float C = rho / c.rho0 - 1.0f;
    const float floorc = -c.cohesion;
    C = C > floorc ? C : floorc;
    sum_g2 = sum_g2 + dot(gi, gi);
    const float D = sum_g2 + c.eps;
    const float lambda = -C / D;
    PL[i] = f4(xi, lambda);
    lr[i] = make_float2(lambda, rho);
    aux[i] = f4(gi, D);
}

// CS_ComputeDeltaP.glsl:120-166 + the clamp and add of CS_ApplyDeltaP.glsl:80-106;
// tape: the corrected position and the raw delta-p.
__global__ __launch_bounds__(kAB) void k_rec_deltap(StepConst c, const float4* __restrict__ PL,
                                                    const uint32_t* __restrict__ nbr, const uint32_t* __restrict__ cnt,
                                                    float4* __restrict__ P, float4* __restrict__ Ptape,
                                                    float4* __restrict__ dptape) {
    const uint32_t i = agid();
    if (i >= c.n) return;
    const float4 xi4 = PL[i];
    const V3 xi = xyz(xi4);
    const float li = xi4.w;
    V3 dp = zero3();
    walk4(nbr, c.n, i, cnt[i], PL, [&](uint32_t, const float4& a) {
        const V3 d = sub(xi, xyz(a));
        const float r2 = dot(d, d);
        const V3 gw = spiky(c, d, r2);
        const float sc = scorr_r2(c, r2);
        dp = add(dp, scl(gw, li + a.w + sc));
    });
    dp = scl(dp, c.inv_rho0);
    dptape[i] = f4(dp);
    const float maxdp = 0.5f * c.h;
    const float l2 = dot(dp, dp);
    const float max2 = maxdp * maxdp;
    if (l2 > max2) {
        const float l = sqrtf(l2);
        dp = scl(dp, maxdp / l);
    }
    const V3 p = add(xi, dp);
    P[i] = f4(p);
    Ptape[i] = f4(p);
}

// CS_Integrate.glsl:135-280 (no SDF) over the list; tape: the XSPH sum.
__global__ __launch_bounds__(kAB) void k_rec_integrate(StepConst c, const float4* __restrict__ P,
                                                       const float4* __restrict__ vs, const uint32_t* __restrict__ perm,
                                                       float4* __restrict__ pos, float4* __restrict__ vel,
                                                       const uint32_t* __restrict__ nbr,
                                                       const uint32_t* __restrict__ cnt, float4* __restrict__ curl_out,
                                                       float4* __restrict__ posS, float4* __restrict__ visc_tape) {
    const uint32_t k = agid();
    if (k >= c.n) return;
    const uint32_t id = perm[k];
    const float4 op4 = pos[id];
    const V3 old_pos = xyz(op4);
    const float4 ov4 = vs[k];
    const V3 old_vel = xyz(ov4);
    V3 xp = xyz(P[k]);
    const float inv = c.inv_rho0;
    const bool do_visc = c.enable_visc == 1u;
    const bool do_vort = c.enable_vort == 1u;
    V3 visc_delta = zero3();
    V3 visc = zero3();
    if (do_visc || do_vort) {
        V3 curl = zero3();
        const uint32_t m = cnt[k];
        for (uint32_t r = 0; r < m; ++r) {
            const uint32_t j = nbr[(size_t)r * c.n + k];
            const V3 d = sub(xp, xyz(P[j]));
            const float r2 = dot(d, d);
            const V3 vij = sub(xyz(vs[j]), old_vel);
            if (do_visc) {
                const float w = poly6_r2(c, r2);
                visc = add(visc, scl(scl(vij, w), inv));
            }
            if (do_vort) {
                const V3 gw = spiky(c, d, r2);
                curl = add(curl, cross(vij, gw));
            }
        }
        if (do_visc) visc_delta = scl(visc, c.viscosity);
        if (do_vort) curl_out[k] = f4(scl(curl, inv));
    }
    visc_tape[k] = f4(visc);
    const Domain& dom = c.dom;
    bool cxh = false, cyh = false, czh = false;
    if (xp.x <= dom.bmin[0]) { xp.x = dom.bmin[0]; cxh = true; }
    if (xp.x >= dom.bmax[0]) { xp.x = dom.bmax[0]; cxh = true; }
    if (xp.y <= dom.bmin[1]) { xp.y = dom.bmin[1]; cyh = true; }
    if (xp.y >= dom.bmax[1]) { xp.y = dom.bmax[1]; cyh = true; }
    if (xp.z <= dom.bmin[2]) { xp.z = dom.bmin[2]; czh = true; }
    if (xp.z >= dom.bmax[2]) { xp.z = dom.bmax[2]; czh = true; }
    const float inv_dt = (c.sub_dt > 0.0f) ? (1.0f / c.sub_dt) : 0.0f;
    V3 v = scl(sub(xp, old_pos), inv_dt);
    v = add(v, visc_delta);
    const float sp2 = dot(v, v);
    if (sp2 > 2500.0f) v = scl(v, 50.0f / sqrtf(sp2));
    if (cxh) v.x = v.x * -c.damping;
    if (cyh) v.y = v.y * -c.damping;
    if (czh) v.z = v.z * -c.damping;
    pos[id] = make_float4(xp.x, xp.y, xp.z, op4.w);
    vel[id] = make_float4(v.x, v.y, v.z, ov4.w);
    posS[k] = f4(xp);
}

// CS_VorticityConfinement.glsl:133-167 over the list; tape: eta.
__global__ __launch_bounds__(kAB) void k_rec_vorticity(StepConst c, const float4* __restrict__ posS,
                                                       const float4* __restrict__ curl,
                                                       const uint32_t* __restrict__ perm, float4* __restrict__ vel,
                                                       const uint32_t* __restrict__ nbr,
                                                       const uint32_t* __restrict__ cnt, float4* __restrict__ eta_tape) {
    const uint32_t k = agid();
    if (k >= c.n) return;
    const V3 pi = xyz(posS[k]);
    V3 eta = zero3();
    const uint32_t m = cnt[k];
    for (uint32_t r = 0; r < m; ++r) {
        const uint32_t j = nbr[(size_t)r * c.n + k];
        const V3 d = sub(pi, xyz(posS[j]));
        const V3 gw = spiky(c, d, dot(d, d));
        const float om = len(xyz(curl[j]));
        eta = add(eta, scl(gw, om));
    }
    eta = scl(eta, c.inv_rho0);
    eta_tape[k] = f4(eta);
    const float el = len(eta);
    if (el < 1e-6f) return;
    const V3 nn = dvs(eta, el);
    const V3 om = xyz(curl[k]);
    const V3 dv = scl(scl(cross(nn, om), c.vort_eps), c.sub_dt);
    const uint32_t id = perm[k];
    const float4 v4 = vel[id];
    vel[id] = make_float4(v4.x + dv.x, v4.y + dv.y, v4.z + dv.z, v4.w);
}


// ---- FP64 arithmetic --------------------------------------------------------
// The reverse passes compute in FP64.  A pair (i, j) contributes to the
// position adjoint of both ends with nearly opposite terms (one computed by
// slot i, one by slot j); in FP32 the rounding of each large term would
// survive their cancellation.  The FP64 trajectory mode runs the forward in
// FP64 too (kernels *64 below).

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 dmk(double x, double y, double z) { return D3{x, y, z}; }
__device__ __forceinline__ D3 dz3() { return D3{0.0, 0.0, 0.0}; }
__device__ __forceinline__ D3 d3(float4 v) { return D3{(double)v.x, (double)v.y, (double)v.z}; }
__device__ __forceinline__ D3 d3(double4 v) { return D3{v.x, v.y, v.z}; }
__device__ __forceinline__ double4 d4(D3 v, double w = 0.0) { return make_double4(v.x, v.y, v.z, w); }
__device__ __forceinline__ D3 add(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 sub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 scl(D3 a, double s) { return D3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ D3 dvs(D3 a, double s) { return D3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double len(D3 a) { return sqrt(dot(a, a)); }
__device__ __forceinline__ D3 cross(D3 a, D3 b) {
    return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ D3 qsum3(D3 v) { return D3{qsum(v.x), qsum(v.y), qsum(v.z)}; }

// The constants of a step in FP64: the FP32 step constants widened (FP32
// trajectory), or PBFluids::setParams (src/fluid/PBFluids.cpp:44-101,
// Kernels.cpp:9-27) evaluated in FP64 (FP64 trajectory).
struct C64 {
    double g[3], sub_dt, h, rho0, eps, wq, k_corr, n_corr, viscosity, damping, cohesion, w0_self, poly6, spiky,
        inv_rho0, vort_eps;
    double bmin[3], bmax[3];
    uint32_t n, hash_mask, enable_scorr, enable_visc, enable_vort;
};

// The pair functions of CS_ComputeLambdas.glsl:74-92 / CS_ComputeDeltaP.glsl:94-104
// in FP64, with the oracle twin's operation order.
__device__ __forceinline__ double d_poly6(const C64& k, double r2) {
    const double h2 = k.h * k.h;
    if (r2 > h2) return 0.0;
    const double q = h2 - r2;
    return k.poly6 * q * q * q;
}
__device__ __forceinline__ D3 d_poly6_grad(const C64& k, D3 r, double r2) {
    const double h2 = k.h * k.h;
    if (r2 > h2) return dz3();
    const double q = h2 - r2;
    return scl(r, -6.0 * k.poly6 * q * q);
}
__device__ __forceinline__ D3 d_spiky(const C64& k, D3 r, double r2) {
    if (r2 > k.h * k.h || r2 < 1e-10) return dz3();
    const double d = sqrt(r2);
    const double hd = k.h - d;
    return scl(r, -k.spiky * hd * hd / d);
}
// J(r) v for the spiky gradient field g(r) = phi(|r|) r, phi(d) = -K (h - d)^2 / d:
// J = phi I + phi'(d) / d r r^T (symmetric, even in r), zero where g is zero.
__device__ __forceinline__ D3 d_spiky_jvp(const C64& k, D3 r, double r2, D3 v) {
    if (r2 > k.h * k.h || r2 < 1e-10) return dz3();
    const double d = sqrt(r2);
    const double hd = k.h - d;
    const double phi = -k.spiky * hd * hd / d;
    const double dphi = k.spiky * (2.0 * hd / d + hd * hd / (d * d));
    return add(scl(v, phi), scl(r, dphi / d * dot(r, v)));
}
// x^n by binary powering for integer n in [0, 64]; n = 4 and n = 3 (s_corr's
// exponent and its derivative's at the configured n = 4) written out as the
// powering's own products: (x x)(x x), and x (x x) after 1 x = x.
__device__ __forceinline__ double d_powr(double x, double n) {
    if (n == 4.0) {
        const double x2 = x * x;
        return x2 * x2;
    }
    if (n == 3.0) return x * (x * x);
    if (n >= 0.0 && n <= 64.0 && n == floor(n)) {
        uint32_t e = (uint32_t)n;
        double result = 1.0, base = x;
        while (e) {
            if (e & 1u) result = result * base;
            e >>= 1u;
            if (e) base = base * base;
        }
        return result;
    }
    return pow(x, n);
}
__device__ __forceinline__ double d_scorr(const C64& k, double r2) {
    if (k.enable_scorr == 0u) return 0.0;
    const double w = d_poly6(k, r2);
    if (k.wq <= 1e-12 || w <= 0.0) return 0.0;
    const double ratio = w / k.wq;
    if (ratio <= 0.0) return 0.0;
    return -k.k_corr * d_powr(ratio, k.n_corr);
}
// -k n (W/wq)^(n-1) / wq dW/dr
__device__ __forceinline__ D3 d_scorr_grad(const C64& k, D3 r, double r2) {
    if (k.enable_scorr == 0u) return dz3();
    const double w = d_poly6(k, r2);
    if (k.wq <= 1e-12 || w <= 0.0 || k.n_corr == 0.0) return dz3();
    const double p = d_powr(w / k.wq, k.n_corr - 1.0);
    return scl(d_poly6_grad(k, r, r2), -k.k_corr * k.n_corr * p / k.wq);
}

// |v.xyz| in the precision of the tape, as the forward computed it.
__device__ __forceinline__ float slen3(float4 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
__device__ __forceinline__ double slen3(double4 v) { return sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
__device__ __forceinline__ bool memcmp_eq(float4 a, float4 b) {
    return __float_as_uint(a.x) == __float_as_uint(b.x) && __float_as_uint(a.y) == __float_as_uint(b.y) &&
           __float_as_uint(a.z) == __float_as_uint(b.z) && __float_as_uint(a.w) == __float_as_uint(b.w);
}
__device__ __forceinline__ bool memcmp_eq(double4 a, double4 b) {
    return __double_as_longlong(a.x) == __double_as_longlong(b.x) &&
           __double_as_longlong(a.y) == __double_as_longlong(b.y) &&
           __double_as_longlong(a.z) == __double_as_longlong(b.z) &&
           __double_as_longlong(a.w) == __double_as_longlong(b.w);
}

template <class T4>
struct Prec;
template <>
struct Prec<float4> {
    using S = float;
    using T2 = float2;
};
template <>
struct Prec<double4> {
    using S = double;
    using T2 = double2;
};

__global__ __launch_bounds__(kAB) void k_f2d(uint32_t n, const float4* __restrict__ in, double4* __restrict__ out,
                                             int keep_w) {
    const uint32_t i = agid();
    if (i >= n) return;
    const float4 v = in[i];
    out[i] = make_double4(v.x, v.y, v.z, keep_w ? (double)v.w : 0.0);
}

__global__ __launch_bounds__(kAB) void k_d2f(uint32_t n, const double4* __restrict__ in, float4* __restrict__ out,
                                             int keep_w) {
    const uint32_t i = agid();
    if (i >= n) return;
    const double4 v = in[i];
    out[i] = make_float4((float)v.x, (float)v.y, (float)v.z, keep_w ? (float)v.w : 0.0f);
}

__global__ __launch_bounds__(kAB) void k_zero_f64(double* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * kAB + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// ---- FP64 trajectory: forward / recording kernels ----------------------------
// The same passes as the FP32 path (predict, neighbour lists, lambda, delta-p,
// integrate, vorticity) on FP64 state, with the FP64 constants.

// CS_PredictAndHash.glsl:88-130 (no interaction force).
__global__ __launch_bounds__(kAB) void k_predict64(C64 k, const double4* __restrict__ pos, double4* __restrict__ vel,
                                                   double4* __restrict__ xstar, uint32_t* __restrict__ keys) {
    const uint32_t id = agid();
    if (id >= k.n) return;
    const D3 p = d3(pos[id]);
    const double4 v4 = vel[id];
    D3 v = d3(v4);
    v = add(v, scl(dmk(k.g[0], k.g[1], k.g[2]), k.sub_dt));
    const D3 xp = add(p, scl(v, k.sub_dt));
    vel[id] = d4(v, v4.w);
    xstar[id] = d4(xp);
    const int32_t cx = (int32_t)floor(xp.x / k.h), cy = (int32_t)floor(xp.y / k.h), cz = (int32_t)floor(xp.z / k.h);
    keys[id] = cell_hash(cx, cy, cz, k.hash_mask);
}

__global__ __launch_bounds__(kAB) void k_reorder64(uint32_t n, const uint32_t* __restrict__ perm,
                                                   const double4* __restrict__ xstar, const double4* __restrict__ vel,
                                                   double4* __restrict__ P, double4* __restrict__ vs) {
    const uint32_t s = agid();
    if (s >= n) return;
    const uint32_t id = perm[s];
    P[s] = xstar[id];
    vs[s] = vel[id];
}

// Neighbour list over FP64 positions, in the reference's visit order.
__global__ __launch_bounds__(kAB) void k_edges64(C64 k, const double4* __restrict__ X, const int2* __restrict__ cells,
                                                 uint32_t* __restrict__ nbr, uint32_t* __restrict__ cnt,
                                                 uint32_t* __restrict__ tin, uint32_t* __restrict__ tcnt,
                                                 uint32_t* __restrict__ overflow) {
    const uint32_t g = agid(), i = g / kQ, lane = g % kQ;
    const bool live = i < k.n;
    uint32_t m = 0, indeg = 0;
    if (live) {
        const D3 xi = d3(X[i]);
        const double h2 = k.h * k.h;
        const int32_t cx = (int32_t)floor(xi.x / k.h), cy = (int32_t)floor(xi.y / k.h),
                      cz = (int32_t)floor(xi.z / k.h);
        m = quad_list(X, cells, BucketWalk(cx, cy, cz, k.hash_mask, 0u), k.n, i, lane, nbr, tin, tcnt, overflow,
                      indeg, [&](const double4& xk) {
                          const D3 r = sub(xi, d3(xk));
                          return dot(r, r) <= h2;
                      });
    }
    store_count(m, indeg, live, i, lane, cnt, overflow);
}

__global__ __launch_bounds__(kAB) void k_rec_lambda64(C64 k, const double4* __restrict__ P,
                                                      const uint32_t* __restrict__ nbr,
                                                      const uint32_t* __restrict__ cnt, double4* __restrict__ PL,
                                                      double2* __restrict__ lr, double4* __restrict__ aux) {
    const uint32_t i = agid();
    if (i >= k.n) return;
    const D3 xi = d3(P[i]);
    const double inv = k.inv_rho0;
    double rho = k.w0_self;
    double sum_g2 = 0.0;
    D3 gi = dz3();
    walk4(nbr, k.n, i, cnt[i], P, [&](uint32_t, const double4& xj) {
        const D3 d = sub(xi, d3(xj));
        const double r2 = dot(d, d);
        rho = rho + d_poly6(k, r2);
        const D3 gw = d_spiky(k, d, r2);
        const double gj2 = dot(gw, gw);
        sum_g2 = sum_g2 + gj2 * (inv * inv);
        gi = add(gi, gw);
    });
    gi = scl(gi, inv);
    double C = rho / k.rho0 - 1.0;
    const double floorc = -k.cohesion;
    C = C > floorc ? C : floorc;
    sum_g2 = sum_g2 + dot(gi, gi);
    const double D = sum_g2 + k.eps;
    const double lambda = -C / D;
    PL[i] = d4(xi, lambda);
    lr[i] = make_double2(lambda, rho);
    aux[i] = d4(gi, D);
}

__global__ __launch_bounds__(kAB) void k_rec_deltap64(C64 k, const double4* __restrict__ PL,
                                                      const uint32_t* __restrict__ nbr,
                                                      const uint32_t* __restrict__ cnt, double4* __restrict__ P,
                                                      double4* __restrict__ Ptape, double4* __restrict__ dptape) {
    const uint32_t i = agid();
    if (i >= k.n) return;
    const double4 xi4 = PL[i];
    const D3 xi = d3(xi4);
    const double li = xi4.w;
    D3 dp = dz3();
    walk4(nbr, k.n, i, cnt[i], PL, [&](uint32_t, const double4& a) {
        const D3 d = sub(xi, d3(a));
        const double r2 = dot(d, d);
        dp = add(dp, scl(d_spiky(k, d, r2), li + a.w + d_scorr(k, r2)));
    });
    dp = scl(dp, k.inv_rho0);
    dptape[i] = d4(dp);
    const double maxdp = 0.5 * k.h;
    const double l2 = dot(dp, dp);
    if (l2 > maxdp * maxdp) dp = scl(dp, maxdp / sqrt(l2));
    const D3 p = add(xi, dp);
    P[i] = d4(p);
    Ptape[i] = d4(p);
}

__global__ __launch_bounds__(kAB) void k_rec_integrate64(C64 k, const double4* __restrict__ P,
                                                         const double4* __restrict__ vs,
                                                         const uint32_t* __restrict__ perm,
                                                         double4* __restrict__ pos, double4* __restrict__ vel,
                                                         const uint32_t* __restrict__ nbr,
                                                         const uint32_t* __restrict__ cnt,
                                                         double4* __restrict__ curl_out, double4* __restrict__ posS,
                                                         double4* __restrict__ visc_tape) {
    const uint32_t s = agid();
    if (s >= k.n) return;
    const uint32_t id = perm[s];
    const double4 op4 = pos[id];
    const D3 old_pos = d3(op4);
    const double4 ov4 = vs[s];
    const D3 old_vel = d3(ov4);
    D3 xp = d3(P[s]);
    const double inv = k.inv_rho0;
    const bool do_visc = k.enable_visc == 1u;
    const bool do_vort = k.enable_vort == 1u;
    D3 visc = dz3(), visc_delta = dz3();
    if (do_visc || do_vort) {
        D3 curl = dz3();
        const uint32_t m = cnt[s];
        for (uint32_t r = 0; r < m; ++r) {
            const uint32_t j = nbr[(size_t)r * k.n + s];
            const D3 d = sub(xp, d3(P[j]));
            const double r2 = dot(d, d);
            const D3 vij = sub(d3(vs[j]), old_vel);
            if (do_visc) visc = add(visc, scl(scl(vij, d_poly6(k, r2)), inv));
            if (do_vort) curl = add(curl, cross(vij, d_spiky(k, d, r2)));
        }
        if (do_visc) visc_delta = scl(visc, k.viscosity);
        if (do_vort) curl_out[s] = d4(scl(curl, inv));
    }
    visc_tape[s] = d4(visc);
    bool cxh = false, cyh = false, czh = false;
    if (xp.x <= k.bmin[0]) { xp.x = k.bmin[0]; cxh = true; }
    if (xp.x >= k.bmax[0]) { xp.x = k.bmax[0]; cxh = true; }
    if (xp.y <= k.bmin[1]) { xp.y = k.bmin[1]; cyh = true; }
    if (xp.y >= k.bmax[1]) { xp.y = k.bmax[1]; cyh = true; }
    if (xp.z <= k.bmin[2]) { xp.z = k.bmin[2]; czh = true; }
    if (xp.z >= k.bmax[2]) { xp.z = k.bmax[2]; czh = true; }
    const double inv_dt = (k.sub_dt > 0.0) ? (1.0 / k.sub_dt) : 0.0;
    D3 v = scl(sub(xp, old_pos), inv_dt);
    v = add(v, visc_delta);
    const double sp2 = dot(v, v);
    if (sp2 > 2500.0) v = scl(v, 50.0 / sqrt(sp2));
    if (cxh) v.x = v.x * -k.damping;
    if (cyh) v.y = v.y * -k.damping;
    if (czh) v.z = v.z * -k.damping;
    pos[id] = d4(xp, op4.w);
    vel[id] = d4(v, ov4.w);
    posS[s] = d4(xp);
}

__global__ __launch_bounds__(kAB) void k_rec_vorticity64(C64 k, const double4* __restrict__ posS,
                                                         const double4* __restrict__ curl,
                                                         const uint32_t* __restrict__ perm,
                                                         double4* __restrict__ vel, const uint32_t* __restrict__ nbr,
                                                         const uint32_t* __restrict__ cnt,
                                                         double4* __restrict__ eta_tape) {
    const uint32_t s = agid();
    if (s >= k.n) return;
    const D3 pi = d3(posS[s]);
    D3 eta = dz3();
    const uint32_t m = cnt[s];
    for (uint32_t r = 0; r < m; ++r) {
        const uint32_t j = nbr[(size_t)r * k.n + s];
        const D3 d = sub(pi, d3(posS[j]));
        eta = add(eta, scl(d_spiky(k, d, dot(d, d)), len(d3(curl[j]))));
    }
    eta = scl(eta, k.inv_rho0);
    eta_tape[s] = d4(eta);
    const double el = len(eta);
    if (el < 1e-6) return;
    const D3 nn = dvs(eta, el);
    const D3 dv = scl(scl(cross(nn, d3(curl[s])), k.vort_eps), k.sub_dt);
    const uint32_t id = perm[s];
    const double4 v4 = vel[id];
    vel[id] = make_double4(v4.x + dv.x, v4.y + dv.y, v4.z + dv.z, v4.w);
}

// ---- reverse passes -----------------------------------------------------------
// Templates over the tape type: float4 (FP32 trajectory) or double4 (FP64).
// Branch decisions (wall flags, speed clamp, delta-p clamp, cohesion floor,
// vorticity threshold) are re-evaluated in the tape's precision exactly as
// the forward took them; the derivative is that of the taken branch.

// Vorticity confinement, slot i's own terms: v_i += a (n x w_i), a = eps_v dt,
// n = eta / |eta|, eta = inv sum_j |w_j| g(p_i - p_j).
template <class T4>
__global__ __launch_bounds__(kAB) void k_rev_vort(C64 k, const T4* __restrict__ posS, const T4* __restrict__ curl,
                                                  const T4* __restrict__ eta_tape, const uint32_t* __restrict__ perm,
                                                  const double4* __restrict__ gvel, const uint32_t* __restrict__ nbr,
                                                  const uint32_t* __restrict__ cnt, double4* __restrict__ gS_own,
                                                  double4* __restrict__ gw_own, double4* __restrict__ ep,
                                                  double* __restrict__ es) {
    using S = typename Prec<T4>::S;
    const uint32_t g = agid(), i = g / kQ, lane = g % kQ;
    if (i >= k.n) return;
    const uint32_t m = cnt[i];
    const T4 e4 = eta_tape[i];
    if (slen3(e4) < S(1e-6)) {  // the forward's test
        for (uint32_t r = lane; r < m; r += kQ) {
            ep[(size_t)r * k.n + i] = make_double4(0.0, 0.0, 0.0, 0.0);
            es[(size_t)r * k.n + i] = 0.0;
        }
        if (lane == 0) {
            gS_own[i] = make_double4(0.0, 0.0, 0.0, 0.0);
            gw_own[i] = make_double4(0.0, 0.0, 0.0, 0.0);
        }
        return;
    }
    const D3 eta = d3(e4);
    const double el = len(eta);
    const double inv = k.inv_rho0;
    const double a = k.vort_eps * k.sub_dt;
    const D3 nn = scl(eta, 1.0 / el);
    const D3 om = d3(curl[i]);
    const D3 vb = d3(gvel[perm[i]]);
    const D3 nb = scl(cross(om, vb), a);
    if (lane == 0) gw_own[i] = d4(scl(cross(vb, nn), a));
    const D3 etab = scl(sub(nb, scl(nn, dot(nn, nb))), 1.0 / el);
    const D3 pi = d3(posS[i]);
    D3 gs = dz3();
    for (uint32_t r = lane; r < m; r += kQ) {
        const size_t e = (size_t)r * k.n + i;
        const uint32_t j = nbr[e];
        const D3 d = sub(pi, d3(posS[j]));
        const double r2 = dot(d, d);
        const double mj = len(d3(curl[j]));
        const D3 rb = scl(d_spiky_jvp(k, d, r2, etab), inv * mj);
        gs = add(gs, rb);
        ep[e] = d4(rb);
        es[e] = inv * dot(d_spiky(k, d, r2), etab);
    }
    gs = qsum3(gs);
    if (lane == 0) gS_own[i] = d4(gs);
}

// Vorticity, the terms of the pairs reaching j: position adjoint into the
// id-order state adjoint, |w_j| adjoint into w_j.
template <class T4>
__global__ __launch_bounds__(kAB) void k_rev_vort_t(uint32_t n, const T4* __restrict__ curl,
                                                    const uint32_t* __restrict__ perm, const uint32_t* __restrict__ tin,
                                                    const uint32_t* __restrict__ tcnt, const double4* __restrict__ ep,
                                                    const double* __restrict__ es, const double4* __restrict__ gS_own,
                                                    const double4* __restrict__ gw_own, double4* __restrict__ gpos,
                                                    double4* __restrict__ gw) {
    const uint32_t g = agid(), j = g / kR, lane = g % kR;
    if (j >= n) return;
    double r[4];
    row_sum<4>(tin, n, j, min(tcnt[j], kInCap), lane, [&](uint32_t e, double* v) {
        const double4 a = ep[e];
        v[0] = a.x;
        v[1] = a.y;
        v[2] = a.z;
        v[3] = es[e];
    }, r);
    if (lane != 0) return;
    const D3 gj = sub(d3(gS_own[j]), dmk(r[0], r[1], r[2]));
    const uint32_t id = perm[j];
    gpos[id] = d4(add(d3(gpos[id]), gj));
    const D3 wj = d3(curl[j]);
    const double mj = len(wj);
    D3 w = d3(gw_own[j]);
    if (mj > 0.0) w = add(w, scl(wj, r[3] / mj));
    gw[j] = d4(w);
}

// Integrate, slot s's own terms.  gpos / gvel: adjoint of the substep's output
// state (id order); gw: adjoint of the curl (null when no later pass reads it).
template <class T4>
__global__ __launch_bounds__(kAB) void k_rev_integrate(C64 k, const T4* __restrict__ xold, const T4* __restrict__ vs,
                                                       const T4* __restrict__ P, const T4* __restrict__ visc_tape,
                                                       const uint32_t* __restrict__ perm,
                                                       const uint32_t* __restrict__ nbr,
                                                       const uint32_t* __restrict__ cnt,
                                                       const double4* __restrict__ gpos,
                                                       const double4* __restrict__ gvel,
                                                       const double4* __restrict__ gw, double4* __restrict__ gP_own,
                                                       double4* __restrict__ gvs_own, double4* __restrict__ gxold,
                                                       double* __restrict__ acc_visc, double4* __restrict__ ep,
                                                       double4* __restrict__ ev) {
    using S = typename Prec<T4>::S;
    const uint32_t g = agid(), s = g / kQ, lane = g % kQ;
    if (s >= k.n) return;
    const uint32_t id = perm[s];
    // the forward's decisions, in its precision: wall flags and the speed clamp
    const T4 p4 = P[s];
    S xpx = p4.x, xpy = p4.y, xpz = p4.z;
    bool cxh = false, cyh = false, czh = false;
    if (xpx <= S(k.bmin[0])) { xpx = S(k.bmin[0]); cxh = true; }
    if (xpx >= S(k.bmax[0])) { xpx = S(k.bmax[0]); cxh = true; }
    if (xpy <= S(k.bmin[1])) { xpy = S(k.bmin[1]); cyh = true; }
    if (xpy >= S(k.bmax[1])) { xpy = S(k.bmax[1]); cyh = true; }
    if (xpz <= S(k.bmin[2])) { xpz = S(k.bmin[2]); czh = true; }
    if (xpz >= S(k.bmax[2])) { xpz = S(k.bmax[2]); czh = true; }
    const bool do_visc = k.enable_visc == 1u;
    const bool do_vort = k.enable_vort == 1u;
    const S inv_dt_s = (S(k.sub_dt) > S(0)) ? (S(1) / S(k.sub_dt)) : S(0);
    const T4 xo = xold[id], vt = visc_tape[s];
    S vx = (xpx - xo.x) * inv_dt_s, vy = (xpy - xo.y) * inv_dt_s, vz = (xpz - xo.z) * inv_dt_s;
    if (do_visc) {
        vx = vx + vt.x * S(k.viscosity);
        vy = vy + vt.y * S(k.viscosity);
        vz = vz + vt.z * S(k.viscosity);
    }
    const S sp2 = vx * vx + vy * vy + vz * vz;
    const double inv_dt = (double)inv_dt_s;
    D3 gv = d3(gvel[id]);
    if (cxh) gv.x = gv.x * -k.damping;
    if (cyh) gv.y = gv.y * -k.damping;
    if (czh) gv.z = gv.z * -k.damping;
    if (sp2 > S(2500)) {
        const D3 v = dmk(vx, vy, vz);
        const double l = len(v);
        const D3 vh = scl(v, 1.0 / l);
        gv = scl(sub(gv, scl(vh, dot(vh, gv))), 50.0 / l);
    }
    const D3 visc = d3(vt);
    D3 vb = dz3();
    if (do_visc) {
        if (lane == 0) acc_visc[id] += dot(gv, visc);
        vb = scl(gv, k.viscosity);
    }
    const D3 gxp = add(d3(gpos[id]), scl(gv, inv_dt));
    if (lane == 0) gxold[id] = d4(scl(gv, -inv_dt));
    // the slot's own term (wall-masked gxp) enters the quad's sum once, in lane 0
    D3 gp = lane == 0 ? dmk(cxh ? 0.0 : gxp.x, cyh ? 0.0 : gxp.y, czh ? 0.0 : gxp.z) : dz3();
    const D3 wb = (do_vort && gw) ? d3(gw[s]) : dz3();
    D3 gvi = dz3();
    const double inv = k.inv_rho0;
    const uint32_t m = cnt[s];
    if (do_visc || do_vort) {
        const D3 p = d3(p4);
        const D3 vi = d3(vs[s]);
        for (uint32_t r = lane; r < m; r += kQ) {
            const size_t e = (size_t)r * k.n + s;
            const uint32_t j = nbr[e];
            const D3 d = sub(p, d3(P[j]));
            const double r2 = dot(d, d);
            const D3 dv = sub(d3(vs[j]), vi);
            D3 dvb = dz3();
            D3 rb = dz3();
            if (do_visc) {
                dvb = scl(vb, inv * d_poly6(k, r2));
                rb = scl(d_poly6_grad(k, d, r2), inv * dot(dv, vb));
            }
            if (do_vort) {
                dvb = add(dvb, scl(cross(d_spiky(k, d, r2), wb), inv));
                rb = add(rb, d_spiky_jvp(k, d, r2, scl(cross(wb, dv), inv)));
            }
            gvi = sub(gvi, dvb);
            gp = add(gp, rb);
            ev[e] = d4(dvb);
            ep[e] = d4(rb);
        }
    } else {
        for (uint32_t r = lane; r < m; r += kQ) {
            ep[(size_t)r * k.n + s] = make_double4(0.0, 0.0, 0.0, 0.0);
            ev[(size_t)r * k.n + s] = make_double4(0.0, 0.0, 0.0, 0.0);
        }
    }
    gp = qsum3(gp);
    gvi = qsum3(gvi);
    if (lane == 0) {
        gP_own[s] = d4(gp);
        gvs_own[s] = d4(gvi);
    }
}

__global__ __launch_bounds__(kAB) void k_rev_pairs_vec2(uint32_t n, const uint32_t* __restrict__ tin,
                                                        const uint32_t* __restrict__ tcnt,
                                                        const double4* __restrict__ ep, const double4* __restrict__ ev,
                                                        const double4* __restrict__ gP_own,
                                                        const double4* __restrict__ gvs_own, double4* __restrict__ gP,
                                                        double4* __restrict__ gvs) {
    const uint32_t g = agid(), j = g / kR, lane = g % kR;
    if (j >= n) return;
    double r[6];
    row_sum<6>(tin, n, j, min(tcnt[j], kInCap), lane, [&](uint32_t e, double* v) {
        const double4 a = ep[e], b = ev[e];
        v[0] = a.x;
        v[1] = a.y;
        v[2] = a.z;
        v[3] = b.x;
        v[4] = b.y;
        v[5] = b.z;
    }, r);
    if (lane != 0) return;
    gP[j] = d4(sub(d3(gP_own[j]), dmk(r[0], r[1], r[2])));
    gvs[j] = d4(add(d3(gvs_own[j]), dmk(r[3], r[4], r[5])));
}

// Clamp-and-add of CS_ApplyDeltaP.glsl:80-106 and delta-p
// (CS_ComputeDeltaP.glsl:120-166) of one iteration, slot i's own terms.
// gP: adjoint of the corrected positions; out: adjoint of the positions the
// iteration started from (own terms) and of lambda.
template <class T4>
__global__ __launch_bounds__(kAB) void k_rev_deltap(C64 k, const T4* __restrict__ Pprev,
                                                    const typename Prec<T4>::T2* __restrict__ lr,
                                                    const T4* __restrict__ dptape, const uint32_t* __restrict__ nbr,
                                                    const uint32_t* __restrict__ cnt, const double4* __restrict__ gP,
                                                    double4* __restrict__ gQ_own, double* __restrict__ gl_own,
                                                    double4* __restrict__ ep, double* __restrict__ es) {
    using S = typename Prec<T4>::S;
    const uint32_t gid = agid(), i = gid / kQ, lane = gid % kQ;
    if (i >= k.n) return;
    const D3 g = d3(gP[i]);
    const T4 dp4 = dptape[i];
    const S maxdp = S(0.5) * S(k.h);
    D3 gdp = g;
    if (dp4.x * dp4.x + dp4.y * dp4.y + dp4.z * dp4.z > maxdp * maxdp) {  // the forward's clamp decision
        const D3 dp = d3(dp4);
        const double l = len(dp);
        const D3 dh = scl(dp, 1.0 / l);
        gdp = scl(sub(g, scl(dh, dot(dh, g))), (double)maxdp / l);
    }
    const double inv = k.inv_rho0;
    const D3 xi = d3(Pprev[i]);
    const double li = (double)lr[i].x;
    D3 gq = lane == 0 ? g : dz3();  // the identity term once per quad
    double gl = 0.0;
    const uint32_t m = cnt[i];
    for (uint32_t r = lane; r < m; r += kQ) {
        const size_t e = (size_t)r * k.n + i;
        const uint32_t j = nbr[e];
        const D3 d = sub(xi, d3(Pprev[j]));
        const double r2 = dot(d, d);
        const double a = li + (double)lr[j].x + d_scorr(k, r2);
        const double t = inv * dot(d_spiky(k, d, r2), gdp);
        gl += t;
        const D3 rb = add(scl(d_spiky_jvp(k, d, r2, gdp), inv * a), scl(d_scorr_grad(k, d, r2), t));
        gq = add(gq, rb);
        ep[e] = d4(rb);
        es[e] = t;
    }
    gq = qsum3(gq);
    gl = qsum(gl);
    if (lane != 0) return;
    gQ_own[i] = d4(gq);
    gl_own[i] = gl;
}

__global__ __launch_bounds__(kAB) void k_rev_pairs_vec_scalar(uint32_t n, const uint32_t* __restrict__ tin,
                                                              const uint32_t* __restrict__ tcnt,
                                                              const double4* __restrict__ ep,
                                                              const double* __restrict__ es,
                                                              const double4* __restrict__ gQ_own,
                                                              const double* __restrict__ gl_own,
                                                              double4* __restrict__ gQ, double* __restrict__ gl) {
    const uint32_t g = agid(), j = g / kR, lane = g % kR;
    if (j >= n) return;
    double r[4];
    row_sum<4>(tin, n, j, min(tcnt[j], kInCap), lane, [&](uint32_t e, double* v) {
        const double4 a = ep[e];
        v[0] = a.x;
        v[1] = a.y;
        v[2] = a.z;
        v[3] = es[e];
    }, r);
    if (lane != 0) return;
    gQ[j] = d4(sub(d3(gQ_own[j]), dmk(r[0], r[1], r[2])));
    gl[j] = gl_own[j] + r[3];
}

// Lambda (CS_ComputeLambdas.glsl:108-162), slot i's own terms, added to gQ.
template <class T4>
__global__ __launch_bounds__(kAB) void k_rev_lambda(C64 k, const T4* __restrict__ Pprev,
                                                    const typename Prec<T4>::T2* __restrict__ lr,
                                                    const T4* __restrict__ aux, const uint32_t* __restrict__ perm,
                                                    const uint32_t* __restrict__ nbr, const uint32_t* __restrict__ cnt,
                                                    const double* __restrict__ gl, double4* __restrict__ gQ,
                                                    double* __restrict__ acc_eps, double4* __restrict__ ep) {
    using S = typename Prec<T4>::S;
    const uint32_t g = agid(), i = g / kQ, lane = g % kQ;
    if (i >= k.n) return;
    const double lb = gl[i];
    const S rho_s = lr[i].y;
    const T4 a4 = aux[i];
    const D3 gi = d3(a4);
    const double D = (double)a4.w;
    // the forward's cohesion-floor decision and clamped value, in its precision
    const S craw = rho_s / S(k.rho0) - S(1);
    const bool active = craw > -S(k.cohesion);
    const double C = active ? (double)craw : -(double)S(k.cohesion);
    const double Cb = -lb / D;
    const double Db = lb * C / (D * D);
    if (lane == 0) acc_eps[perm[i]] += Db;
    const double inv = k.inv_rho0;
    const double rhob = active ? Cb / k.rho0 : 0.0;
    const double sb = 2.0 * inv * inv * Db;
    const D3 Gb = scl(gi, 2.0 * Db * inv);
    const D3 xi = d3(Pprev[i]);
    D3 acc = dz3();
    const uint32_t m = cnt[i];
    for (uint32_t r = lane; r < m; r += kQ) {
        const size_t e = (size_t)r * k.n + i;
        const uint32_t j = nbr[e];
        const D3 d = sub(xi, d3(Pprev[j]));
        const double r2 = dot(d, d);
        const D3 gw = d_spiky(k, d, r2);
        const D3 rb = add(scl(d_poly6_grad(k, d, r2), rhob), d_spiky_jvp(k, d, r2, add(scl(gw, sb), Gb)));
        acc = add(acc, rb);
        ep[e] = d4(rb);
    }
    acc = qsum3(acc);
    if (lane == 0) gQ[i] = d4(add(d3(gQ[i]), acc));
}

__global__ __launch_bounds__(kAB) void k_rev_pairs_sub(uint32_t n, const uint32_t* __restrict__ tin,
                                                       const uint32_t* __restrict__ tcnt,
                                                       const double4* __restrict__ ep, double4* __restrict__ gQ) {
    const uint32_t g = agid(), j = g / kR, lane = g % kR;
    if (j >= n) return;
    double r[3];
    row_sum<3>(tin, n, j, min(tcnt[j], kInCap), lane, [&](uint32_t e, double* v) {
        const double4 a = ep[e];
        v[0] = a.x;
        v[1] = a.y;
        v[2] = a.z;
    }, r);
    if (lane == 0) gQ[j] = d4(sub(d3(gQ[j]), dmk(r[0], r[1], r[2])));
}

// Predict (CS_PredictAndHash.glsl:88-130 without the interaction force):
// x* = x + v* dt, v* = v + g dt.
__global__ __launch_bounds__(kAB) void k_rev_predict(C64 k, const uint32_t* __restrict__ perm,
                                                     const double4* __restrict__ gP0, const double4* __restrict__ gvs,
                                                     const double4* __restrict__ gxold, double4* __restrict__ gpos,
                                                     double4* __restrict__ gvel) {
    const uint32_t s = agid();
    if (s >= k.n) return;
    const uint32_t id = perm[s];
    const D3 gx = d3(gP0[s]);
    gpos[id] = d4(add(gx, d3(gxold[id])));
    gvel[id] = d4(add(d3(gvs[s]), scl(gx, k.sub_dt)));
}

__global__ __launch_bounds__(kAB) void k_reduce2(uint32_t n, const double* __restrict__ a,
                                                 const double* __restrict__ b, double* __restrict__ out) {
    __shared__ double sa[kAB], sb[kAB];
    double x = 0.0, y = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += kAB) {
        x += a[i];
        y += b[i];
    }
    sa[threadIdx.x] = x;
    sb[threadIdx.x] = y;
    __syncthreads();
    for (int s = kAB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sa[threadIdx.x] += sa[threadIdx.x + s];
            sb[threadIdx.x] += sb[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sa[0];
        out[1] = sb[0];
    }
}

// Largest |component| of the state adjoint, rounded to FP32 (as ordered bits
// of non-negative floats; a NaN orders above +inf), into the kSpread slots of
// out (position) and out + kSpread (velocity).
__global__ __launch_bounds__(kAB) void k_absmax(uint32_t n, const double4* __restrict__ gpos,
                                                const double4* __restrict__ gvel, uint32_t* __restrict__ out) {
    const uint32_t i = agid();
    uint32_t bp = 0u, bq = 0u;
    if (i < n) {
        const double4 p = gpos[i], v = gvel[i];
        const float m = (float)fmax(fmax(fabs(p.x), fabs(p.y)), fabs(p.z));
        const float q = (float)fmax(fmax(fabs(v.x), fabs(v.y)), fabs(v.z));
        bp = __float_as_uint(m);
        bq = __float_as_uint(q);
        if (p.x != p.x || p.y != p.y || p.z != p.z) bp = 0x7fc00000u;
        if (v.x != v.x || v.y != v.y || v.z != v.z) bq = 0x7fc00000u;
    }
    block_max<2>({bp, bq}, {out, out + kSpread});
}

// Bitwise comparison of the replayed state with the forward's checkpoint.
template <class T4>
__global__ __launch_bounds__(kAB) void k_compare(uint32_t n, const T4* __restrict__ a, const T4* __restrict__ b,
                                                 uint32_t* __restrict__ count) {
    const uint32_t i = agid();
    if (i >= n) return;
    const T4 x = a[i], y = b[i];
    const bool same = memcmp_eq(x, y);
    if (!same) atomicAdd(count, 1u);
}

struct EdgeList {
    uint32_t* nbr = nullptr;   // kEdgeCap x n
    uint32_t* cnt = nullptr;   // n
    uint32_t* tin = nullptr;   // kInCap x n
    uint32_t* tcnt = nullptr;  // n
};

// The recorded trajectory in one precision: checkpoints of every frame start
// and the tape of the frame being reversed.  The FP64 trajectory also keeps
// its working state here (the FP32 one runs in the core's buffers).
template <class T4>
struct Tape {
    using T2 = typename Prec<T4>::T2;
    T4* ck_pos = nullptr;  // (frames + 1) x n; the last is the forward's result
    T4* ck_vel = nullptr;
    T4* xold = nullptr;    // S x n, id order
    T4* vs = nullptr;      // S x n
    T4* P = nullptr;       // S x (I + 1) x n
    T2* lr = nullptr;      // S x I x n
    T4* aux = nullptr;     // S x I x n
    T4* dp = nullptr;      // S x I x n
    T4* visc = nullptr;    // S x n
    T4* eta = nullptr;     // n
    T4 *pos = nullptr, *vel = nullptr, *xstar = nullptr, *Pw = nullptr, *PL = nullptr, *vsw = nullptr,
       *curl = nullptr, *posS = nullptr;
    bool ready = false;

    hipError_t alloc(size_t n, size_t frames, size_t S, size_t I, bool work) {
        hipError_t e = hipSuccess;
#define ALLOC(p, cnt) \
    if ((e = dev_alloc(&(p), (cnt))) != hipSuccess) return e;
        ALLOC(ck_pos, (frames + 1) * n);
        ALLOC(ck_vel, (frames + 1) * n);
        ALLOC(xold, S * n);
        ALLOC(vs, S * n);
        ALLOC(P, S * (I + 1) * n);
        ALLOC(lr, S * I * n);
        ALLOC(aux, S * I * n);
        ALLOC(dp, S * I * n);
        ALLOC(visc, S * n);
        ALLOC(eta, n);
        if (work)
            for (T4** p : {&pos, &vel, &xstar, &Pw, &PL, &vsw, &curl, &posS}) ALLOC(*p, n);
#undef ALLOC
        ready = true;
        return hipSuccess;
    }
    void release() {
        for (T4** p : {&ck_pos, &ck_vel, &xold, &vs, &P, &aux, &dp, &visc, &eta, &pos, &vel, &xstar, &Pw, &PL, &vsw,
                       &curl, &posS})
            dev_free(p);
        dev_free(&lr);
        ready = false;
    }
};

}  // namespace

struct fsi_adjoint {
    Core core;
    fsi_fluid_config cfg{};
    fsi_pbf_constants k{};
    uint32_t n = 0, frames = 0, S = 0, I = 0;
    uint32_t precision = 32;  // of the trajectory: 32 (the solver's kernels) or 64
    double visc64 = 0.0, eps64 = 0.0;  // the parameters of the FP64 trajectory
    bool forward_done = false;
    Tape<float4> t32;
    Tape<double4> t64;
    // frames x S x n: the sorted permutation of every substep of the forward,
    // which the replay takes instead of sorting the same keys again
    uint32_t* fwd_perm = nullptr;
    std::vector<EdgeList> lists;  // S x (I + 1) + 1
    uint32_t* list_mem = nullptr;
    // adjoint work (FP64)
    double4 *gpos = nullptr, *gvel = nullptr, *gxold = nullptr;
    double4 *gP = nullptr, *gQ = nullptr, *gown = nullptr, *gvs = nullptr, *gvs_own = nullptr;
    double4 *gw = nullptr, *gw_own = nullptr;
    double *gl = nullptr, *gl_own = nullptr, *acc_visc = nullptr, *acc_eps = nullptr;
    double4 *ep = nullptr, *ev = nullptr;
    double* es = nullptr;
    // kFlagWords: [0] list overflow bits, [1] replayed records that differ
    // from the forward, then the slots of the largest neighbour count and of
    // the largest in-degree of any list
    uint32_t* flags = nullptr;
    uint32_t stats[4] = {0, 0, 0, 0};  // flags of the last backward (reduce_flags)
    uint32_t* norms = nullptr;          // frames x 2 x kSpread: max |adjoint| of pos, vel at each frame start
    std::vector<uint32_t> host_norms;   // frames x 2
    std::vector<uint32_t> host_words;   // staging for flags and norms
    double* dred = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
};

namespace {

int bind(fsi_adjoint* a) {
    if (!a) return fail(FSI_ERR_INVALID_ARG, "handle is NULL");
    FSI_HIP_TRY(hipSetDevice(a->core.device));
    return FSI_OK;
}

void free_all(fsi_adjoint* a) {
    a->t32.release();
    a->t64.release();
    dev_free(&a->fwd_perm);
    dev_free(&a->list_mem);
    for (double4** p : {&a->gpos, &a->gvel, &a->gxold, &a->gP, &a->gQ, &a->gown, &a->gvs, &a->gvs_own, &a->gw,
                        &a->gw_own, &a->ep, &a->ev})
        dev_free(p);
    for (double** p : {&a->gl, &a->gl_own, &a->acc_visc, &a->acc_eps, &a->es}) dev_free(p);
    dev_free(&a->flags);
    dev_free(&a->norms);
    dev_free(&a->dred);
}

StepConst step_const(const fsi_adjoint* a) {
    const float zero[3] = {0, 0, 0};
    StepConst c = step_const_common(a->cfg, a->k, a->n, false, zero, zero);
    Domain& d = c.dom;
    for (int x = 0; x < 3; ++x) {
        d.bmin[x] = a->k.bounds_min[x];
        d.bmax[x] = a->k.bounds_max[x];
    }
    d.hash_mask = a->k.hash_mask;
    d.hash_base = 0;
    set_binning(d, a->k.bounds_min, a->k.bounds_max, a->k.h, 7, 0, false);
    return c;
}

// The step constants in FP64.  FP32 trajectory: the FP32 constants widened,
// so that every decision of the reverse passes is the forward's.  FP64
// trajectory: the derived coefficients of PBFluids::setParams
// (src/fluid/PBFluids.cpp:44-101; wq from Kernels.cpp:9-27) evaluated in FP64
// from the configuration.
C64 c64_of(const fsi_adjoint* a, const StepConst& c) {
    C64 k{};
    for (int x = 0; x < 3; ++x) {
        k.g[x] = c.g[x];
        k.bmin[x] = c.dom.bmin[x];
        k.bmax[x] = c.dom.bmax[x];
    }
    k.sub_dt = c.sub_dt;
    k.h = c.h;
    k.rho0 = c.rho0;
    k.eps = c.eps;
    k.wq = c.wq;
    k.k_corr = c.k_corr;
    k.n_corr = c.n_corr;
    k.viscosity = c.viscosity;
    k.damping = c.damping;
    k.cohesion = c.cohesion;
    k.w0_self = c.w0_self;
    k.poly6 = c.poly6;
    k.spiky = c.spiky;
    k.inv_rho0 = c.inv_rho0;
    k.vort_eps = c.vort_eps;
    k.n = c.n;
    k.hash_mask = c.dom.hash_mask;
    k.enable_scorr = c.enable_scorr;
    k.enable_visc = c.enable_visc;
    k.enable_vort = c.enable_vort;
    if (a->precision == 64) {
        k.viscosity = a->visc64;
        k.eps = a->eps64;
        const double PI = 3.14159265358979323846;
        const double h = a->cfg.h;
        const double h2 = h * h, h3 = h2 * h, h6 = h3 * h3, h9 = h6 * h3;
        k.sub_dt = (double)a->cfg.dt / (double)a->cfg.substep_iterations;
        k.w0_self = (315.0 / (64.0 * PI)) / h3;
        k.poly6 = 315.0 / (64.0 * PI * h9);
        k.spiky = 45.0 / (PI * h6);
        k.inv_rho0 = 1.0 / (double)a->cfg.rho0;
        const double dq = (double)a->cfg.delta_q * h;
        const double r2 = dq * dq;
        if (r2 > h2) {
            k.wq = 0.0;
        } else {
            const double h4 = h2 * h2, diff = h2 - r2;
            k.wq = (315.0 / (64.0 * PI) / (h4 * h4 * h)) * (diff * diff * diff);
        }
    }
    return k;
}

hipError_t alloc_common(fsi_adjoint* a) {
    const size_t n = a->n, S = a->S, I = a->I;
    hipError_t e = hipSuccess;
#define ALLOC(p, cnt) \
    if ((e = dev_alloc(&(p), (cnt))) != hipSuccess) return e;
    ALLOC(a->fwd_perm, (size_t)a->frames * S * n);
    const size_t nl = S * (I + 1) + 1;
    const size_t per = ((size_t)kEdgeCap + kInCap + 2) * n;
    ALLOC(a->list_mem, nl * per);
    a->lists.resize(nl);
    for (size_t l = 0; l < nl; ++l) {
        uint32_t* base = a->list_mem + l * per;
        a->lists[l].nbr = base;
        a->lists[l].tin = base + (size_t)kEdgeCap * n;
        a->lists[l].cnt = base + ((size_t)kEdgeCap + kInCap) * n;
        a->lists[l].tcnt = base + ((size_t)kEdgeCap + kInCap + 1) * n;
    }
    for (double4** p : {&a->gpos, &a->gvel, &a->gxold, &a->gP, &a->gQ, &a->gown, &a->gvs, &a->gvs_own, &a->gw,
                        &a->gw_own})
        ALLOC(*p, n);
    for (double** p : {&a->gl, &a->gl_own, &a->acc_visc, &a->acc_eps}) ALLOC(*p, n);
    ALLOC(a->ep, (size_t)kEdgeCap * n);
    ALLOC(a->ev, (size_t)kEdgeCap * n);
    ALLOC(a->es, (size_t)kEdgeCap * n);
    ALLOC(a->flags, kFlagWords);
    ALLOC(a->norms, 2 * kSpread * (size_t)a->frames);
    ALLOC(a->dred, 2);
#undef ALLOC
    return hipSuccess;
}

// The flag words as (overflow bits, replay mismatches, largest out-degree,
// largest in-degree), the maxima reduced over their slots.
void reduce_flags(const uint32_t* w, uint32_t out[4]) {
    out[0] = w[0];
    out[1] = w[1];
    out[2] = out[3] = 0u;
    for (uint32_t s = 0; s < kSpread; ++s) {
        out[2] = std::max(out[2], w[kOutMax + s]);
        out[3] = std::max(out[3], w[kInMax + s]);
    }
}

// Stream order with the caller: our work waits for the caller's stream, and
// the caller's stream waits for our work.  NULL is the null stream, which the
// handle's non-blocking stream does not follow by itself.
hipError_t join_in(fsi_adjoint* a, void* stream) {
    hipError_t e = hipEventRecord(a->ev_in, (hipStream_t)stream);
    if (e != hipSuccess) return e;
    return hipStreamWaitEvent(a->core.stream, a->ev_in, 0);
}
hipError_t join_out(fsi_adjoint* a, void* stream) {
    hipError_t e = hipEventRecord(a->ev_out, a->core.stream);
    if (e != hipSuccess) return e;
    return hipStreamWaitEvent((hipStream_t)stream, a->ev_out, 0);
}

// The in-degree counters of every list, before a frame's list builds.
__global__ __launch_bounds__(kAB) void k_zero_tcnt(uint32_t* __restrict__ base, size_t stride, uint32_t n) {
    const uint32_t i = agid();
    if (i < n) base[(size_t)blockIdx.y * stride + i] = 0u;
}

hipError_t zero_counters(fsi_adjoint* a) {
    const size_t per = ((size_t)kEdgeCap + kInCap + 2) * a->n;
    uint32_t* base = a->list_mem + ((size_t)kEdgeCap + kInCap + 1) * a->n;
    hipLaunchKernelGGL(k_zero_tcnt, dim3((a->n + kAB - 1) / kAB, (unsigned)a->lists.size()), dim3(kAB), 0,
                       a->core.stream, base, per, a->n);
    return hipGetLastError();
}

// The list L over the positions X and its transposed list (whose row
// counters zero_counters cleared for the frame).
hipError_t build_list(fsi_adjoint* a, const StepConst& c, const EdgeList& L, const float4* X) {
    hipLaunchKernelGGL(k_edges, qgrid(a->n), dim3(kAB), 0, a->core.stream, c, X, a->core.d.cells, L.nbr, L.cnt,
                       L.tin, L.tcnt, a->flags);
    return hipGetLastError();
}

hipError_t build_list(fsi_adjoint* a, const C64& k, const EdgeList& L, const double4* X) {
    hipLaunchKernelGGL(k_edges64, qgrid(a->n), dim3(kAB), 0, a->core.stream, k, X, a->core.d.cells, L.nbr, L.cnt,
                       L.tin, L.tcnt, a->flags);
    return hipGetLastError();
}

const EdgeList& list_of(const fsi_adjoint* a, uint32_t s, uint32_t k) { return a->lists[(size_t)s * (a->I + 1) + k]; }

// Replays frame f of the FP32 trajectory from the state in d.pos / d.vel,
// writing the tape.  The sort of each substep is the forward's (the same
// keys of the same states): its permutation is taken from fwd_perm.
int record_frame(fsi_adjoint* a, const StepConst& c, uint32_t f) {
    DeviceState& d = a->core.d;
    Tape<float4>& t = a->t32;
    hipStream_t st = a->core.stream;
    const size_t n = a->n;
    const uint32_t I = a->I;
    FSI_HIP_TRY(zero_counters(a));
    for (uint32_t s = 0; s < a->S; ++s) {
        FSI_HIP_TRY(hipMemcpyAsync(t.xold + s * n, d.pos, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
        FSI_HIP_TRY(launch_predict(c, d, st));
        FSI_HIP_TRY(hipMemcpyAsync(d.perm, a->fwd_perm + ((size_t)f * a->S + s) * n, sizeof(uint32_t) * n,
                                   hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_sorted_keys, agrid(a->n), dim3(kAB), 0, st, a->n, d.keys, d.perm, d.skeys);
        FSI_HIP_TRY(launch_cells(c, d, st));
        FSI_HIP_TRY(launch_reorder(c, d, st));
        float4* Ps = t.P + (size_t)s * (I + 1) * n;
        FSI_HIP_TRY(hipMemcpyAsync(t.vs + s * n, d.vs, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
        FSI_HIP_TRY(hipMemcpyAsync(Ps, d.P, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
        for (uint32_t it = 0; it < I; ++it) {
            const EdgeList& L = list_of(a, s, it);
            FSI_HIP_TRY(build_list(a, c, L, d.P));
            const size_t q = ((size_t)s * I + it) * n;
            hipLaunchKernelGGL(k_rec_lambda, agrid(a->n), dim3(kAB), 0, st, c, d.P, L.nbr, L.cnt, d.PL, t.lr + q,
                               t.aux + q);
            hipLaunchKernelGGL(k_rec_deltap, agrid(a->n), dim3(kAB), 0, st, c, d.PL, L.nbr, L.cnt, d.P,
                               Ps + (it + 1) * n, t.dp + q);
        }
        const EdgeList& L = list_of(a, s, I);
        FSI_HIP_TRY(build_list(a, c, L, d.P));
        hipLaunchKernelGGL(k_rec_integrate, agrid(a->n), dim3(kAB), 0, st, c, d.P, d.vs, d.perm, d.pos, d.vel, L.nbr,
                           L.cnt, d.curl, d.posS, t.visc + s * n);
    }
    if (c.enable_vort) {
        const EdgeList& L = a->lists.back();
        FSI_HIP_TRY(build_list(a, c, L, d.posS));
        hipLaunchKernelGGL(k_rec_vorticity, agrid(a->n), dim3(kAB), 0, st, c, d.posS, d.curl, d.perm, d.vel, L.nbr,
                           L.cnt, t.eta);
    }
    FSI_HIP_TRY(hipGetLastError());
    return FSI_OK;
}

// Frame f of the FP64 trajectory from the state in t64.pos / t64.vel,
// writing the tape (the forward runs this too: the lists are the search).
// The forward sorts each substep and keeps the permutation in fwd_perm; the
// replay of the backward takes it from there.
int record_frame64(fsi_adjoint* a, const StepConst& c, const C64& k, uint32_t f, bool forward) {
    DeviceState& d = a->core.d;
    Tape<double4>& t = a->t64;
    hipStream_t st = a->core.stream;
    const size_t n = a->n;
    const uint32_t I = a->I;
    const int key_bits = log2u(a->cfg.hash_size);
    const dim3 g = agrid(a->n), b(kAB);
    FSI_HIP_TRY(zero_counters(a));
    for (uint32_t s = 0; s < a->S; ++s) {
        FSI_HIP_TRY(hipMemcpyAsync(t.xold + s * n, t.pos, sizeof(double4) * n, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_predict64, g, b, 0, st, k, t.pos, t.vel, t.xstar, d.keys);
        uint32_t* fp = a->fwd_perm + ((size_t)f * a->S + s) * n;
        if (forward) {
            FSI_HIP_TRY(
                radix_sort_pairs(d.keys, d.n, key_bits, a->core.sort_digit_bits, d.sort, &d.skeys, &d.perm, st));
            FSI_HIP_TRY(hipMemcpyAsync(fp, d.perm, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
        } else {
            FSI_HIP_TRY(hipMemcpyAsync(d.perm, fp, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
            hipLaunchKernelGGL(k_sorted_keys, g, b, 0, st, a->n, d.keys, d.perm, d.skeys);
        }
        FSI_HIP_TRY(launch_cells(c, d, st));
        hipLaunchKernelGGL(k_reorder64, g, b, 0, st, a->n, d.perm, t.xstar, t.vel, t.Pw, t.vsw);
        double4* Ps = t.P + (size_t)s * (I + 1) * n;
        FSI_HIP_TRY(hipMemcpyAsync(t.vs + s * n, t.vsw, sizeof(double4) * n, hipMemcpyDeviceToDevice, st));
        FSI_HIP_TRY(hipMemcpyAsync(Ps, t.Pw, sizeof(double4) * n, hipMemcpyDeviceToDevice, st));
        for (uint32_t it = 0; it < I; ++it) {
            const EdgeList& L = list_of(a, s, it);
            FSI_HIP_TRY(build_list(a, k, L, t.Pw));
            const size_t q = ((size_t)s * I + it) * n;
            hipLaunchKernelGGL(k_rec_lambda64, g, b, 0, st, k, t.Pw, L.nbr, L.cnt, t.PL, t.lr + q, t.aux + q);
            hipLaunchKernelGGL(k_rec_deltap64, g, b, 0, st, k, t.PL, L.nbr, L.cnt, t.Pw, Ps + (it + 1) * n, t.dp + q);
        }
        const EdgeList& L = list_of(a, s, I);
        FSI_HIP_TRY(build_list(a, k, L, t.Pw));
        hipLaunchKernelGGL(k_rec_integrate64, g, b, 0, st, k, t.Pw, t.vsw, d.perm, t.pos, t.vel, L.nbr, L.cnt, t.curl,
                           t.posS, t.visc + s * n);
    }
    if (k.enable_vort) {
        const EdgeList& L = a->lists.back();
        FSI_HIP_TRY(build_list(a, k, L, t.posS));
        hipLaunchKernelGGL(k_rec_vorticity64, g, b, 0, st, k, t.posS, t.curl, d.perm, t.vel, L.nbr, L.cnt, t.eta);
    }
    FSI_HIP_TRY(hipGetLastError());
    return FSI_OK;
}

// Runs the recorded frame backwards: gpos / gvel (id order) hold the adjoint
// of the frame's final state on entry and of its initial state on exit.
// perms: the sorted permutation of each substep (S x n); posS / curl: the
// last substep's positions and curl (sorted slots).
template <class T4>
int reverse_frame(fsi_adjoint* a, const C64& k, const Tape<T4>& t, const uint32_t* perms, const T4* posS,
                  const T4* curl) {
    hipStream_t st = a->core.stream;
    const uint32_t n = a->n;
    const size_t ns = n;
    const uint32_t I = a->I, S = a->S;
    const dim3 g = agrid(n), gq = qgrid(n), gr = rgrid(n), b(kAB);
    bool have_gw = false;
    if (k.enable_vort) {
        const EdgeList& L = a->lists.back();
        const uint32_t* perm = perms + (S - 1) * ns;
        hipLaunchKernelGGL((k_rev_vort<T4>), gq, b, 0, st, k, posS, curl, t.eta, perm, a->gvel, L.nbr, L.cnt, a->gown,
                           a->gw_own, a->ep, a->es);
        hipLaunchKernelGGL((k_rev_vort_t<T4>), gr, b, 0, st, n, curl, perm, L.tin, L.tcnt, a->ep, a->es, a->gown,
                           a->gw_own, a->gpos, a->gw);
        have_gw = true;
    }
    for (uint32_t s = S; s-- > 0;) {
        const uint32_t* perm = perms + s * ns;
        const T4* Ps = t.P + (size_t)s * (I + 1) * ns;
        const EdgeList& Li = list_of(a, s, I);
        hipLaunchKernelGGL((k_rev_integrate<T4>), gq, b, 0, st, k, t.xold + s * ns, t.vs + s * ns, Ps + I * ns,
                           t.visc + s * ns, perm, Li.nbr, Li.cnt, a->gpos, a->gvel,
                           (have_gw && s == S - 1) ? a->gw : nullptr, a->gown, a->gvs_own, a->gxold, a->acc_visc,
                           a->ep, a->ev);
        hipLaunchKernelGGL(k_rev_pairs_vec2, gr, b, 0, st, n, Li.tin, Li.tcnt, a->ep, a->ev, a->gown, a->gvs_own, a->gP,
                           a->gvs);
        for (uint32_t it = I; it-- > 0;) {
            const EdgeList& L = list_of(a, s, it);
            const size_t q = ((size_t)s * I + it) * ns;
            const T4* Pprev = Ps + it * ns;
            hipLaunchKernelGGL((k_rev_deltap<T4>), gq, b, 0, st, k, Pprev, t.lr + q, t.dp + q, L.nbr, L.cnt, a->gP,
                               a->gown, a->gl_own, a->ep, a->es);
            hipLaunchKernelGGL(k_rev_pairs_vec_scalar, gr, b, 0, st, n, L.tin, L.tcnt, a->ep, a->es, a->gown, a->gl_own,
                               a->gQ, a->gl);
            hipLaunchKernelGGL((k_rev_lambda<T4>), gq, b, 0, st, k, Pprev, t.lr + q, t.aux + q, perm, L.nbr, L.cnt,
                               a->gl, a->gQ, a->acc_eps, a->ep);
            hipLaunchKernelGGL(k_rev_pairs_sub, gr, b, 0, st, n, L.tin, L.tcnt, a->ep, a->gQ);
            std::swap(a->gP, a->gQ);
        }
        hipLaunchKernelGGL(k_rev_predict, g, b, 0, st, k, perm, a->gP, a->gvs, a->gxold, a->gpos, a->gvel);
    }
    FSI_HIP_TRY(hipGetLastError());
    return FSI_OK;
}

// Sets the constants for the given parameters; returns the step constants.
StepConst prepare(fsi_adjoint* a) {
    pack_constants(a->cfg, a->cfg.bounds_min, a->cfg.bounds_max, std::clamp(a->cfg.bound_damping, 0.0f, 1.0f), a->n,
                   a->k);
    return step_const(a);
}

// Forward of either precision.  f64_io: states are n double4 (FP64
// trajectory only), else n float4.
int forward_impl(fsi_adjoint* a, const void* pos0, const void* vel0, bool f64_io, double viscosity, double eps,
                 void* pos_out, void* vel_out, void* stream) {
    int rc = bind(a);
    if (rc) return rc;
    if (!pos0 || !vel0) return fail(FSI_ERR_INVALID_ARG, "NULL state");
    if (!std::isfinite(viscosity) || !std::isfinite(eps)) return fail(FSI_ERR_INVALID_ARG, "non-finite parameter");
    if (f64_io && a->precision != 64) return fail(FSI_ERR_STATE, "FP64 states need the FP64 trajectory");
    if (!(a->precision == 32 ? a->t32.ready : a->t64.ready)) return fail(FSI_ERR_STATE, "handle has no tape");
    a->forward_done = false;
    a->cfg.viscosity = (float)viscosity;
    a->cfg.eps = (float)eps;
    a->visc64 = f64_io ? viscosity : (double)a->cfg.viscosity;
    a->eps64 = f64_io ? eps : (double)a->cfg.eps;
    const StepConst c = prepare(a);
    Core& core = a->core;
    DeviceState& d = core.d;
    hipStream_t st = core.stream;
    const size_t n = a->n;
    FSI_HIP_TRY(join_in(a, stream));
    if (a->precision == 64) {
        const C64 k = c64_of(a, c);
        Tape<double4>& t = a->t64;
        const dim3 g = agrid(a->n), b(kAB);
        if (f64_io) {
            FSI_HIP_TRY(hipMemcpyAsync(t.pos, pos0, sizeof(double4) * n, hipMemcpyDeviceToDevice, st));
            FSI_HIP_TRY(hipMemcpyAsync(t.vel, vel0, sizeof(double4) * n, hipMemcpyDeviceToDevice, st));
        } else {
            hipLaunchKernelGGL(k_f2d, g, b, 0, st, a->n, (const float4*)pos0, t.pos, 1);
            hipLaunchKernelGGL(k_f2d, g, b, 0, st, a->n, (const float4*)vel0, t.vel, 1);
        }
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(kAB), 0, st, a->flags, (size_t)kFlagWords);
        for (uint32_t f = 0; f < a->frames; ++f) {
            FSI_HIP_TRY(hipMemcpyAsync(t.ck_pos + f * n, t.pos, sizeof(double4) * n, hipMemcpyDeviceToDevice, st));
            FSI_HIP_TRY(hipMemcpyAsync(t.ck_vel + f * n, t.vel, sizeof(double4) * n, hipMemcpyDeviceToDevice, st));
            if ((rc = record_frame64(a, c, k, f, true))) return rc;
        }
        FSI_HIP_TRY(hipMemcpyAsync(t.ck_pos + a->frames * n, t.pos, sizeof(double4) * n, hipMemcpyDeviceToDevice, st));
        FSI_HIP_TRY(hipMemcpyAsync(t.ck_vel + a->frames * n, t.vel, sizeof(double4) * n, hipMemcpyDeviceToDevice, st));
        if (f64_io) {
            if (pos_out) FSI_HIP_TRY(hipMemcpyAsync(pos_out, t.pos, sizeof(double4) * n, hipMemcpyDeviceToDevice, st));
            if (vel_out) FSI_HIP_TRY(hipMemcpyAsync(vel_out, t.vel, sizeof(double4) * n, hipMemcpyDeviceToDevice, st));
        } else {
            if (pos_out) hipLaunchKernelGGL(k_d2f, g, b, 0, st, a->n, t.pos, (float4*)pos_out, 1);
            if (vel_out) hipLaunchKernelGGL(k_d2f, g, b, 0, st, a->n, t.vel, (float4*)vel_out, 1);
        }
        FSI_HIP_TRY(hipGetLastError());
        uint32_t flags[4];
        a->host_words.resize(kFlagWords);
        FSI_HIP_TRY(hipMemcpyAsync(a->host_words.data(), a->flags, sizeof(uint32_t) * kFlagWords,
                                   hipMemcpyDeviceToHost, st));
        FSI_HIP_TRY(join_out(a, stream));
        FSI_HIP_TRY(hipStreamSynchronize(st));
        reduce_flags(a->host_words.data(), flags);
        if (flags[0])
            return fail(FSI_ERR_CAPACITY, "more than 96 neighbours (" + std::to_string(flags[2]) +
                                              ") or 128 pairs into one particle (" + std::to_string(flags[3]) +
                                              ") in a pass");
        a->forward_done = true;
        return FSI_OK;
    }
    Tape<float4>& t = a->t32;
    d.tiled = c.dom.dshift == kTileShift;
    FramePlan plan;
    plan.substeps = a->S;
    plan.max_iter = a->I;
    plan.key_bits = log2u(a->cfg.hash_size);
    plan.vorticity = a->cfg.enable_vorticity != 0;
    FSI_HIP_TRY(hipMemcpyAsync(d.pos, pos0, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
    FSI_HIP_TRY(hipMemcpyAsync(d.vel, vel0, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
    for (uint32_t f = 0; f < a->frames; ++f) {
        FSI_HIP_TRY(hipMemcpyAsync(t.ck_pos + f * n, d.pos, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
        FSI_HIP_TRY(hipMemcpyAsync(t.ck_vel + f * n, d.vel, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
        core.record_perm = a->fwd_perm + (size_t)f * a->S * n;
        rc = core_frame(core, c, plan);
        core.record_perm = nullptr;
        if (rc) return rc;
    }
    FSI_HIP_TRY(hipMemcpyAsync(t.ck_pos + a->frames * n, d.pos, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
    FSI_HIP_TRY(hipMemcpyAsync(t.ck_vel + a->frames * n, d.vel, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
    if (pos_out) FSI_HIP_TRY(hipMemcpyAsync(pos_out, d.pos, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
    if (vel_out) FSI_HIP_TRY(hipMemcpyAsync(vel_out, d.vel, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
    FSI_HIP_TRY(join_out(a, stream));
    a->forward_done = true;
    return FSI_OK;
}

// Backward of either precision; f64_io as in forward_impl.
int backward_impl(fsi_adjoint* a, const void* grad_pos, const void* grad_vel, bool f64_io, void* grad_pos0,
                  void* grad_vel0, double* grad_params, void* stream) {
    int rc = bind(a);
    if (rc) return rc;
    if (!a->forward_done) return fail(FSI_ERR_STATE, "backward before forward");
    if (f64_io && a->precision != 64) return fail(FSI_ERR_STATE, "FP64 states need the FP64 trajectory");
    const StepConst c = step_const(a);
    const C64 k = c64_of(a, c);
    Core& core = a->core;
    DeviceState& d = core.d;
    hipStream_t st = core.stream;
    const size_t n = a->n;
    const dim3 g = agrid(a->n), b(kAB);
    FSI_HIP_TRY(join_in(a, stream));
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(kAB), 0, st, a->flags, (size_t)kFlagWords);
    hipLaunchKernelGGL(k_zero_u32, agrid(2 * kSpread * a->frames), dim3(kAB), 0, st, a->norms,
                       2 * kSpread * (size_t)a->frames);
    hipLaunchKernelGGL(k_zero_f64, g, b, 0, st, a->acc_visc, n);
    hipLaunchKernelGGL(k_zero_f64, g, b, 0, st, a->acc_eps, n);
    for (int q = 0; q < 2; ++q) {
        const void* src = q == 0 ? grad_pos : grad_vel;
        double4* dst = q == 0 ? a->gpos : a->gvel;
        if (!src)
            hipLaunchKernelGGL(k_zero_f64, agrid(4 * a->n), b, 0, st, (double*)dst, 4 * n);
        else if (f64_io)
            FSI_HIP_TRY(hipMemcpyAsync(dst, src, sizeof(double4) * n, hipMemcpyDeviceToDevice, st));
        else
            hipLaunchKernelGGL(k_f2d, g, b, 0, st, a->n, (const float4*)src, dst, 0);
    }
    for (uint32_t f = a->frames; f-- > 0;) {
        if (a->precision == 64) {
            Tape<double4>& t = a->t64;
            FSI_HIP_TRY(hipMemcpyAsync(t.pos, t.ck_pos + f * n, sizeof(double4) * n, hipMemcpyDeviceToDevice, st));
            FSI_HIP_TRY(hipMemcpyAsync(t.vel, t.ck_vel + f * n, sizeof(double4) * n, hipMemcpyDeviceToDevice, st));
            if ((rc = record_frame64(a, c, k, f, false))) return rc;
            hipLaunchKernelGGL((k_compare<double4>), g, b, 0, st, a->n, t.pos, t.ck_pos + (f + 1) * n, a->flags + 1);
            hipLaunchKernelGGL((k_compare<double4>), g, b, 0, st, a->n, t.vel, t.ck_vel + (f + 1) * n, a->flags + 1);
            if ((rc = reverse_frame(a, k, t, a->fwd_perm + (size_t)f * a->S * n, (const double4*)t.posS,
                                    (const double4*)t.curl)))
                return rc;
        } else {
            Tape<float4>& t = a->t32;
            FSI_HIP_TRY(hipMemcpyAsync(d.pos, t.ck_pos + f * n, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
            FSI_HIP_TRY(hipMemcpyAsync(d.vel, t.ck_vel + f * n, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
            if ((rc = record_frame(a, c, f))) return rc;
            hipLaunchKernelGGL((k_compare<float4>), g, b, 0, st, a->n, d.pos, t.ck_pos + (f + 1) * n, a->flags + 1);
            hipLaunchKernelGGL((k_compare<float4>), g, b, 0, st, a->n, d.vel, t.ck_vel + (f + 1) * n, a->flags + 1);
            if ((rc = reverse_frame(a, k, t, a->fwd_perm + (size_t)f * a->S * n, (const float4*)d.posS,
                                    (const float4*)d.curl)))
                return rc;
        }
        hipLaunchKernelGGL(k_absmax, g, b, 0, st, a->n, a->gpos, a->gvel, a->norms + 2 * kSpread * (size_t)f);
    }
    hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kAB), 0, st, a->n, a->acc_visc, a->acc_eps, a->dred);
    for (int q = 0; q < 2; ++q) {
        void* dst = q == 0 ? grad_pos0 : grad_vel0;
        const double4* src = q == 0 ? a->gpos : a->gvel;
        if (!dst) continue;
        if (f64_io)
            FSI_HIP_TRY(hipMemcpyAsync(dst, src, sizeof(double4) * n, hipMemcpyDeviceToDevice, st));
        else
            hipLaunchKernelGGL(k_d2f, g, b, 0, st, a->n, src, (float4*)dst, 0);
    }
    FSI_HIP_TRY(hipGetLastError());
    double red[2] = {0.0, 0.0};
    uint32_t* flags = a->stats;
    FSI_HIP_TRY(hipMemcpyAsync(red, a->dred, sizeof(red), hipMemcpyDeviceToHost, st));
    const size_t nnorm = 2 * kSpread * (size_t)a->frames;
    a->host_words.resize(kFlagWords + nnorm);
    FSI_HIP_TRY(hipMemcpyAsync(a->host_words.data(), a->flags, sizeof(uint32_t) * kFlagWords, hipMemcpyDeviceToHost,
                               st));
    FSI_HIP_TRY(hipMemcpyAsync(a->host_words.data() + kFlagWords, a->norms, sizeof(uint32_t) * nnorm,
                               hipMemcpyDeviceToHost, st));
    FSI_HIP_TRY(join_out(a, stream));
    FSI_HIP_TRY(hipStreamSynchronize(st));
    reduce_flags(a->host_words.data(), flags);
    a->host_norms.assign(2 * (size_t)a->frames, 0u);
    for (size_t q = 0; q < a->host_norms.size(); ++q)
        for (uint32_t w = 0; w < kSpread; ++w)
            a->host_norms[q] = std::max(a->host_norms[q], a->host_words[kFlagWords + q * kSpread + w]);
    if (flags[1])
        return fail(FSI_ERR_STATE, "replay differs from the forward (" + std::to_string(flags[1]) + " records)");
    if (flags[0])
        return fail(FSI_ERR_CAPACITY, "more than 96 neighbours (" + std::to_string(flags[2]) +
                                          ") or 128 pairs into one particle (" + std::to_string(flags[3]) +
                                          ") in a pass");
    if (grad_params) {
        grad_params[0] = red[0];
        grad_params[1] = red[1];
    }
    return FSI_OK;
}

}  // namespace

extern "C" {

int fsi_adj_create(const fsi_fluid_config* cfg, uint32_t n, uint32_t frames, fsi_adjoint** out) {
    if (!out) return fail(FSI_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    int rc = check_config(cfg);
    if (rc) return rc;
    if (cfg->enable_apbf) return fail(FSI_ERR_INVALID_ARG, "the adjoint does not differentiate APBF levels of detail");
    // pair numbers r * n + i fit 32 bits
    if (n == 0 || n > (1u << 25)) return fail(FSI_ERR_INVALID_ARG, "particle count must be 1..2^25");
    if (frames == 0 || frames > 100000u) return fail(FSI_ERR_INVALID_ARG, "frames must be 1..100000");
    if (cfg->solver_iterations == 0 || cfg->solver_iterations > 64u)
        return fail(FSI_ERR_INVALID_ARG, "solver_iterations must be 1..64 for the adjoint");
    if (cfg->substep_iterations > 64u) return fail(FSI_ERR_INVALID_ARG, "substep_iterations must be 1..64");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(FSI_ERR_NO_DEVICE, "no HIP device visible");
    fsi_adjoint* a = new (std::nothrow) fsi_adjoint();
    if (!a) return fail(FSI_ERR_INVALID_ARG, "out of host memory");
    auto bail = [&](int code) {
        free_all(a);
        core_destroy(a->core);
        delete a;
        return code;
    };
    if ((rc = core_init(a->core))) return bail(rc);
    a->cfg = *cfg;
    a->cfg.particle_count = n;
    a->n = n;
    a->frames = frames;
    a->S = cfg->substep_iterations;
    a->I = cfg->solver_iterations;
    hipError_t e = core_alloc_particles(a->core, n);
    if (e == hipSuccess) e = core_alloc_cells(a->core, cfg->hash_size);
    if (e == hipSuccess) e = alloc_common(a);
    if (e == hipSuccess) e = a->t32.alloc(n, frames, a->S, a->I, false);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&a->ev_in, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&a->ev_out, hipEventDisableTiming);
    if (e != hipSuccess) return bail(fail(FSI_ERR_HIP, std::string("allocation: ") + hipGetErrorString(e)));
    a->core.d.batched = false;
    a->core.d.any_sdf = false;
    *out = a;
    return FSI_OK;
}

int fsi_adj_destroy(fsi_adjoint* a) {
    if (!a) return FSI_OK;
    (void)hipSetDevice(a->core.device);
    if (a->core.stream) (void)hipStreamSynchronize(a->core.stream);
    if (a->ev_in) (void)hipEventDestroy(a->ev_in);
    if (a->ev_out) (void)hipEventDestroy(a->ev_out);
    free_all(a);
    core_destroy(a->core);
    delete a;
    return FSI_OK;
}

int fsi_adj_set_precision(fsi_adjoint* a, uint32_t bits) {
    int rc = bind(a);
    if (rc) return rc;
    if (bits != 32 && bits != 64) return fail(FSI_ERR_INVALID_ARG, "precision must be 32 or 64");
    if (bits == a->precision) return FSI_OK;
    FSI_HIP_TRY(hipStreamSynchronize(a->core.stream));
    Tape<float4>& t32 = a->t32;
    Tape<double4>& t64 = a->t64;
    hipError_t e = hipSuccess;
    if (bits == 64) {
        t32.release();
        e = t64.alloc(a->n, a->frames, a->S, a->I, true);
    } else {
        t64.release();
        e = t32.alloc(a->n, a->frames, a->S, a->I, false);
    }
    if (e != hipSuccess) {
        t32.release();
        t64.release();
        a->forward_done = false;
        return fail(FSI_ERR_HIP, std::string("allocation: ") + hipGetErrorString(e) + " (handle has no tape now)");
    }
    a->precision = bits;
    a->forward_done = false;
    return FSI_OK;
}

int fsi_adj_forward(fsi_adjoint* a, const float* pos0, const float* vel0, float viscosity, float eps, float* pos_out,
                    float* vel_out, void* stream) {
    return forward_impl(a, pos0, vel0, false, viscosity, eps, pos_out, vel_out, stream);
}

int fsi_adj_forward64(fsi_adjoint* a, const double* pos0, const double* vel0, double viscosity, double eps,
                      double* pos_out, double* vel_out, void* stream) {
    return forward_impl(a, pos0, vel0, true, viscosity, eps, pos_out, vel_out, stream);
}

int fsi_adj_backward(fsi_adjoint* a, const float* grad_pos, const float* grad_vel, float* grad_pos0,
                     float* grad_vel0, double* grad_params, void* stream) {
    return backward_impl(a, grad_pos, grad_vel, false, grad_pos0, grad_vel0, grad_params, stream);
}

int fsi_adj_backward64(fsi_adjoint* a, const double* grad_pos, const double* grad_vel, double* grad_pos0,
                       double* grad_vel0, double* grad_params, void* stream) {
    return backward_impl(a, grad_pos, grad_vel, true, grad_pos0, grad_vel0, grad_params, stream);
}

int fsi_adj_frame_norms(fsi_adjoint* a, float* out) {
    if (!a || !out) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    if (a->host_norms.size() != 2 * (size_t)a->frames) return fail(FSI_ERR_STATE, "no backward has run");
    for (size_t i = 0; i < a->host_norms.size(); ++i) {
        const uint32_t b = a->host_norms[i];
        float f;
        std::memcpy(&f, &b, sizeof(f));
        out[i] = f;
    }
    return FSI_OK;
}

int fsi_adj_stats(fsi_adjoint* a, uint32_t out[4]) {
    if (!a || !out) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    std::memcpy(out, a->stats, sizeof(a->stats));
    return FSI_OK;
}

}  // extern "C"
