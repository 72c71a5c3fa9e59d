// pbf_kernels.hip — the PBF substep kernels for gfx950.
//
// Storage: the public state (pos, vel) stays in particle-id order, as the
// reference's Particle SSBO.  After the sort, the predicted positions and
// velocities are gathered into key-sorted arrays, so the members of one hash
// bucket are contiguous (one bucket of ~8 particles is one or two 128-byte
// lines) and every gather pass reads them without the reference's
// hashGrid[k].y -> solver[j] indirection.  A particle's neighbour sums still
// run bucket by bucket in the reference order (z, y, x offsets; slots in
// sorted order), so each sum is the same sequence of FP32 operations.
//
// Pairs with r^2 > h^2 contribute exact zeros in the reference (poly6 and the
// spiky gradient vanish there, CS_ComputeLambdas.glsl:76,86) and are skipped;
// an accumulator that starts at +0 never becomes -0, so skipping them leaves
// every sum bit-identical for finite inputs.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbf_gather.h"
#include "pbf_launch.h"

namespace fsi {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint32_t gid() { return blockIdx.x * kBlock + threadIdx.x; }

// Workgroups are handed to the 8 XCDs of the MI355X round-robin (block b runs
// on XCD b % 8).  Renumber them so that each XCD walks one contiguous range
// of the processing order, i.e. one compact region of space whose buckets and
// cell-table lines stay in that XCD's own L2.
__device__ __forceinline__ uint32_t xcd_gid() {
    const uint32_t b = blockIdx.x, nb = gridDim.x;
    const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u, l = b >> 3;
    const uint32_t lb = (x < r) ? x * (q + 1) + l : r * (q + 1) + (x - r) * q + l;
    return lb * kBlock + threadIdx.x;
}

inline dim3 grid_for(uint32_t n) { return dim3((n + kBlock - 1) / kBlock); }

// The same for the gather kernels, whose workgroup size BS (128 or 256) is a
// tuning knob (DeviceState::block).
template <int BS>
__device__ __forceinline__ uint32_t xcd_gid_bs() {
    const uint32_t b = blockIdx.x, nb = gridDim.x;
    const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u, l = b >> 3;
    const uint32_t lb = (x < r) ? x * (q + 1) + l : r * (q + 1) + (x - r) * q + l;
    return lb * BS + threadIdx.x;
}
template <int BS>
inline dim3 grid_bs(uint32_t n) { return dim3((n + BS - 1) / BS); }

// The storage slot thread t takes: order[t], or t itself when the storage is
// in processing order (order == nullptr).
__device__ __forceinline__ uint32_t slot_at(const uint32_t* __restrict__ order, uint32_t t) {
    return order != nullptr ? order[t] : t;
}

// Per-lane LDS list of the two-phase form (entries), and the fill level at
// which the whole wave flushes.
constexpr uint32_t kLdsCap = 24;
constexpr uint32_t kFlushAt = 16;

// The neighbour list: kNbrCap slots per thread.  The 64 threads of a wave
// own one block of 64 x kNbrCap slots, in which entries r..r+3 (r a multiple
// of four) of all 64 threads form one contiguous 1 KiB run, four slots per
// thread: entry r of thread t is slot
//   (t >> 6) * 64 * kNbrCap + (r >> 2) * 256 + (t & 63) * 4 + (r & 3).
// The walks read four entries per thread with one 16-byte load and the list
// builder writes them with one 16-byte store, so each wave access covers
// whole cache lines once.  (One row per thread instead made every wave load
// touch 64 lines, re-fetched four times each as the row was walked, and the
// lines of the waves in flight did not fit in the L2.)
// Thread t's entries r..r+3 (r a multiple of four) as one uint4: wave_run(...)[(r >> 2) * 64].
// The wave's block base is wave-uniform (read from the first active lane).
__device__ __forceinline__ uint4* wave_run(uint32_t* list, uint32_t t) {
    return reinterpret_cast<uint4*>(list + (size_t)__builtin_amdgcn_readfirstlane(t >> 6) *
                                               (64u * (uint32_t)kNbrCap)) + (t & 63u);
}
__device__ __forceinline__ const uint4* wave_run(const uint32_t* list, uint32_t t) {
    return reinterpret_cast<const uint4*>(list + (size_t)__builtin_amdgcn_readfirstlane(t >> 6) *
                                                     (64u * (uint32_t)kNbrCap)) + (t & 63u);
}

// Per-lane staging ring of the list builder (entries; a power of two and a
// multiple of four).
constexpr uint32_t kLRing = 16;  // holds three groups and a partly filled one
// The list builder's bucket table: per lane, the 27 buckets of the walk as
// start << L | length, L = the leading zero bits of n (at least 5: a start
// is below n < 2^27), and an empty entry after them.  A bucket longer than
// 2^L - 1 slots (never at 1M particles: L = 11) sends the particle to the
// bucket scan of the walks.
constexpr uint32_t kBucketEntries = 28;

// The neighbour-list builder of the gathers (the first of their two
// kernels).  Thread t takes the particle in slot order[t], whose position is
// X: it scans the 27 buckets around it in scan_buckets' order and gives every
// candidate the quantised pre-test (qdist2 <= qpass, see pbf_device.h); the
// slots that pass are its list, in visit order (the particle's own slot
// passes, r = 0, and so may a candidate just outside the support: the walks
// give every entry the exact test).
//
// First the lane reads the cell-table entries of its 27 buckets into its
// column of an LDS table.  Then it streams through them on its own: every
// iteration tests the next four slots of its current bucket and moves to
// the next bucket when this one is done, so the lanes of a wave run about as
// many iterations as their longest bucket sequence needs.  (A walk in step
// over the 27 positions runs, at each position, as many groups as the
// longest bucket of the wave's lanes there: the lanes sit in several cells
// whose buckets differ in length, and at 1M particles that is half as many
// groups again.)  An iteration loads three groups of four of the lane's
// stream before it tests any, so a wave waits once per twelve slots.
//
// Slots are staged in a per-lane LDS ring and written to HBM four at a time
// (the layout above), so every store covers whole lines.  list_count[t] is
// the number of entries, also when it exceeds kNbrCap (then only the first
// kNbrCap are kept and the walks scan the buckets instead; a bucket longer
// than the table holds gives kNbrCap + 1 without a list).  With lod, an
// APBF-inactive particle builds no list.  With TAPS, cand_out[k] counts the
// candidates other than the particle itself, as the reference visits them.
// X and Q hold n slots (n < 2^27).
template <bool TAPS, bool B, int BS>
__global__ __launch_bounds__(BS) void k_build_list(StepConst c, const uint32_t* __restrict__ order,
                                                       const float4* __restrict__ X, const uint2* __restrict__ Q,
                                                       const uint32_t* __restrict__ lod, uint32_t iter,
                                                       const int2* __restrict__ cells,
                                                       uint32_t* __restrict__ list_count, uint32_t* __restrict__ list,
                                                       uint32_t* __restrict__ cand_out) {
    const uint32_t t = xcd_gid_bs<BS>();
    if (t >= c.n) return;
    const uint32_t k = slot_at(order, t);
    if (lod != nullptr && c.enable_apbf != 0u && lod[k] < iter) return;
    const V3 x = xyz(X[k]);
    const uint2 qi = quantize(x, c.qscale);
    const int32_t qpass = qpass_of(x, c.qscale);
    // the staging rings first (LDS offset 0, so that a ring address is an
    // OR of the entry and lane bits), then the bucket tables
    __shared__ uint32_t lds[(kLRing + kBucketEntries) * BS];
    uint32_t* mine = &lds[kLRing * BS + threadIdx.x];
    uint4* run = wave_run(list, t);
    int32_t cx, cy, cz;
    cell_of(x, c.h, cx, cy, cz);
    const BucketWalk bw(cx, cy, cz, domain_of<B>(c, k));
    // the bucket table (an empty bucket is (-1, -1) in the cell table)
    const uint32_t lbits = (uint32_t)__builtin_clz(c.n), lmax = (1u << lbits) - 1u;  // L and 2^L - 1
    uint32_t ncand = 0;
    bool longb = false;
#pragma unroll 1
    for (int q0 = 0; q0 < 27; q0 += 9) {
        int2 se[9];
#pragma unroll
        for (int j = 0; j < 9; ++j) se[j] = cells[bw.bucket(q0 + j)];
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            const uint32_t len = (uint32_t)(se[j].y - se[j].x);
            if (TAPS) ncand += len - ((k - (uint32_t)se[j].x) < len ? 1u : 0u);
            longb = longb || len > lmax;
            mine[(q0 + j) * BS] = len != 0u ? ((uint32_t)se[j].x << lbits) | (len & lmax) : 0u;
        }
    }
    mine[27 * BS] = 0u;
    // the records of slots s..s+3 (two 16-byte loads of Q, which holds n + 4
    // records: every group of a bucket lies inside it; records past the
    // bucket are not valid candidates)
    struct G4 {
        uint4 a, b;
    };
    typedef uint4 uint4_a8 __attribute__((aligned(8)));
    const char* qbytes = reinterpret_cast<const char*>(Q);
    auto loadg = [&](uint32_t s) {
        const char* p = qbytes + (size_t)s * sizeof(uint2);
        return G4{*reinterpret_cast<const uint4_a8*>(p), *reinterpret_cast<const uint4_a8*>(p + 16)};
    };
    // ring positions in bytes: entry e of this lane's ring is at
    // e * step + lane bytes, so an address is one AND-OR of the position
    constexpr uint32_t step = (uint32_t)BS * 4u, wrap = (kLRing - 1u) * step;
    const uint32_t laneb = threadIdx.x * 4u;
    char* const lbase = reinterpret_cast<char*>(lds);
    uint32_t tail = 0, done = 0;  // entries appended / written out (a multiple of four), times step
    // writes entries done..done+3 (the last group may hold stale slots past
    // the count: they are never read as entries)
    auto copy4 = [&]() {
        const uint32_t* r = reinterpret_cast<const uint32_t*>(lbase + ((done & wrap) | laneb));
        const uint4 v = make_uint4(r[0], r[BS], r[2 * BS], r[3 * BS]);
        if (done < (uint32_t)kNbrCap * step) run[(done / (4u * step)) * 64u] = v;
        done += 4u * step;
    };
    // branch-free append: the slot is always written and the tail advances
    // only for a valid slot that passes the pre-test
    auto visit = [&](uint32_t s, uint2 qs, bool valid) {
        *reinterpret_cast<uint32_t*>(lbase + ((tail & wrap) | laneb)) = s;
        tail += (qdist2(qi, qs) <= qpass) && valid ? step : 0u;
    };
    // the stream: a group is (start << L | slots left), the table's own
    // encoding (0 when nothing is left); p is the table entry it came from
    // and p = 27 the empty entry after the last bucket.  An iteration tests
    // three groups of four of the stream, A (g, p), B and C after it (each
    // the next four slots of the previous group's bucket or the first four
    // of the next bucket).  All three are loaded at the top of the
    // iteration, and the table entries the next iteration needs are read
    // while they are tested.
    auto entry = [&](uint32_t q) { return mine[min(q, 27u) * BS]; };
    // four slots on in the same bucket: start + 4, left - 4 (left > 4, so
    // neither field carries into the other)
    const uint32_t adv = (4u << lbits) - 4u;
    // the group after (g0, p0), given e = the entry after p0
    auto after = [&](uint32_t g0, uint32_t p0, uint32_t e, uint32_t& g1, uint32_t& p1) {
        const bool same = (g0 & lmax) > 4u;
        g1 = same ? g0 + adv : e;
        p1 = same ? p0 : min(p0 + 1u, 27u);
    };
    uint32_t p = longb ? 27u : 0u;
    uint32_t g = entry(p), gB, pB, gC, pC;
    after(g, p, entry(p + 1u), gB, pB);
    after(gB, pB, entry(pB + 1u), gC, pC);
    // a group with no slots left is 0 and loads slots 0..3 (Q holds n + 4
    // records); a group with slots loads inside Q as well
    auto test4 = [&](const G4& q4, uint32_t g0) {
        const uint32_t s0 = g0 >> lbits, r0 = g0 & lmax;
        visit(s0, make_uint2(q4.a.x, q4.a.y), 0u < r0);
        visit(s0 + 1, make_uint2(q4.a.z, q4.a.w), 1u < r0);
        visit(s0 + 2, make_uint2(q4.b.x, q4.b.y), 2u < r0);
        visit(s0 + 3, make_uint2(q4.b.z, q4.b.w), 3u < r0);
    };
#pragma unroll 1
    while (__any(p < 27u || g != 0u)) {
        const G4 ga = loadg(g >> lbits), gb = loadg(gB >> lbits), gc = loadg(gC >> lbits);
        // the entries after pC (the next three groups start at or after it)
        const uint32_t f1 = entry(pC + 1u), f2 = entry(pC + 2u), f3 = entry(pC + 3u);
        if (tail - done > (kLRing - 12u) * step) copy4();  // room for twelve more
        if (tail - done > (kLRing - 12u) * step) copy4();
        if (tail - done > (kLRing - 12u) * step) copy4();
        test4(ga, g);
        test4(gb, gB);
        test4(gc, gC);
        // the next A follows C, the next B follows it, the next C follows
        // that; each takes the entry after its predecessor's (pC + 1, 2, 3)
        after(gC, pC, f1, g, p);
        after(g, p, p == pC ? f1 : f2, gB, pB);
        after(gB, pB, pB == pC ? f1 : (pB == pC + 1u ? f2 : f3), gC, pC);
    }
#pragma unroll 1
    while (tail > done) copy4();
    tail /= step;
    list_count[t] = longb ? (uint32_t)kNbrCap + 1u : tail;
    if (TAPS) cand_out[k] = ncand;
}

// The walk of a gather (the second kernel): hands thread t's neighbours to
// body(j, xj, r, r2, vj) in visit order, with xj = X[j], r = xi - xj and, with
// VEL, vj = V[j] (loaded with xj, ahead of the test).  Every list entry gets
// the exact test (j != self, r^2 <= h^2; the reference adds exact zeros for
// the pairs it rejects).  A list that overflowed kNbrCap is not walked: the
// 27 buckets are scanned again.  Four entries per 16-byte load, and their
// records through a buffer resource over X (n slots): entries past the count
// are never used, and a stale one past n reads zeros.  Without VEL an
// iteration takes eight entries (lambda 209.8 -> 205.8 us, delta-p 63.9 ->
// 60.0 us at 1M particles); with VEL, whose sixteen records per iteration
// would take 32 more registers, it takes four.
template <bool TAPS, bool VEL, class Body>
__device__ __forceinline__ void gather_list(V3 xi, float h, const Domain& dm, const int2* __restrict__ cells,
                                            const float4* __restrict__ X, const float4* __restrict__ V, uint32_t n,
                                            uint32_t self, uint32_t t, const uint32_t* __restrict__ list_count,
                                            const uint32_t* __restrict__ list, Body&& body, uint32_t& nnbr) {
    const float h2 = h * h;
    const uint32_t cnt = list_count[t];
    if (cnt <= (uint32_t)kNbrCap) {
        const uint4* run = wave_run(list, t);
        const __amdgpu_buffer_rsrc_t xr =
            __builtin_amdgcn_make_buffer_rsrc((void*)X, (short)0, (int)(n * (uint32_t)sizeof(float4)), 0x00020000);
        const __amdgpu_buffer_rsrc_t vr =
            __builtin_amdgcn_make_buffer_rsrc((void*)V, (short)0, (int)(n * (uint32_t)sizeof(float4)), 0x00020000);
        auto load = [&](const __amdgpu_buffer_rsrc_t& rs, uint32_t j) {
            const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(j * (uint32_t)sizeof(float4)), 0, 0);
            return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                               __uint_as_float(v[3]));
        };
        // with VEL the velocity loads stay behind a test of V: the compiler
        // then issues the four position loads before the four velocity loads
        // (k_integrate 139 us per launch at 1M particles; 159 us with the
        // loads interleaved as it schedules them when the test is constant)
        const bool withv = VEL && V != nullptr;
        auto exact = [&](uint32_t j, float4 xj, float4 vj) {
            const V3 r = sub(xi, xyz(xj));
            const float r2 = dot(r, r);
            if (j != self && r2 <= h2) {
                if (TAPS) ++nnbr;
                body(j, xj, r, r2, vj);
            }
        };
        if constexpr (!VEL) {
            // eight entries per iteration: two list loads, then eight record
            // loads in flight together (the second four re-read the last
            // group when the list ends inside the first)
            const uint32_t last4 = (cnt - 1u) & ~3u;
#pragma unroll 1
            for (uint32_t e = 0; e < cnt; e += 8) {
                const uint4 j = run[(e >> 2) * 64u];
                const uint4 jb = run[(min(e + 4u, last4) >> 2) * 64u];
                const float4 a0 = load(xr, j.x), a1 = load(xr, j.y), a2 = load(xr, j.z), a3 = load(xr, j.w);
                const float4 b0 = load(xr, jb.x), b1 = load(xr, jb.y), b2 = load(xr, jb.z), b3 = load(xr, jb.w);
                const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                exact(j.x, a0, z);
                if (e + 1 < cnt) exact(j.y, a1, z);
                if (e + 2 < cnt) exact(j.z, a2, z);
                if (e + 3 < cnt) exact(j.w, a3, z);
                if (e + 4 < cnt) exact(jb.x, b0, z);
                if (e + 5 < cnt) exact(jb.y, b1, z);
                if (e + 6 < cnt) exact(jb.z, b2, z);
                if (e + 7 < cnt) exact(jb.w, b3, z);
            }
            return;
        }
#pragma unroll 1
        for (uint32_t e = 0; e < cnt; e += 4) {
            const uint4 j = run[(e >> 2) * 64u];
            const float4 a0 = load(xr, j.x), a1 = load(xr, j.y), a2 = load(xr, j.z), a3 = load(xr, j.w);
            float4 v0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), v1 = v0, v2 = v0, v3 = v0;
            if (withv) {
                v0 = load(vr, j.x);
                v1 = load(vr, j.y);
                v2 = load(vr, j.z);
                v3 = load(vr, j.w);
            }
            exact(j.x, a0, v0);
            if (e + 1 < cnt) exact(j.y, a1, v1);
            if (e + 2 < cnt) exact(j.z, a2, v2);
            if (e + 3 < cnt) exact(j.w, a3, v3);
        }
    } else {
        scan_buckets(xi, h, dm, cells, X, self, [&](uint32_t j, float4 xj, V3 r, float r2) {
            if (r2 <= h2) {
                if (TAPS) ++nnbr;
                body(j, xj, r, r2, V != nullptr ? V[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f));
            }
        });
    }
}

// Neighbour gather, bit-identical to the reference's loop in both forms:
//   V = 0  single loop: body(j, xj, r, r2) for each candidate in the support.
//   V = 1  list walk (gather_list) over the list k_build_list wrote for this
//          pass: the quantised pre-test of the builder drops most of the ~85%
//          of candidates outside the support, and the lanes of a wave run the
//          expensive arithmetic (square root, divisions) together on the
//          entries instead of under divergence.
// Pairs with r^2 > h^2 are skipped: in the reference they add exact zeros.
template <int V, bool TAPS, bool VEL = false, class Body>
__device__ __forceinline__ void gather(V3 xi, float h, const Domain& dm, const int2* __restrict__ cells,
                                       const float4* __restrict__ X, const float4* __restrict__ Vel, uint32_t n,
                                       uint32_t self, uint32_t t, const uint32_t* __restrict__ list_count,
                                       const uint32_t* __restrict__ list, Body&& body, uint32_t& ncand,
                                       uint32_t& nnbr) {
    if constexpr (V == 0) {
        const float h2 = h * h;
        scan_buckets(xi, h, dm, cells, X, self, [&](uint32_t k, float4 xk, V3 r, float r2) {
            if (TAPS) ++ncand;
            if (r2 <= h2) {
                if (TAPS) ++nnbr;
                body(k, xk, r, r2, Vel != nullptr ? Vel[k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f));
            }
        });
    } else {
        gather_list<TAPS, VEL>(xi, h, dm, cells, X, Vel, n, self, t, list_count, list, body, nnbr);
    }
}

// CS_PredictAndHash.glsl:88-130
template <bool B>
__global__ __launch_bounds__(kBlock) void k_predict(StepConst c, const float4* __restrict__ pos,
                                                    float4* __restrict__ vel, float4* __restrict__ xstar,
                                                    uint32_t* __restrict__ keys) {
    const uint32_t id = gid();
    if (id >= c.n) return;
    const float4 p4 = pos[id];
    const float4 v4 = vel[id];
    const V3 p = xyz(p4);
    V3 v = xyz(v4);
    const float dt = c.sub_dt;
    v = add(v, scl(mk(c.g[0], c.g[1], c.g[2]), dt));
    if (c.interacting == 1u) {
        const V3 to = sub(mk(c.ipos[0], c.ipos[1], c.ipos[2]), p);
        const float dist = len(to);
        if (dist < c.irad && dist > 1e-5f) {
            const float falloff = 1.0f - (dist / c.irad);
            const V3 nrm = dvs(to, len(to));
            v = add(v, scl(scl(scl(nrm, c.istr), falloff), dt));
        }
    }
    const V3 xp = add(p, scl(v, dt));
    vel[id] = make_float4(v.x, v.y, v.z, v4.w);
    xstar[id] = make_float4(xp.x, xp.y, xp.z, 0.0f);
    int32_t cx, cy, cz;
    cell_of(xp, c.h, cx, cy, cz);
    const Domain& dom = domain_of<B>(c, id);
    keys[id] = dom.hash_base + cell_hash(cx, cy, cz, dom.hash_mask);
}

// CS_BuildOffsets.glsl:55-78, clearMode = 1
__global__ __launch_bounds__(kBlock) void k_clear_cells(int2* __restrict__ cells, uint32_t H) {
    const uint32_t b = gid();
    if (b < H) cells[b] = make_int2(-1, -1);
}

// CS_BuildOffsets.glsl:55-78, clearMode = 0
__global__ __launch_bounds__(kBlock) void k_build_cells(const uint32_t* __restrict__ skeys, uint32_t n,
                                                        int2* __restrict__ cells) {
    const uint32_t i = gid();
    if (i >= n) return;
    const uint32_t k = skeys[i];
    if (i == 0 || skeys[i - 1] != k) cells[k].x = (int32_t)i;
    if (i == n - 1 || skeys[i + 1] != k) cells[k].y = (int32_t)(i + 1);
}

// Gather the predicted state into key-sorted order (lambda starts at 0,
// CS_PredictAndHash.glsl:121) and bin every slot by the Morton code of its
// cell; the bins give the spatial processing order of the gather passes.
template <bool B>
__global__ __launch_bounds__(kBlock) void k_reorder(StepConst c, const uint32_t* __restrict__ perm,
                                                    const float4* __restrict__ xstar,
                                                    const float4* __restrict__ vel, float4* __restrict__ P,
                                                    uint2* __restrict__ Q, float4* __restrict__ PL,
                                                    float4* __restrict__ vs, uint32_t* __restrict__ bins,
                                                    uint32_t* __restrict__ bin_of, uint32_t* __restrict__ bin_rank) {
    const uint32_t k = gid();
    if (k >= c.n) return;
    const uint32_t id = perm[k];
    const float4 x = xstar[id];
    P[k] = x;
    Q[k] = quantize(xyz(x), c.qscale);
    PL[k] = x;
    vs[k] = vel[id];
    int32_t cx, cy, cz;
    cell_of(xyz(x), c.h, cx, cy, cz);
    const uint32_t b = morton_bin(domain_of<B>(c, k), cx, cy, cz);
    bin_of[k] = b;
    // one atomic per distinct bin in the wave (the members of a hash bucket
    // are consecutive slots and share a bin)
    const uint32_t lane = __lane_id();
    unsigned long long active = __ballot(1);
    uint32_t rank = 0;
    while (active) {
        const int leader = __ffsll((long long)active) - 1;
        const uint32_t lb = __shfl(b, leader);
        const unsigned long long same = __ballot(b == lb) & active;
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(&bins[lb], (uint32_t)__popcll(same));
        base = __shfl(base, leader);
        if ((same >> lane) & 1ull) rank = base + (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
        active &= ~same;
    }
    bin_rank[k] = rank;
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
    const uint32_t i = gid();
    if (i < n) p[i] = v;
}

// Block ranges of an SDF grid (Domain::sdf_brange): block (bx, by, bz) covers
// voxels bx*kSdfBlock .. bx*kSdfBlock + kSdfBlock on each axis (clamped to the
// grid), so that it holds both voxels of every trilinear tap pair whose lower
// voxel lies in the block.  A non-finite voxel makes the block (-inf, +inf).
__global__ __launch_bounds__(kBlock) void k_sdf_block_range(const float* __restrict__ grid, int32_t dx, int32_t dy,
                                                            int32_t dz, int32_t bx, int32_t by, int32_t bz,
                                                            float2* __restrict__ out) {
    const uint32_t b = gid();
    if (b >= (uint32_t)bx * (uint32_t)by * (uint32_t)bz) return;
    const int32_t ix = (int32_t)(b % (uint32_t)bx), iy = (int32_t)((b / (uint32_t)bx) % (uint32_t)by),
                  iz = (int32_t)(b / ((uint32_t)bx * (uint32_t)by));
    const int32_t x0 = ix * kSdfBlock, y0 = iy * kSdfBlock, z0 = iz * kSdfBlock;
    const int32_t x1 = min(x0 + kSdfBlock, dx - 1), y1 = min(y0 + kSdfBlock, dy - 1), z1 = min(z0 + kSdfBlock, dz - 1);
    float lo = __builtin_inff(), hi = -__builtin_inff();
    bool finite = true;
    for (int32_t z = z0; z <= z1; ++z)
        for (int32_t y = y0; y <= y1; ++y)
            for (int32_t x = x0; x <= x1; ++x) {
                const float v = grid[(size_t)x + (size_t)dx * ((size_t)y + (size_t)dy * (size_t)z)];
                finite = finite && fabsf(v) <= 3.402823466e38f;
                lo = fminf(lo, v);
                hi = fmaxf(hi, v);
            }
    out[b] = finite ? make_float2(lo, hi) : make_float2(-__builtin_inff(), __builtin_inff());
}

// Sorted-slot processing order: thread t takes slot t.
__global__ __launch_bounds__(kBlock) void k_order_slots(uint32_t n, uint32_t* __restrict__ order) {
    const uint32_t k = gid();
    if (k < n) order[k] = k;
}

__global__ __launch_bounds__(kBlock) void k_order(uint32_t n, const uint32_t* __restrict__ base,
                                                  const uint32_t* __restrict__ bin_of,
                                                  const uint32_t* __restrict__ bin_rank, uint32_t* __restrict__ order) {
    const uint32_t k = gid();
    if (k >= n) return;
    order[base[bin_of[k]] + bin_rank[k]] = k;
}

// Relabelled storage (processing order 2).  Every hash bucket goes whole
// into the Morton bin of the cell of its first particle, and the particle
// arrays of the substep are stored bin by bin, each bucket contiguous and in
// its sorted-slot order: a bucket's slots [start, end) become the storage
// slots [m, m + end - start), so a gather that visits buckets in the
// reference order and each bucket in slot order sees the same sequence, while
// the particles a wave takes (consecutive storage slots) and their neighbours
// lie close together in memory.  The first kernel sizes the bins (one atomic
// per bucket, on its first slot), the second places every slot and writes the
// cell table over storage slots (mcells).
template <bool B>
__global__ __launch_bounds__(kBlock) void k_bucket_bins(StepConst c, const uint32_t* __restrict__ skeys,
                                                        const uint32_t* __restrict__ perm,
                                                        const float4* __restrict__ xstar,
                                                        const int2* __restrict__ cells, uint32_t* __restrict__ bins,
                                                        uint32_t* __restrict__ bin_of, uint32_t* __restrict__ bin_rank) {
    const uint32_t k = gid();
    if (k >= c.n) return;
    const uint32_t key = skeys[k];
    if (k != 0 && skeys[k - 1] == key) return;  // not the bucket's first slot
    const int2 se = cells[key];
    int32_t cx, cy, cz;
    cell_of(xyz(xstar[perm[k]]), c.h, cx, cy, cz);
    const uint32_t b = morton_bin(domain_of<B>(c, k), cx, cy, cz);
    bin_of[k] = b;
    bin_rank[k] = atomicAdd(&bins[b], (uint32_t)(se.y - se.x));
}

template <bool B>
__global__ __launch_bounds__(kBlock) void k_relabel(StepConst c, const uint32_t* __restrict__ skeys,
                                                    const uint32_t* __restrict__ perm,
                                                    const float4* __restrict__ xstar, const float4* __restrict__ vel,
                                                    const int2* __restrict__ cells, const uint32_t* __restrict__ base,
                                                    const uint32_t* __restrict__ bin_of,
                                                    const uint32_t* __restrict__ bin_rank, float4* __restrict__ P,
                                                    uint2* __restrict__ Q, float4* __restrict__ PL,
                                                    float4* __restrict__ vs, uint32_t* __restrict__ mperm,
                                                    uint32_t* __restrict__ bin_m, int2* __restrict__ mcells) {
    const uint32_t k = gid();
    if (k >= c.n) return;
    const uint32_t key = skeys[k];
    const int2 se = cells[key];
    const uint32_t head = (uint32_t)se.x;
    const uint32_t b = bin_of[head];
    const uint32_t m = base[b] + bin_rank[head] + (k - head);
    const uint32_t id = perm[k];
    const float4 x = xstar[id];
    P[m] = x;
    Q[m] = quantize(xyz(x), c.qscale);
    if (PL != nullptr) PL[m] = x;  // lambda = 0 (read before the lambda pass writes it only under APBF)
    vs[m] = vel[id];
    mperm[m] = id;
    if (bin_m != nullptr) bin_m[m] = b;  // read by the tiled gathers only
    if (k == head) mcells[key] = make_int2((int32_t)m, (int32_t)(m + (uint32_t)(se.y - se.x)));
}

// CS_ComputeLOD.glsl:66-79
__global__ __launch_bounds__(kBlock) void k_lod(StepConst c, const float4* __restrict__ P,
                                                uint32_t* __restrict__ lod) {
    const uint32_t k = gid();
    if (k >= c.n) return;
    const V3 x = xyz(P[k]);
    const float dist = len(sub(mk(c.cam[0], c.cam[1], c.cam[2]), x));
    float t = dist / c.lod_max_dist;
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    const float m = (float)c.max_lod * (1.0f - t) + (float)c.min_lod * t;
    uint32_t l = (uint32_t)roundf(m);
    l = l < c.min_lod ? c.min_lod : (l > c.max_lod ? c.max_lod : l);
    lod[k] = l;
}

// CS_ComputeLambdas.glsl:108-162.  In the list form (V = 1) the neighbours
// come from the list k_build_list wrote for these positions; the delta-p pass
// of the same iteration walks it again.
template <int V, bool TAPS, bool B, int BS>
__global__ __launch_bounds__(BS) void k_lambda(StepConst c, const uint32_t* __restrict__ order,
                                                   const float4* __restrict__ P, float4* __restrict__ PL,
                                                   float* __restrict__ rho_out, const uint32_t* __restrict__ lod,
                                                   const int2* __restrict__ cells, uint32_t iter,
                                                   const uint32_t* __restrict__ list_count,
                                                   const uint32_t* __restrict__ list, uint32_t* __restrict__ cand_out,
                                                   uint32_t* __restrict__ nbr_out) {
    const uint32_t t = xcd_gid_bs<BS>();
    if (t >= c.n) return;
    const uint32_t k = slot_at(order, t);
    const V3 xi = xyz(P[k]);
    if (c.enable_apbf != 0u && lod[k] < iter) {
        // inactive: keep the last lambda, refresh the position
        PL[k] = make_float4(xi.x, xi.y, xi.z, PL[k].w);
        return;
    }
    const float inv = c.inv_rho0;
    float rho = c.w0_self;
    float sum_g2 = 0.0f;
    V3 gi = mk(0.0f, 0.0f, 0.0f);
    uint32_t ncand = 0, nnbr = 0;
    auto body = [&](uint32_t, float4, V3 r, float r2, float4) {
        rho = rho + poly6_r2(c, r2);
        const V3 gw = spiky(c, r, r2);
        const float gj2 = dot(gw, gw);
        sum_g2 = sum_g2 + gj2 * (inv * inv);
        gi = add(gi, gw);
    };
    gather<V, TAPS>(xi, c.h, domain_of<B>(c, k), cells, P, nullptr, c.n, k, t, list_count, list, body, ncand,
                    nnbr);
    gi = scl(gi, inv);
    float C = rho / c.rho0 - 1.0f;
    const float floorc = -c.cohesion;
    C = C > floorc ? C : floorc;
    sum_g2 = sum_g2 + dot(gi, gi);
    const float lambda = -C / (sum_g2 + c.eps);
    PL[k] = make_float4(xi.x, xi.y, xi.z, lambda);
    rho_out[k] = rho;
    if (TAPS) {
        if (V == 0) cand_out[k] = ncand;  // the list form counts them in k_build_list
        nbr_out[k] = nnbr;
    }
}

// CS_ComputeDeltaP.glsl:120-166 fused with the clamp and add of
// CS_ApplyDeltaP.glsl:80-106 and, for a single scene, the SDF projection of
// :108-121 (in the batched form the scene's Domain comes from memory, and its
// projection would double the registers of the gather: k_apply_sdf follows
// there).  Reads (x, lambda) from PL, writes the corrected position to P.  In
// the list form the neighbours come from the list the lambda pass of this
// iteration walked (same positions, same order).
template <int V, bool TAPS, bool B, int BS>
__global__ __launch_bounds__(BS) void k_deltap(StepConst c, const uint32_t* __restrict__ order,
                                                   const float4* __restrict__ PL, float4* __restrict__ P,
                                                   uint2* __restrict__ Q, const uint32_t* __restrict__ lod,
                                                   const int2* __restrict__ cells, uint32_t iter,
                                                   const uint32_t* __restrict__ list_count,
                                                   const uint32_t* __restrict__ list, float4* __restrict__ dp_tap) {
    const uint32_t t = xcd_gid_bs<BS>();
    if (t >= c.n) return;
    const uint32_t k = slot_at(order, t);
    if (c.enable_apbf != 0u && lod[k] < iter) return;  // P[k] already holds x
    const float4 xi4 = PL[k];
    const V3 xi = xyz(xi4);
    const float li = xi4.w;
    V3 dp = mk(0.0f, 0.0f, 0.0f);
    auto body = [&](float lj, V3 r, float r2) {
        const V3 gw = spiky(c, r, r2);
        const float sc = scorr_r2(c, r2);
        dp = add(dp, scl(gw, li + lj + sc));
    };
    uint32_t ncand = 0, nnbr = 0;
    gather<V, false>(xi, c.h, domain_of<B>(c, k), cells, PL, nullptr, c.n, k, t, list_count, list,
                     [&](uint32_t, float4 xj, V3 r, float r2, float4) { body(xj.w, r, r2); }, ncand, nnbr);
    dp = scl(dp, c.inv_rho0);
    if (TAPS) dp_tap[k] = make_float4(dp.x, dp.y, dp.z, 0.0f);
    // apply: clamp |dp| <= h/2, x += dp, SDF projection
    const float maxdp = 0.5f * c.h;
    const float l2 = dot(dp, dp);
    const float max2 = maxdp * maxdp;
    if (l2 > max2) {
        const float l = sqrtf(l2);
        dp = scl(dp, maxdp / l);
    }
    V3 p = add(xi, dp);
    with_domain<B>(c, k, [&](const Domain& dom) {
        if (dom.enable_sdf == 1u) {
            V3 nrm;
            sdf_project(dom, p, nrm);
        }
    });
    P[k] = make_float4(p.x, p.y, p.z, 0.0f);
    Q[k] = quantize(p, c.qscale);
}

// The SDF projection of CS_ApplyDeltaP.glsl:108-121 on the corrected
// positions, in place, after the tiled delta-p gathers (k_deltap projects
// itself).
template <bool B>
__global__ __launch_bounds__(kBlock) void k_apply_sdf(StepConst c, const uint32_t* __restrict__ order,
                                                      float4* __restrict__ P, uint2* __restrict__ Q,
                                                      const uint32_t* __restrict__ lod, uint32_t iter) {
    const uint32_t t = xcd_gid();
    if (t >= c.n) return;
    const uint32_t k = slot_at(order, t);
    if (c.enable_apbf != 0u && lod[k] < iter) return;
    const Domain& dom = domain_of<B>(c, k);
    if (dom.enable_sdf != 1u) return;
    V3 p = xyz(P[k]);
    V3 nrm;
    if (sdf_project(dom, p, nrm)) {
        P[k] = make_float4(p.x, p.y, p.z, 0.0f);
        Q[k] = quantize(p, c.qscale);
    }
}

// The rest of CS_Integrate.glsl:206-280 per particle: AABB clamp, SDF
// projection, velocity from the displacement plus the XSPH change, speed
// clamp, wall damping, removal of the normal velocity at an SDF contact; the
// committed state goes to pos / vel (id order) and posS (sorted).  posS.w
// carries |curl| (the |omega_j| of CS_VorticityConfinement.glsl:156, the same
// FP32 length), so that the vorticity gather reads one record per neighbour.
// xp = P[k], vw = vs[k].w, visc_delta and om = |curl| as the gather left them.
__device__ __forceinline__ void finish_particle(const StepConst& c, const Domain& dom, uint32_t k, V3 xp, float vw,
                                                V3 visc_delta, float om, const uint32_t* __restrict__ perm,
                                                float4* __restrict__ pos, float4* __restrict__ vel,
                                                float4* __restrict__ posS, uint2* __restrict__ Q) {
    const uint32_t id = perm[k];
    const float4 op4 = pos[id];
    const V3 old_pos = xyz(op4);
    bool cxh = false, cyh = false, czh = false;
    if (xp.x <= dom.bmin[0]) { xp.x = dom.bmin[0]; cxh = true; }
    if (xp.x >= dom.bmax[0]) { xp.x = dom.bmax[0]; cxh = true; }
    if (xp.y <= dom.bmin[1]) { xp.y = dom.bmin[1]; cyh = true; }
    if (xp.y >= dom.bmax[1]) { xp.y = dom.bmax[1]; cyh = true; }
    if (xp.z <= dom.bmin[2]) { xp.z = dom.bmin[2]; czh = true; }
    if (xp.z >= dom.bmax[2]) { xp.z = dom.bmax[2]; czh = true; }
    bool hit = false;
    V3 nrm = mk(0.0f, 0.0f, 0.0f);
    if (dom.enable_sdf == 1u) hit = sdf_project(dom, xp, nrm);
    const float inv_dt = (c.sub_dt > 0.0f) ? (1.0f / c.sub_dt) : 0.0f;
    V3 v = scl(sub(xp, old_pos), inv_dt);
    v = add(v, visc_delta);
    const float sp2 = dot(v, v);
    if (sp2 > 2500.0f) v = scl(v, 50.0f / sqrtf(sp2));
    if (cxh) v.x = v.x * -c.damping;
    if (cyh) v.y = v.y * -c.damping;
    if (czh) v.z = v.z * -c.damping;
    if (hit) {
        const float vn = dot(v, nrm);
        v = sub(v, scl(nrm, vn));
    }
    pos[id] = make_float4(xp.x, xp.y, xp.z, op4.w);
    vel[id] = make_float4(v.x, v.y, v.z, vw);
    posS[k] = make_float4(xp.x, xp.y, xp.z, om);
    Q[k] = quantize(xp, c.qscale);  // the vorticity gather's candidates are the committed positions
}

// CS_Integrate.glsl:135-280.  The gather: XSPH sum and curl over the
// neighbours (CS_Integrate.glsl:160-205); neighbour velocities come from the
// sorted post-predict snapshot vs (the Jacobi form of the racy read at :171).
// It writes the curl and goes on with finish_particle (FIN), or writes the
// XSPH velocity change for k_integrate_finish.
template <int V, bool B, int BS, bool FIN>
__global__ __launch_bounds__(BS) void k_integrate(StepConst c, const uint32_t* __restrict__ order,
                                                      const float4* __restrict__ P, const float4* __restrict__ vs,
                                                      const int2* __restrict__ cells,
                                                      const uint32_t* __restrict__ list_count,
                                                      const uint32_t* __restrict__ list, float4* __restrict__ vdelta,
                                                      float4* __restrict__ curl_out, const uint32_t* __restrict__ perm,
                                                      float4* __restrict__ pos, float4* __restrict__ vel,
                                                      float4* __restrict__ posS, uint2* __restrict__ Q) {
    const uint32_t t = xcd_gid_bs<BS>();
    if (t >= c.n) return;
    const uint32_t k = slot_at(order, t);
    const float4 vs4 = vs[k];
    const V3 old_vel = xyz(vs4);
    const V3 xp = xyz(P[k]);
    const float inv = c.inv_rho0;
    const bool do_visc = c.enable_visc == 1u;
    const bool do_vort = c.enable_vort == 1u;
    V3 visc = mk(0.0f, 0.0f, 0.0f);
    V3 curl = mk(0.0f, 0.0f, 0.0f);
    uint32_t ncand = 0, nnbr = 0;
    gather<V, false, true>(xp, c.h, domain_of<B>(c, k), cells, P, vs, c.n, k, t, list_count, list,
                     [&](uint32_t, float4, V3 r, float r2, float4 vj) {
        const V3 vij = sub(xyz(vj), old_vel);
        if (do_visc) {
            const float w = poly6_r2(c, r2);
            visc = add(visc, scl(scl(vij, w), inv));
        }
        if (do_vort) {
            const V3 gw = spiky(c, r, r2);
            curl = add(curl, cross(vij, gw));
        }
    }, ncand, nnbr);
    const V3 vd = do_visc ? scl(visc, c.viscosity) : mk(0.0f, 0.0f, 0.0f);
    const V3 cw = scl(curl, inv);
    if (do_vort) curl_out[k] = make_float4(cw.x, cw.y, cw.z, 0.0f);
    if constexpr (FIN) {
        const float om = do_vort ? len(cw) : 0.0f;
        with_domain<B>(c, k, [&](const Domain& dom) {
            finish_particle(c, dom, k, xp, vs4.w, vd, om, perm, pos, vel, posS, Q);
        });
    } else {
        if (do_visc) vdelta[k] = make_float4(vd.x, vd.y, vd.z, 0.0f);
    }
}

// finish_particle on frames with neither viscosity nor vorticity (no gather).
template <bool B>
__global__ __launch_bounds__(kBlock) void k_integrate_finish(StepConst c, const uint32_t* __restrict__ order,
                                                             const float4* __restrict__ P,
                                                             const float4* __restrict__ vs,
                                                             const float4* __restrict__ vdelta,
                                                             const uint32_t* __restrict__ perm,
                                                             const float4* __restrict__ curl,
                                                             float4* __restrict__ pos, float4* __restrict__ vel,
                                                             float4* __restrict__ posS, uint2* __restrict__ Q) {
    const uint32_t t = xcd_gid();
    if (t >= c.n) return;
    const uint32_t k = slot_at(order, t);
    const V3 visc_delta = c.enable_visc == 1u ? xyz(vdelta[k]) : mk(0.0f, 0.0f, 0.0f);
    const float om = c.enable_vort == 1u ? len(xyz(curl[k])) : 0.0f;
    with_domain<B>(c, k, [&](const Domain& dom) {
        finish_particle(c, dom, k, xyz(P[k]), vs[k].w, visc_delta, om, perm, pos, vel, posS, Q);
    });
}

// CS_VorticityConfinement.glsl:133-167 (pass 1)
template <int V, bool B, int BS>
__global__ __launch_bounds__(BS) void k_vorticity(StepConst c, const uint32_t* __restrict__ order,
                                                      const float4* __restrict__ posS, const float4* __restrict__ curl,
                                                      const uint32_t* __restrict__ perm,
                                                      float4* __restrict__ vel, const int2* __restrict__ cells,
                                                      const uint32_t* __restrict__ list_count,
                                                      const uint32_t* __restrict__ list) {
    const uint32_t t = xcd_gid_bs<BS>();
    if (t >= c.n) return;
    const uint32_t k = slot_at(order, t);
    
    const V3 pi = xyz(posS[k]);
    V3 eta = mk(0.0f, 0.0f, 0.0f);
    uint32_t ncand = 0, nnbr = 0;
    gather<V, false>(pi, c.h, domain_of<B>(c, k), cells, posS, nullptr, c.n, k, t, list_count, list,
                     [&](uint32_t, float4 xj, V3 r, float r2, float4) {
        const V3 gw = spiky(c, r, r2);
        eta = add(eta, scl(gw, xj.w));  // xj.w = |omega_j|
    }, ncand, nnbr);
    eta = scl(eta, c.inv_rho0);
    const float el = len(eta);
    if (el < 1e-6f) return;
    const V3 nn = dvs(eta, el);
    const V3 om = xyz(curl[k]);
    const V3 dv = scl(scl(cross(nn, om), c.vort_eps), c.sub_dt);
    const uint32_t id = perm[k];
    const float4 v4 = vel[id];
    vel[id] = make_float4(v4.x + dv.x, v4.y + dv.y, v4.z + dv.z, v4.w);
}

// CS_ContactFilter.glsl:68-83, as a deterministic compaction in id order.
template <bool B>
__global__ __launch_bounds__(kBlock) void k_contact_flag(StepConst c, const float4* __restrict__ pos, float band,
                                                         uint32_t* __restrict__ flags) {
    const uint32_t id = gid();
    if (id >= c.n) return;
    const Domain& dom = domain_of<B>(c, id);
    if (dom.enable_sdf != 1u) {
        flags[id] = 0u;
        return;
    }
    const float d = sdf_query(dom, xyz(pos[id]));
    const float ad = d < 0.0f ? -d : d;
    flags[id] = ad < band ? 1u : 0u;
}

template <bool B>
__global__ __launch_bounds__(kBlock) void k_contact_write(StepConst c, const float4* __restrict__ pos,
                                                          const float4* __restrict__ vel,
                                                          const uint32_t* __restrict__ flags,
                                                          const uint32_t* __restrict__ slots,
                                                          float4* __restrict__ out, uint32_t* __restrict__ count) {
    const uint32_t id = gid();
    if (id >= c.n) return;
    const bool hit = flags[id] != 0u;
    if (hit) {
        const uint32_t s = slots[id];
        const float4 p4 = pos[id];
        out[2 * (size_t)s] = make_float4(p4.x, p4.y, p4.z, 0.0f);
        out[2 * (size_t)s + 1] = vel[id];
    }
    if (id == c.n - 1) count[0] = slots[id] + (hit ? 1u : 0u);
}

// ---- tiled gathers (gather form 2) -----------------------------------------
// The domains are binned by Morton bins of 4x4x4 cells; a tile is a run of 2^M
// consecutive bins (the cell-tile shape, a tuning knob): 4x4x4 cells for M = 0,
// 8x4x4 for M = 1, 8x8x4 for M = 2.  One workgroup takes a tile: it loads the
// cell-table entries of the tile's cells and a one-cell halo around them and
// copies those buckets' records into LDS once, in bucket order; every particle
// of the tile whose current cell is inside the tile then visits its 27 buckets
// from LDS, in the reference order (an aliased bucket appears once per halo
// cell that hashes to it, as it is visited once per occurrence).  A particle
// that has moved out of its tile, or a tile whose buckets exceed the LDS
// capacity, takes the global single-loop gather.  The pass is persistent:
// kTileGrid workgroups take tiles from a queue per XCD, each XCD over a
// contiguous Morton range, until the range is done, so tiles of uneven
// occupancy (uneven scenes in a batch) balance within the XCD.

template <int M>
struct TileShape {
    static constexpr int TX = (1 << kTileShift) << ((M + 2) / 3);
    static constexpr int TY = (1 << kTileShift) << ((M + 1) / 3);
    static constexpr int TZ = (1 << kTileShift) << (M / 3);
    static constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2;
    static constexpr int HC = HX * HY * HZ;                   // halo cells
    static constexpr uint32_t Cap = M == 0 ? 2560u : M == 1 ? 3584u : 5120u;  // staged records
};
constexpr uint32_t kTileGrid = 768;  // persistent workgroups: 3 per CU, a multiple of 8
constexpr uint32_t kDeferGrid = 256; // workgroups of the deferred passes

template <int M>
struct TileStage {
    int32_t o[3];  // first cell of the tile
    uint32_t total;
    uint32_t next;  // the tile this workgroup took from its queue
    int32_t hstart[TileShape<M>::HC];  // first slot of the bucket of halo cell h, -1 empty
    uint32_t hcnt[TileShape<M>::HC];
    uint32_t hoff[TileShape<M>::HC];   // its first record in x[]
    float4 x[TileShape<M>::Cap];
};

// The next tile of this workgroup's XCD range (tiles [nt*x/8, nt*(x+1)/8) for
// XCD x), taken from the XCD's queue head heads[x] (zeroed before the pass),
// or UINT32_MAX when the range is done; the same for the whole workgroup.  The
// caller's previous tile ended with a barrier, so ts.next is free.
template <int M>
__device__ __forceinline__ uint32_t next_tile(TileStage<M>& ts, uint32_t nt, uint32_t* __restrict__ heads) {
    const uint32_t x = blockIdx.x & 7u;
    const uint32_t begin = (uint32_t)(((uint64_t)nt * x) >> 3);
    const uint32_t end = (uint32_t)(((uint64_t)nt * (x + 1)) >> 3);
    if (threadIdx.x == 0) ts.next = begin + atomicAdd(&heads[x], 1u);
    __syncthreads();
    const uint32_t t = ts.next;
    return t < end ? t : UINT32_MAX;
}

// Stages the tile whose first particle is at [start] of the processing order;
// returns whether its buckets fit.  Ends with a barrier.
template <int M, bool B>
__device__ bool stage_tile(const StepConst& c, TileStage<M>& ts, const float4* __restrict__ X,
                           const int2* __restrict__ cells, const uint32_t* __restrict__ order,
                           const uint32_t* __restrict__ bin_of, uint32_t start) {
    using S = TileShape<M>;
    const uint32_t k0 = slot_at(order, start);
    const Domain& dm = domain_of<B>(c, k0);
    const uint32_t code = (bin_of[k0] - dm.bin_base) & ~((1u << M) - 1u);  // the tile's first bin
    const int32_t ox = dm.dcmin[0] + (int32_t)(compact7(code) << kTileShift);
    const int32_t oy = dm.dcmin[1] + (int32_t)(compact7(code >> 1) << kTileShift);
    const int32_t oz = dm.dcmin[2] + (int32_t)(compact7(code >> 2) << kTileShift);
    const uint32_t tid = threadIdx.x;
    for (uint32_t hc = tid; hc < (uint32_t)S::HC; hc += kBlock) {
        const int32_t hx = (int32_t)(hc % S::HX), hy = (int32_t)((hc / S::HX) % S::HY),
                      hz = (int32_t)(hc / (S::HX * S::HY));
        const int2 se = cells[dm.hash_base + cell_hash(ox - 1 + hx, oy - 1 + hy, oz - 1 + hz, dm.hash_mask)];
        ts.hstart[hc] = se.x;
        ts.hcnt[hc] = se.x < 0 ? 0u : (uint32_t)(se.y - se.x);
    }
    if (tid == 0) {
        ts.o[0] = ox;
        ts.o[1] = oy;
        ts.o[2] = oz;
    }
    __syncthreads();
    if (tid < 64) {  // wave 0: exclusive scan of the counts, kPer per lane
        constexpr int kPer = (S::HC + 63) / 64;
        uint32_t local = 0;
        for (int j = 0; j < kPer; ++j) {
            const uint32_t h = kPer * tid + j;
            if (h < (uint32_t)S::HC) local += ts.hcnt[h];
        }
        uint32_t v = local;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t u = __shfl_up(v, d, 64);
            if ((int)tid >= d) v += u;
        }
        uint32_t run = v - local;
        for (int j = 0; j < kPer; ++j) {
            const uint32_t h = kPer * tid + j;
            if (h < (uint32_t)S::HC) {
                ts.hoff[h] = run;
                run += ts.hcnt[h];
            }
        }
        if (tid == 63) ts.total = v;
    }
    __syncthreads();
    const uint32_t total = ts.total;
    const bool ok = total <= S::Cap;
    if (ok) {
        for (uint32_t e = tid; e < total; e += kBlock) {
            uint32_t lo = 0, hi = S::HC - 1;  // last halo cell whose records start at or before e
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (ts.hoff[mid] <= e) lo = mid;
                else hi = mid - 1;
            }
            ts.x[e] = X[ts.hstart[lo] + (int32_t)(e - ts.hoff[lo])];
        }
    }
    __syncthreads();
    return ok;
}

// Whether the cell (lx, ly, lz) relative to the tile is inside it.
template <int M>
__device__ __forceinline__ bool in_tile(int32_t lx, int32_t ly, int32_t lz) {
    using S = TileShape<M>;
    return lx >= 0 && lx < S::TX && ly >= 0 && ly < S::TY && lz >= 0 && lz < S::TZ;
}

// Phase 1 of the two-phase gather over the staged tile: as collect_flush, with
// candidates read from LDS and their record numbers in x[] collected.  (lx,
// ly, lz) is the particle's cell relative to the tile, inside it.
template <int M, bool TAPS, class Flush>
__device__ __forceinline__ uint32_t collect_flush_tile(const TileStage<M>& ts, V3 x, float h, int32_t lx, int32_t ly,
                                                       int32_t lz, uint32_t self, uint16_t* col, Flush&& flush,
                                                       uint32_t& ncand) {
    const float h2 = h * h;
    uint32_t cnt = 0, total = 0;
#pragma unroll 1
    for (int q = 0; q < 27; ++q) {
        const int hc = (lx + q % 3) + TileShape<M>::HX * (ly + (q / 3) % 3) +
                       TileShape<M>::HX * TileShape<M>::HY * (lz + q / 9);
        const uint32_t nb = ts.hcnt[hc];
        const uint32_t base = ts.hoff[hc];
        const uint32_t s0 = (uint32_t)ts.hstart[hc];
        for (uint32_t e = 0; e < nb; ++e) {
            if (s0 + e == self) continue;
            if (TAPS) ++ncand;
            const V3 r = sub(x, xyz(ts.x[base + e]));
            if (dot(r, r) <= h2) {
                col[cnt * kBlock] = (uint16_t)(base + e);
                if (++cnt == kLdsCap) {
                    flush(total, cnt);
                    total += cnt;
                    cnt = 0;
                }
            }
        }
        if (__any(cnt >= kFlushAt)) {
            flush(total, cnt);
            total += cnt;
            cnt = 0;
        }
    }
    flush(total, cnt);
    return total + cnt;
}

// The end of CS_ComputeLambdas.glsl:108-162 for one particle, given its sums.
template <bool TAPS>
__device__ __forceinline__ void lambda_store(const StepConst& c, uint32_t k, V3 xi, float rho, float sum_g2, V3 gi,
                                             uint32_t ncand, uint32_t nnbr, float4* __restrict__ PL,
                                             float* __restrict__ rho_out, uint32_t* __restrict__ cand_out,
                                             uint32_t* __restrict__ nbr_out) {
    gi = scl(gi, c.inv_rho0);
    float C = rho / c.rho0 - 1.0f;
    const float floorc = -c.cohesion;
    C = C > floorc ? C : floorc;
    sum_g2 = sum_g2 + dot(gi, gi);
    const float lambda = -C / (sum_g2 + c.eps);
    PL[k] = make_float4(xi.x, xi.y, xi.z, lambda);
    rho_out[k] = rho;
    if (TAPS) {
        cand_out[k] = ncand;
        nbr_out[k] = nnbr;
    }
}

// CS_ComputeLambdas.glsl:108-162, tiled.  Particles outside their tile are
// appended to defer[] for k_lambda_deferred.
template <int M, bool TAPS, bool B>
__global__ __launch_bounds__(kBlock) void k_lambda_tile(StepConst c, const uint32_t* __restrict__ tiles,
                                                        const uint32_t* __restrict__ ntiles,
                                                        const uint32_t* __restrict__ order,
                                                        const uint32_t* __restrict__ bin_of,
                                                        const float4* __restrict__ P, float4* __restrict__ PL,
                                                        float* __restrict__ rho_out, const uint32_t* __restrict__ lod,
                                                        const int2* __restrict__ cells, uint32_t iter,
                                                        uint32_t* __restrict__ cand_out, uint32_t* __restrict__ nbr_out,
                                                        uint32_t* __restrict__ defer, uint32_t* __restrict__ ndefer) {
    __shared__ TileStage<M> ts;
    __shared__ uint16_t lst[kLdsCap * kBlock];
    uint16_t* col = &lst[threadIdx.x];
    const uint32_t nt = *ntiles;
    const float inv = c.inv_rho0;
    for (uint32_t tile; (tile = next_tile<M>(ts, nt, ndefer + 1)) != UINT32_MAX;) {
        const uint32_t start = tiles[tile], end = tiles[tile + 1];
        const bool ok = stage_tile<M, B>(c, ts, P, cells, order, bin_of, start);
        for (uint32_t t = start + threadIdx.x; t < end; t += kBlock) {
            const uint32_t k = slot_at(order, t);
            const V3 xi = xyz(P[k]);
            if (c.enable_apbf != 0u && lod[k] < iter) {
                PL[k] = make_float4(xi.x, xi.y, xi.z, PL[k].w);
                continue;
            }
            int32_t cx, cy, cz;
            cell_of(xi, c.h, cx, cy, cz);
            const int32_t lx = cx - ts.o[0], ly = cy - ts.o[1], lz = cz - ts.o[2];
            if (!(ok && in_tile<M>(lx, ly, lz))) {
                defer[atomicAdd(ndefer, 1u)] = k;
                continue;
            }
            float rho = c.w0_self;
            float sum_g2 = 0.0f;
            V3 gi = mk(0.0f, 0.0f, 0.0f);
            uint32_t ncand = 0;
            const uint32_t nnbr = collect_flush_tile<M, TAPS>(ts, xi, c.h, lx, ly, lz, k, col, [&](uint32_t, uint32_t cnt) {
#pragma unroll 1
                for (uint32_t e = 0; e < cnt; ++e) {
                    const V3 d = sub(xi, xyz(ts.x[col[e * kBlock]]));
                    const float r2 = dot(d, d);
                    rho = rho + poly6_r2(c, r2);
                    const V3 gw = spiky(c, d, r2);
                    const float gj2 = dot(gw, gw);
                    sum_g2 = sum_g2 + gj2 * (inv * inv);
                    gi = add(gi, gw);
                }
            }, ncand);
            lambda_store<TAPS>(c, k, xi, rho, sum_g2, gi, ncand, nnbr, PL, rho_out, cand_out, nbr_out);
        }
        __syncthreads();
    }
}

// The particles the tiled lambda pass deferred, single-loop gather from HBM.
template <bool TAPS, bool B>
__global__ __launch_bounds__(kBlock) void k_lambda_deferred(StepConst c, const uint32_t* __restrict__ defer,
                                                            const uint32_t* __restrict__ ndefer,
                                                            const float4* __restrict__ P, float4* __restrict__ PL,
                                                            float* __restrict__ rho_out, const int2* __restrict__ cells,
                                                            uint32_t* __restrict__ cand_out,
                                                            uint32_t* __restrict__ nbr_out) {
    const uint32_t nd = *ndefer;
    const float inv = c.inv_rho0;
    const float h2 = c.h * c.h;
    for (uint32_t i = gid(); i < nd; i += gridDim.x * kBlock) {
        const uint32_t k = defer[i];
        const V3 xi = xyz(P[k]);
        float rho = c.w0_self;
        float sum_g2 = 0.0f;
        V3 gi = mk(0.0f, 0.0f, 0.0f);
        uint32_t ncand = 0, nnbr = 0;
        scan_buckets(xi, c.h, domain_of<B>(c, k), cells, P, k, [&](uint32_t, float4, V3 r, float r2) {
            if (TAPS) ++ncand;
            if (r2 <= h2) {
                if (TAPS) ++nnbr;
                rho = rho + poly6_r2(c, r2);
                const V3 gw = spiky(c, r, r2);
                const float gj2 = dot(gw, gw);
                sum_g2 = sum_g2 + gj2 * (inv * inv);
                gi = add(gi, gw);
            }
        });
        lambda_store<TAPS>(c, k, xi, rho, sum_g2, gi, ncand, nnbr, PL, rho_out, cand_out, nbr_out);
    }
}

// The clamp and add of CS_ApplyDeltaP.glsl:80-106 for one particle.
template <bool TAPS>
__device__ __forceinline__ void deltap_store(const StepConst& c, uint32_t k, V3 xi, V3 dp, float4* __restrict__ P,
                                             uint2* __restrict__ Q, float4* __restrict__ dp_tap) {
    dp = scl(dp, c.inv_rho0);
    if (TAPS) dp_tap[k] = make_float4(dp.x, dp.y, dp.z, 0.0f);
    const float maxdp = 0.5f * c.h;
    const float l2 = dot(dp, dp);
    const float max2 = maxdp * maxdp;
    if (l2 > max2) {
        const float l = sqrtf(l2);
        dp = scl(dp, maxdp / l);
    }
    const V3 p = add(xi, dp);
    P[k] = make_float4(p.x, p.y, p.z, 0.0f);
    Q[k] = quantize(p, c.qscale);
}

// CS_ComputeDeltaP.glsl:120-166 + the clamp and add of CS_ApplyDeltaP.glsl:80-106,
// tiled; particles outside their tile go to defer[] for k_deltap_deferred.
template <int M, bool TAPS, bool B>
__global__ __launch_bounds__(kBlock) void k_deltap_tile(StepConst c, const uint32_t* __restrict__ tiles,
                                                        const uint32_t* __restrict__ ntiles,
                                                        const uint32_t* __restrict__ order,
                                                        const uint32_t* __restrict__ bin_of,
                                                        const float4* __restrict__ PL, float4* __restrict__ P,
                                                        uint2* __restrict__ Q, const uint32_t* __restrict__ lod,
                                                        const int2* __restrict__ cells, uint32_t iter,
                                                        float4* __restrict__ dp_tap, uint32_t* __restrict__ defer,
                                                        uint32_t* __restrict__ ndefer) {
    __shared__ TileStage<M> ts;
    __shared__ uint16_t lst[kLdsCap * kBlock];
    uint16_t* col = &lst[threadIdx.x];
    const uint32_t nt = *ntiles;
    for (uint32_t tile; (tile = next_tile<M>(ts, nt, ndefer + 1)) != UINT32_MAX;) {
        const uint32_t start = tiles[tile], end = tiles[tile + 1];
        const bool ok = stage_tile<M, B>(c, ts, PL, cells, order, bin_of, start);
        for (uint32_t t = start + threadIdx.x; t < end; t += kBlock) {
            const uint32_t k = slot_at(order, t);
            if (c.enable_apbf != 0u && lod[k] < iter) continue;  // P[k] already holds x
            const float4 xi4 = PL[k];
            const V3 xi = xyz(xi4);
            int32_t cx, cy, cz;
            cell_of(xi, c.h, cx, cy, cz);
            const int32_t lx = cx - ts.o[0], ly = cy - ts.o[1], lz = cz - ts.o[2];
            if (!(ok && in_tile<M>(lx, ly, lz))) {
                defer[atomicAdd(ndefer, 1u)] = k;
                continue;
            }
            const float li = xi4.w;
            V3 dp = mk(0.0f, 0.0f, 0.0f);
            uint32_t ncand = 0;
            collect_flush_tile<M, false>(ts, xi, c.h, lx, ly, lz, k, col, [&](uint32_t, uint32_t cnt) {
#pragma unroll 1
                for (uint32_t e = 0; e < cnt; ++e) {
                    const float4 a = ts.x[col[e * kBlock]];
                    const V3 d = sub(xi, xyz(a));
                    const float r2 = dot(d, d);
                    const V3 gw = spiky(c, d, r2);
                    const float sc = scorr_r2(c, r2);
                    dp = add(dp, scl(gw, li + a.w + sc));
                }
            }, ncand);
            deltap_store<TAPS>(c, k, xi, dp, P, Q, dp_tap);
        }
        __syncthreads();
    }
}

template <bool TAPS, bool B>
__global__ __launch_bounds__(kBlock) void k_deltap_deferred(StepConst c, const uint32_t* __restrict__ defer,
                                                            const uint32_t* __restrict__ ndefer,
                                                            const float4* __restrict__ PL, float4* __restrict__ P,
                                                            uint2* __restrict__ Q, const int2* __restrict__ cells,
                                                            float4* __restrict__ dp_tap) {
    const uint32_t nd = *ndefer;
    const float h2 = c.h * c.h;
    for (uint32_t i = gid(); i < nd; i += gridDim.x * kBlock) {
        const uint32_t k = defer[i];
        const float4 xi4 = PL[k];
        const V3 xi = xyz(xi4);
        const float li = xi4.w;
        V3 dp = mk(0.0f, 0.0f, 0.0f);
        scan_buckets(xi, c.h, domain_of<B>(c, k), cells, PL, k, [&](uint32_t, float4 xj, V3 r, float r2) {
            if (r2 <= h2) {
                const V3 gw = spiky(c, r, r2);
                const float sc = scorr_r2(c, r2);
                dp = add(dp, scl(gw, li + xj.w + sc));
            }
        });
        deltap_store<TAPS>(c, k, xi, dp, P, Q, dp_tap);
    }
}

// Tile list: the positions of the processing order where the tile (the bin
// without its low m Morton bits) changes.
__global__ __launch_bounds__(kBlock) void k_tile_flags(uint32_t n, uint32_t m, const uint32_t* __restrict__ order,
                                                       const uint32_t* __restrict__ bin_of,
                                                       uint32_t* __restrict__ tflag) {
    const uint32_t t = gid();
    if (t >= n) return;
    tflag[t] = (t == 0 || (bin_of[slot_at(order, t)] >> m) != (bin_of[slot_at(order, t - 1)] >> m)) ? 1u : 0u;
}

__global__ __launch_bounds__(kBlock) void k_tile_list(uint32_t n, const uint32_t* __restrict__ tflag,
                                                      const uint32_t* __restrict__ tidx, uint32_t* __restrict__ tiles,
                                                      uint32_t* __restrict__ ntiles) {
    const uint32_t t = gid();
    if (t >= n) return;
    if (tflag[t]) tiles[tidx[t]] = t;
    if (t == n - 1) {
        const uint32_t nt = tidx[t] + tflag[t];
        ntiles[0] = nt;
        tiles[nt] = n;
    }
}

}  // namespace

#define L_T(B) L(true, B)
#define L_F(B) L(false, B)
#define FSI_B_SWITCH(b, CALL) \
    if (b) {                      \
        CALL(true);               \
    } else {                      \
        CALL(false);              \
    }

hipError_t launch_predict(const StepConst& c, DeviceState& d, hipStream_t st) {
#define L(B) hipLaunchKernelGGL((k_predict<B>), grid_for(c.n), dim3(kBlock), 0, st, c, d.pos, d.vel, d.xstar, d.keys);
    FSI_B_SWITCH(d.batched, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_cells(const StepConst& c, DeviceState& d, hipStream_t st) {
    hipLaunchKernelGGL(k_clear_cells, grid_for(d.hash_size), dim3(kBlock), 0, st, d.cells, d.hash_size);
    hipLaunchKernelGGL(k_build_cells, grid_for(c.n), dim3(kBlock), 0, st, d.skeys, c.n, d.cells);
    return hipGetLastError();
}

// The staged-tile gathers take Morton tiles of the processing order.
static inline bool tiled_form(const DeviceState& d) { return d.variant == 2 && d.tiled && d.morton; }

// What the passes after the reorder read: the cell table over the storage
// slots, the storage slot -> particle id map and the Morton bin of each
// storage slot (sorted slots unless the storage is relabelled).
static inline const int2* gcells(const DeviceState& d) { return d.relabel ? d.mcells : d.cells; }
static inline const uint32_t* sperm(const DeviceState& d) { return d.relabel ? d.mperm : d.perm; }
static inline const uint32_t* sbins(const DeviceState& d) { return d.relabel ? d.bin_m : d.bin_of; }
static inline const uint32_t* sorder(const DeviceState& d) { return d.relabel ? nullptr : d.order; }

hipError_t launch_reorder(const StepConst& c, DeviceState& d, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_u32, grid_for(kOrderBins), dim3(kBlock), 0, st, d.bins, kOrderBins, 0u);
    hipError_t e = hipSuccess;
    d.stored_relabel = d.relabel;
    if (d.relabel) {
        hipLaunchKernelGGL(k_clear_cells, grid_for(d.hash_size), dim3(kBlock), 0, st, d.mcells, d.hash_size);
#define L(B)                                                                                                        \
    hipLaunchKernelGGL((k_bucket_bins<B>), grid_for(c.n), dim3(kBlock), 0, st, c, d.skeys, d.perm, d.xstar, d.cells, \
                       d.bins, d.bin_of, d.bin_rank);
        FSI_B_SWITCH(d.batched, L)
#undef L
        e = scan_exclusive_u32(d.bins, d.bins, kOrderBins, d.sort.partial, st);
        if (e != hipSuccess) return e;
#define L(B)                                                                                                    \
    hipLaunchKernelGGL((k_relabel<B>), grid_for(c.n), dim3(kBlock), 0, st, c, d.skeys, d.perm, d.xstar, d.vel,  \
                       d.cells, d.bins, d.bin_of, d.bin_rank, d.P, d.Q, c.enable_apbf ? d.PL : nullptr, d.vs,   \
                       d.mperm, tiled_form(d) ? d.bin_m : nullptr, d.mcells);
        FSI_B_SWITCH(d.batched, L)
#undef L
    } else {
#define L(B)                                                                                                    \
    hipLaunchKernelGGL((k_reorder<B>), grid_for(c.n), dim3(kBlock), 0, st, c, d.perm, d.xstar, d.vel, d.P, d.Q, \
                       d.PL, d.vs, d.bins, d.bin_of, d.bin_rank);
        FSI_B_SWITCH(d.batched, L)
#undef L
        e = scan_exclusive_u32(d.bins, d.bins, kOrderBins, d.sort.partial, st);
        if (e != hipSuccess) return e;
        if (d.morton)
            hipLaunchKernelGGL(k_order, grid_for(c.n), dim3(kBlock), 0, st, c.n, d.bins, d.bin_of, d.bin_rank,
                               d.order);
        else
            hipLaunchKernelGGL(k_order_slots, grid_for(c.n), dim3(kBlock), 0, st, c.n, d.order);
    }
    if (tiled_form(d)) {
        hipLaunchKernelGGL(k_tile_flags, grid_for(c.n), dim3(kBlock), 0, st, c.n, (uint32_t)d.tile_shape, sorder(d),
                           sbins(d), d.tflag);
        e = scan_exclusive_u32(d.tflag, d.tidx, c.n, d.sort.partial, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_tile_list, grid_for(c.n), dim3(kBlock), 0, st, c.n, d.tflag, d.tidx, d.tiles, d.ntiles);
    }
    return hipGetLastError();
}

hipError_t launch_sdf_block_range(const float* grid, const int32_t dims[3], float2* out, hipStream_t st) {
    const int32_t bx = sdf_blocks(dims[0]), by = sdf_blocks(dims[1]), bz = sdf_blocks(dims[2]);
    const uint32_t nb = (uint32_t)bx * (uint32_t)by * (uint32_t)bz;
    hipLaunchKernelGGL(k_sdf_block_range, grid_for(nb), dim3(kBlock), 0, st, grid, dims[0], dims[1], dims[2], bx, by, bz,
                       out);
    return hipGetLastError();
}

hipError_t launch_lod(const StepConst& c, DeviceState& d, hipStream_t st) {
    hipLaunchKernelGGL(k_lod, grid_for(c.n), dim3(kBlock), 0, st, c, d.P, d.lod);
    return hipGetLastError();
}

// The list of the list-form gathers over the positions X (quantised in d.Q).
template <bool B, int BS>
static void build_list_bs(const StepConst& c, DeviceState& d, const float4* X, const uint32_t* lod, uint32_t iter,
                          bool taps, hipStream_t st) {
    if (taps)
        hipLaunchKernelGGL((k_build_list<true, B, BS>), grid_bs<BS>(c.n), dim3(BS), 0, st, c, sorder(d), X, d.Q, lod,
                           iter, gcells(d), d.list_count, d.list, d.cand);
    else
        hipLaunchKernelGGL((k_build_list<false, B, BS>), grid_bs<BS>(c.n), dim3(BS), 0, st, c, sorder(d), X, d.Q, lod,
                           iter, gcells(d), d.list_count, d.list, d.cand);
}
template <bool B>
static void build_list_t(const StepConst& c, DeviceState& d, const float4* X, const uint32_t* lod, uint32_t iter,
                         bool taps, hipStream_t st) {
    if (d.block == 128) build_list_bs<B, 128>(c, d, X, lod, iter, taps, st);
    else build_list_bs<B, 256>(c, d, X, lod, iter, taps, st);
}

// gather form V (tuning knob) x batched x workgroup size (tuning knob)
#define FSI_GATHER_SWITCH(CALL)                                                         \
    if (d.variant >= 1) {                                                               \
        if (d.batched) {                                                                \
            if (d.block == 128) { CALL(1, true, 128); } else { CALL(1, true, 256); }    \
        } else {                                                                        \
            if (d.block == 128) { CALL(1, false, 128); } else { CALL(1, false, 256); }  \
        }                                                                               \
    } else {                                                                            \
        if (d.batched) {                                                                \
            if (d.block == 128) { CALL(0, true, 128); } else { CALL(0, true, 256); }    \
        } else {                                                                        \
            if (d.block == 128) { CALL(0, false, 128); } else { CALL(0, false, 256); }  \
        }                                                                               \
    }

hipError_t launch_lambda(const StepConst& c, DeviceState& d, uint32_t iter, bool taps, hipStream_t st) {
    if (tiled_form(d)) {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kBlock), 0, st, d.ndefer, kDeferWords, 0u);
#define LM(M, T, B)                                                                                              \
    hipLaunchKernelGGL((k_lambda_tile<M, T, B>), dim3(kTileGrid), dim3(kBlock), 0, st, c, d.tiles, d.ntiles,     \
                       sorder(d), sbins(d), d.P, d.PL, d.rho, d.lod, gcells(d), iter, d.cand, d.nbr, d.defer, d.ndefer);
#define L(T, B)                                                                                                  \
    if (d.tile_shape == 2) { LM(2, T, B) } else if (d.tile_shape == 1) { LM(1, T, B) } else { LM(0, T, B) }      \
    hipLaunchKernelGGL((k_lambda_deferred<T, B>), dim3(kDeferGrid), dim3(kBlock), 0, st, c, d.defer, d.ndefer, d.P, \
                       d.PL, d.rho, gcells(d), d.cand, d.nbr);
        if (taps) {
            FSI_B_SWITCH(d.batched, L_T)
        } else {
            FSI_B_SWITCH(d.batched, L_F)
        }
#undef L
#undef LM
        return hipGetLastError();
    }
    if (d.variant >= 1) {
#define L(B) build_list_t<B>(c, d, d.P, d.lod, iter, taps, st);
        FSI_B_SWITCH(d.batched, L)
#undef L
    }
#define L(V, B, BS)                                                                                              \
    if (taps)                                                                                                    \
        hipLaunchKernelGGL((k_lambda<V, true, B, BS>), grid_bs<BS>(c.n), dim3(BS), 0, st, c, sorder(d), d.P, d.PL, \
                           d.rho, d.lod, gcells(d), iter, d.list_count, d.list, d.cand, d.nbr);                    \
    else                                                                                                         \
        hipLaunchKernelGGL((k_lambda<V, false, B, BS>), grid_bs<BS>(c.n), dim3(BS), 0, st, c, sorder(d), d.P, d.PL, \
                           d.rho, d.lod, gcells(d), iter, d.list_count, d.list, d.cand, d.nbr);
    FSI_GATHER_SWITCH(L)
#undef L
    return hipGetLastError();
}

hipError_t launch_deltap(const StepConst& c, DeviceState& d, uint32_t iter, bool taps, hipStream_t st) {
    if (tiled_form(d)) {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kBlock), 0, st, d.ndefer, kDeferWords, 0u);
#define LM(M, T, B)                                                                                               \
    hipLaunchKernelGGL((k_deltap_tile<M, T, B>), dim3(kTileGrid), dim3(kBlock), 0, st, c, d.tiles, d.ntiles,      \
                       sorder(d), sbins(d), d.PL, d.P, d.Q, d.lod, gcells(d), iter, d.dp_tap, d.defer, d.ndefer);
#define L(T, B)                                                                                                   \
    if (d.tile_shape == 2) { LM(2, T, B) } else if (d.tile_shape == 1) { LM(1, T, B) } else { LM(0, T, B) }       \
    hipLaunchKernelGGL((k_deltap_deferred<T, B>), dim3(kDeferGrid), dim3(kBlock), 0, st, c, d.defer, d.ndefer, d.PL, \
                       d.P, d.Q, gcells(d), d.dp_tap);
        if (taps) {
            FSI_B_SWITCH(d.batched, L_T)
        } else {
            FSI_B_SWITCH(d.batched, L_F)
        }
#undef L
#undef LM
    } else {
#define L(V, B, BS)                                                                                             \
    if (taps)                                                                                                   \
        hipLaunchKernelGGL((k_deltap<V, true, B, BS>), grid_bs<BS>(c.n), dim3(BS), 0, st, c, sorder(d), d.PL, d.P, \
                           d.Q, d.lod, gcells(d), iter, d.list_count, d.list, d.dp_tap);                          \
    else                                                                                                        \
        hipLaunchKernelGGL((k_deltap<V, false, B, BS>), grid_bs<BS>(c.n), dim3(BS), 0, st, c, sorder(d), d.PL,     \
                           d.P, d.Q, d.lod, gcells(d), iter, d.list_count, d.list, d.dp_tap);
        FSI_GATHER_SWITCH(L)
#undef L
    }
    if (d.any_sdf && tiled_form(d)) {
#define L(B) hipLaunchKernelGGL((k_apply_sdf<B>), grid_for(c.n), dim3(kBlock), 0, st, c, sorder(d), d.P, d.Q, d.lod, iter);
        FSI_B_SWITCH(d.batched, L)
#undef L
    }
    return hipGetLastError();
}

hipError_t launch_integrate(const StepConst& c, DeviceState& d, hipStream_t st) {
    if (c.enable_visc == 1u || c.enable_vort == 1u) {
        if (d.variant >= 1) {
#define L(B) build_list_t<B>(c, d, d.P, nullptr, 0u, false, st);
            FSI_B_SWITCH(d.batched, L)
#undef L
        }
        // the gather finishes its particles itself
#define L(V, B, BS)                                                                                              \
    hipLaunchKernelGGL((k_integrate<V, B, BS, true>), grid_bs<BS>(c.n), dim3(BS), 0, st, c, sorder(d), d.P, d.vs, \
                       gcells(d), d.list_count, d.list, d.vdelta, d.curl, sperm(d), d.pos, d.vel, d.posS, d.Q);
        FSI_GATHER_SWITCH(L)
#undef L
        return hipGetLastError();
    }
#define L(B)                                                                                                   \
    hipLaunchKernelGGL((k_integrate_finish<B>), grid_for(c.n), dim3(kBlock), 0, st, c, sorder(d), d.P, d.vs, d.vdelta, \
                       sperm(d), d.curl, d.pos, d.vel, d.posS, d.Q);
    FSI_B_SWITCH(d.batched, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_vorticity(const StepConst& c, DeviceState& d, hipStream_t st) {
    if (d.variant >= 1) {
#define L(B) build_list_t<B>(c, d, d.posS, nullptr, 0u, false, st);
        FSI_B_SWITCH(d.batched, L)
#undef L
    }
#define L(V, B, BS)                                                                                              \
    hipLaunchKernelGGL((k_vorticity<V, B, BS>), grid_bs<BS>(c.n), dim3(BS), 0, st, c, sorder(d), d.posS, d.curl, \
                       sperm(d), d.vel, gcells(d), d.list_count, d.list);
    FSI_GATHER_SWITCH(L)
#undef L
    return hipGetLastError();
}

hipError_t launch_contacts(const StepConst& c, DeviceState& d, float band, float4* out, uint32_t* count,
                           hipStream_t st) {
#define L(B) hipLaunchKernelGGL((k_contact_flag<B>), grid_for(c.n), dim3(kBlock), 0, st, c, d.pos, band, d.flags);
    FSI_B_SWITCH(d.batched, L)
#undef L
    hipError_t e = scan_exclusive_u32(d.flags, d.slots, c.n, d.sort.partial, st);
    if (e != hipSuccess) return e;
#define L(B)                                                                                                   \
    hipLaunchKernelGGL((k_contact_write<B>), grid_for(c.n), dim3(kBlock), 0, st, c, d.pos, d.vel, d.flags,     \
                       d.slots, out, count);
    FSI_B_SWITCH(d.batched, L)
#undef L
    return hipGetLastError();
}

}  // namespace fsi
