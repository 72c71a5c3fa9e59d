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

#include "pbf_launch.h"

namespace fsi {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint32_t gid() { return blockIdx.x * kBlock + threadIdx.x; }

inline dim3 grid_for(uint32_t n) { return dim3((n + kBlock - 1) / kBlock); }

// Visit the candidate slots of the 27 buckets around cell (cx,cy,cz), in the
// order of CS_ComputeLambdas.glsl:126-150, skipping the particle's own slot.
template <class F>
__device__ __forceinline__ void for_each_candidate(int32_t cx, int32_t cy, int32_t cz, uint32_t mask,
                                                   const int2* __restrict__ cells, uint32_t self, F&& f) {
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const uint32_t b = cell_hash(cx + dx, cy + dy, cz + dz, mask);
                const int2 se = cells[b];
                if (se.x == -1) continue;
                for (int32_t k = se.x; k < se.y; ++k) {
                    if ((uint32_t)k == self) continue;
                    f((uint32_t)k);
                }
            }
}

// CS_PredictAndHash.glsl:88-130
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
    keys[id] = cell_hash(cx, cy, cz, c.hash_mask);
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

// Gather the predicted state into key-sorted order; lambda starts at 0
// (CS_PredictAndHash.glsl:121).
__global__ __launch_bounds__(kBlock) void k_reorder(uint32_t n, const uint32_t* __restrict__ perm,
                                                    const float4* __restrict__ xstar,
                                                    const float4* __restrict__ vel, float4* __restrict__ P,
                                                    float4* __restrict__ PL, float4* __restrict__ vs) {
    const uint32_t k = gid();
    if (k >= n) return;
    const uint32_t id = perm[k];
    const float4 x = xstar[id];
    P[k] = x;
    PL[k] = x;
    vs[k] = vel[id];
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

// CS_ComputeLambdas.glsl:108-162
template <bool TAPS>
__global__ __launch_bounds__(kBlock) void k_lambda(StepConst c, const float4* __restrict__ P,
                                                   float4* __restrict__ PL, float* __restrict__ rho_out,
                                                   const uint32_t* __restrict__ lod, const int2* __restrict__ cells,
                                                   uint32_t iter, uint32_t* __restrict__ cand_out,
                                                   uint32_t* __restrict__ nbr_out) {
    const uint32_t k = gid();
    if (k >= c.n) return;
    const V3 xi = xyz(P[k]);
    if (c.enable_apbf != 0u && lod[k] < iter) {
        // inactive: keep the last lambda, refresh the position
        PL[k] = make_float4(xi.x, xi.y, xi.z, PL[k].w);
        return;
    }
    const float inv = c.inv_rho0;
    const float h2 = c.h * c.h;
    float rho = c.w0_self;
    float sum_g2 = 0.0f;
    V3 gi = mk(0.0f, 0.0f, 0.0f);
    uint32_t ncand = 0, nnbr = 0;
    int32_t cx, cy, cz;
    cell_of(xi, c.h, cx, cy, cz);
    for_each_candidate(cx, cy, cz, c.hash_mask, cells, k, [&](uint32_t j) {
        const V3 r = sub(xi, xyz(P[j]));
        const float r2 = dot(r, r);
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
    gi = scl(gi, inv);
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

// CS_ComputeDeltaP.glsl:120-166 fused with CS_ApplyDeltaP.glsl:80-124: reads
// (x, lambda) from PL, writes the corrected position to P.
template <bool TAPS>
__global__ __launch_bounds__(kBlock) void k_deltap(StepConst c, const float4* __restrict__ PL,
                                                   float4* __restrict__ P, const uint32_t* __restrict__ lod,
                                                   const int2* __restrict__ cells, uint32_t iter,
                                                   float4* __restrict__ dp_tap) {
    const uint32_t k = gid();
    if (k >= c.n) return;
    if (c.enable_apbf != 0u && lod[k] < iter) return;  // P[k] already holds x
    const float4 xi4 = PL[k];
    const V3 xi = xyz(xi4);
    const float li = xi4.w;
    const float h2 = c.h * c.h;
    V3 dp = mk(0.0f, 0.0f, 0.0f);
    int32_t cx, cy, cz;
    cell_of(xi, c.h, cx, cy, cz);
    for_each_candidate(cx, cy, cz, c.hash_mask, cells, k, [&](uint32_t j) {
        const float4 xj4 = PL[j];
        const V3 r = sub(xi, xyz(xj4));
        const float r2 = dot(r, r);
        if (r2 <= h2) {
            const V3 gw = spiky(c, r, r2);
            const float sc = scorr_r2(c, r2);
            dp = add(dp, scl(gw, li + xj4.w + sc));
        }
    });
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
    if (c.enable_sdf == 1u) {
        V3 nrm;
        (void)sdf_project(c, p, nrm);
    }
    P[k] = make_float4(p.x, p.y, p.z, 0.0f);
}

// CS_Integrate.glsl:135-280.  Neighbour velocities come from the sorted
// post-predict snapshot vs (the Jacobi form of the racy read at :171).
__global__ __launch_bounds__(kBlock) void k_integrate(StepConst c, const float4* __restrict__ P,
                                                      const float4* __restrict__ vs,
                                                      const uint32_t* __restrict__ perm,
                                                      float4* __restrict__ pos, float4* __restrict__ vel,
                                                      const int2* __restrict__ cells, float4* __restrict__ curl_out,
                                                      float4* __restrict__ posS) {
    const uint32_t k = gid();
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
    V3 visc_delta = mk(0.0f, 0.0f, 0.0f);
    if (do_visc || do_vort) {
        const float h2 = c.h * c.h;
        V3 visc = mk(0.0f, 0.0f, 0.0f);
        V3 curl = mk(0.0f, 0.0f, 0.0f);
        int32_t cx, cy, cz;
        cell_of(xp, c.h, cx, cy, cz);
        for_each_candidate(cx, cy, cz, c.hash_mask, cells, k, [&](uint32_t j) {
            const V3 r = sub(xp, xyz(P[j]));
            const float r2 = dot(r, r);
            if (r2 <= h2) {
                const V3 vij = sub(xyz(vs[j]), old_vel);
                if (do_visc) {
                    const float w = poly6_r2(c, r2);
                    visc = add(visc, scl(scl(vij, w), inv));
                }
                if (do_vort) {
                    const V3 gw = spiky(c, r, r2);
                    curl = add(curl, cross(vij, gw));
                }
            }
        });
        if (do_visc) visc_delta = scl(visc, c.viscosity);
        if (do_vort) {
            const V3 cw = scl(curl, inv);
            curl_out[k] = make_float4(cw.x, cw.y, cw.z, 0.0f);
        }
    }
    bool cxh = false, cyh = false, czh = false;
    if (xp.x <= c.bmin[0]) { xp.x = c.bmin[0]; cxh = true; }
    if (xp.x >= c.bmax[0]) { xp.x = c.bmax[0]; cxh = true; }
    if (xp.y <= c.bmin[1]) { xp.y = c.bmin[1]; cyh = true; }
    if (xp.y >= c.bmax[1]) { xp.y = c.bmax[1]; cyh = true; }
    if (xp.z <= c.bmin[2]) { xp.z = c.bmin[2]; czh = true; }
    if (xp.z >= c.bmax[2]) { xp.z = c.bmax[2]; czh = true; }
    bool hit = false;
    V3 nrm = mk(0.0f, 0.0f, 0.0f);
    if (c.enable_sdf == 1u) hit = sdf_project(c, xp, nrm);
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
    vel[id] = make_float4(v.x, v.y, v.z, ov4.w);
    posS[k] = make_float4(xp.x, xp.y, xp.z, 0.0f);
}

// CS_VorticityConfinement.glsl:133-167 (pass 1)
__global__ __launch_bounds__(kBlock) void k_vorticity(StepConst c, const float4* __restrict__ posS,
                                                      const float4* __restrict__ curl,
                                                      const uint32_t* __restrict__ perm,
                                                      float4* __restrict__ vel, const int2* __restrict__ cells) {
    const uint32_t k = gid();
    if (k >= c.n) return;
    const V3 pi = xyz(posS[k]);
    const float h2 = c.h * c.h;
    int32_t cx, cy, cz;
    cell_of(pi, c.h, cx, cy, cz);
    V3 eta = mk(0.0f, 0.0f, 0.0f);
    for_each_candidate(cx, cy, cz, c.hash_mask, cells, k, [&](uint32_t j) {
        const V3 r = sub(pi, xyz(posS[j]));
        const float r2 = dot(r, r);
        if (r2 <= h2) {
            const V3 gw = spiky(c, r, r2);
            const float om = len(xyz(curl[j]));
            eta = add(eta, scl(gw, om));
        }
    });
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
__global__ __launch_bounds__(kBlock) void k_contact_flag(StepConst c, const float4* __restrict__ pos, float band,
                                                         uint32_t* __restrict__ flags) {
    const uint32_t id = gid();
    if (id >= c.n) return;
    const float d = sdf_query(c, xyz(pos[id]));
    const float ad = d < 0.0f ? -d : d;
    flags[id] = ad < band ? 1u : 0u;
}

__global__ __launch_bounds__(kBlock) void k_contact_write(StepConst c, const float4* __restrict__ pos,
                                                          const float4* __restrict__ vel,
                                                          const uint32_t* __restrict__ slots, float band,
                                                          float4* __restrict__ out, uint32_t* __restrict__ count) {
    const uint32_t id = gid();
    if (id >= c.n) return;
    const float4 p4 = pos[id];
    const float d = sdf_query(c, xyz(p4));
    const float ad = d < 0.0f ? -d : d;
    if (ad < band) {
        const uint32_t s = slots[id];
        out[2 * (size_t)s] = make_float4(p4.x, p4.y, p4.z, 0.0f);
        out[2 * (size_t)s + 1] = vel[id];
    }
    if (id == c.n - 1) count[0] = slots[id] + (ad < band ? 1u : 0u);
}

}  // namespace

hipError_t launch_predict(const StepConst& c, DeviceState& d, hipStream_t st) {
    hipLaunchKernelGGL(k_predict, grid_for(c.n), dim3(kBlock), 0, st, c, d.pos, d.vel, d.xstar, d.keys);
    return hipGetLastError();
}

hipError_t launch_cells(const StepConst& c, DeviceState& d, hipStream_t st) {
    hipLaunchKernelGGL(k_clear_cells, grid_for(d.hash_size), dim3(kBlock), 0, st, d.cells, d.hash_size);
    hipLaunchKernelGGL(k_build_cells, grid_for(c.n), dim3(kBlock), 0, st, d.skeys, c.n, d.cells);
    return hipGetLastError();
}

hipError_t launch_reorder(const StepConst& c, DeviceState& d, hipStream_t st) {
    hipLaunchKernelGGL(k_reorder, grid_for(c.n), dim3(kBlock), 0, st, c.n, d.perm, d.xstar, d.vel, d.P, d.PL, d.vs);
    return hipGetLastError();
}

hipError_t launch_lod(const StepConst& c, DeviceState& d, hipStream_t st) {
    hipLaunchKernelGGL(k_lod, grid_for(c.n), dim3(kBlock), 0, st, c, d.P, d.lod);
    return hipGetLastError();
}

hipError_t launch_lambda(const StepConst& c, DeviceState& d, uint32_t iter, bool taps, hipStream_t st) {
    if (taps)
        hipLaunchKernelGGL(k_lambda<true>, grid_for(c.n), dim3(kBlock), 0, st, c, d.P, d.PL, d.rho, d.lod, d.cells,
                           iter, d.cand, d.nbr);
    else
        hipLaunchKernelGGL(k_lambda<false>, grid_for(c.n), dim3(kBlock), 0, st, c, d.P, d.PL, d.rho, d.lod, d.cells,
                           iter, d.cand, d.nbr);
    return hipGetLastError();
}

hipError_t launch_deltap(const StepConst& c, DeviceState& d, uint32_t iter, bool taps, hipStream_t st) {
    if (taps)
        hipLaunchKernelGGL(k_deltap<true>, grid_for(c.n), dim3(kBlock), 0, st, c, d.PL, d.P, d.lod, d.cells, iter,
                           d.dp_tap);
    else
        hipLaunchKernelGGL(k_deltap<false>, grid_for(c.n), dim3(kBlock), 0, st, c, d.PL, d.P, d.lod, d.cells, iter,
                           d.dp_tap);
    return hipGetLastError();
}

hipError_t launch_integrate(const StepConst& c, DeviceState& d, hipStream_t st) {
    hipLaunchKernelGGL(k_integrate, grid_for(c.n), dim3(kBlock), 0, st, c, d.P, d.vs, d.perm, d.pos, d.vel, d.cells,
                       d.curl, d.posS);
    return hipGetLastError();
}

hipError_t launch_vorticity(const StepConst& c, DeviceState& d, hipStream_t st) {
    hipLaunchKernelGGL(k_vorticity, grid_for(c.n), dim3(kBlock), 0, st, c, d.posS, d.curl, d.perm, d.vel, d.cells);
    return hipGetLastError();
}

hipError_t launch_contacts(const StepConst& c, DeviceState& d, float band, float4* out, uint32_t* count,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_contact_flag, grid_for(c.n), dim3(kBlock), 0, st, c, d.pos, band, d.flags);
    hipError_t e = scan_exclusive_u32(d.flags, d.flags, c.n, d.sort.partial, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_contact_write, grid_for(c.n), dim3(kBlock), 0, st, c, d.pos, d.vel, d.flags, band, out,
                       count);
    return hipGetLastError();
}

}  // namespace fsi
