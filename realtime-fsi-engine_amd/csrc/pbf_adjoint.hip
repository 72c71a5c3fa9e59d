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
// them into slot j through the transposed list (the pairs that reach j, in
// ascending pair order), so every sum has a fixed order and the gradient is
// reproducible bit for bit.
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

__device__ __forceinline__ V3 zero3() { return mk(0.0f, 0.0f, 0.0f); }
__device__ __forceinline__ float4 f4(V3 v, float w = 0.0f) { return make_float4(v.x, v.y, v.z, w); }

__global__ __launch_bounds__(kAB) void k_zero_u32(uint32_t* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * kAB + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ---- neighbour lists ------------------------------------------------------
// Pair (i, r) -- the r-th neighbour of slot i -- is stored at r * n + i; its
// number in the transposed list is that same index.

__global__ __launch_bounds__(kAB) void k_edges(StepConst c, const float4* __restrict__ X,
                                               const int2* __restrict__ cells, uint32_t* __restrict__ nbr,
                                               uint32_t* __restrict__ cnt, uint32_t* __restrict__ overflow) {
    const uint32_t i = agid();
    if (i >= c.n) return;
    const float h2 = c.h * c.h;
    uint32_t m = 0;
    scan_buckets(xyz(X[i]), c.h, c.dom, cells, X, i, [&](uint32_t k, float4, V3, float r2) {
        if (r2 <= h2) {
            if (m < kEdgeCap) nbr[(size_t)m * c.n + i] = k;
            ++m;
        }
    });
    atomicMax(overflow + 2, m);
    if (m > kEdgeCap) {
        atomicOr(overflow, 1u);
        m = kEdgeCap;
    }
    cnt[i] = m;
}

__global__ __launch_bounds__(kAB) void k_transpose(uint32_t n, const uint32_t* __restrict__ nbr,
                                                   const uint32_t* __restrict__ cnt, uint32_t* __restrict__ tin,
                                                   uint32_t* __restrict__ tcnt, uint32_t* __restrict__ overflow) {
    const uint32_t i = agid();
    if (i >= n) return;
    const uint32_t m = cnt[i];
    for (uint32_t r = 0; r < m; ++r) {
        const size_t e = (size_t)r * n + i;
        const uint32_t j = nbr[e];
        const uint32_t q = atomicAdd(&tcnt[j], 1u);
        if (q < kInCap) tin[(size_t)q * n + j] = (uint32_t)e;
        else atomicOr(overflow, 2u);
    }
}

// Ascending pair order inside every transposed row (insertion sort).
__global__ __launch_bounds__(kAB) void k_tsort(uint32_t n, uint32_t* __restrict__ tin, uint32_t* __restrict__ tcnt,
                                               uint32_t* __restrict__ stats) {
    const uint32_t j = agid();
    if (j >= n) return;
    atomicMax(stats + 3, tcnt[j]);
    const uint32_t m = min(tcnt[j], kInCap);
    tcnt[j] = m;
    for (uint32_t a = 1; a < m; ++a) {
        const uint32_t v = tin[(size_t)a * n + j];
        uint32_t b = a;
        while (b > 0) {
            const uint32_t u = tin[(size_t)(b - 1) * n + j];
            if (u <= v) break;
            tin[(size_t)b * n + j] = u;
            --b;
        }
        tin[(size_t)b * n + j] = v;
    }
}

// ---- recording forward ----------------------------------------------------

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
    const uint32_t m = cnt[i];
    for (uint32_t r = 0; r < m; ++r) {
        const uint32_t j = nbr[(size_t)r * c.n + i];
        const V3 d = sub(xi, xyz(P[j]));
        const float r2 = dot(d, d);
        rho = rho + poly6_r2(c, r2);
        const V3 gw = spiky(c, d, r2);
        const float gj2 = dot(gw, gw);
        sum_g2 = sum_g2 + gj2 * (inv * inv);
        gi = add(gi, gw);
    }
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
    const uint32_t m = cnt[i];
    for (uint32_t r = 0; r < m; ++r) {
        const float4 a = PL[nbr[(size_t)r * c.n + i]];
        const V3 d = sub(xi, xyz(a));
        const float r2 = dot(d, d);
        const V3 gw = spiky(c, d, r2);
        const float sc = scorr_r2(c, r2);
        dp = add(dp, scl(gw, li + a.w + sc));
    }
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

// ---- reverse passes -------------------------------------------------------
// The reverse passes compute in FP64.  A pair (i, j) contributes to the
// position adjoint of both ends with nearly opposite terms (one computed by
// slot i, one by slot j); in FP32 the rounding of each large term would
// survive their cancellation.  The tape (the forward's FP32 values) is read
// as is, so the derivative is that of the FP32 trajectory.

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 dmk(double x, double y, double z) { return D3{x, y, z}; }
__device__ __forceinline__ D3 dz3() { return D3{0.0, 0.0, 0.0}; }
__device__ __forceinline__ D3 d3(float4 v) { return D3{(double)v.x, (double)v.y, (double)v.z}; }
__device__ __forceinline__ D3 d3(double4 v) { return D3{v.x, v.y, v.z}; }
__device__ __forceinline__ double4 d4(D3 v) { return make_double4(v.x, v.y, v.z, 0.0); }
__device__ __forceinline__ D3 add(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 sub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 scl(D3 a, double s) { return D3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double len(D3 a) { return sqrt(dot(a, a)); }
__device__ __forceinline__ D3 cross(D3 a, D3 b) {
    return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// The pair functions of the forward in FP64 (the support test on r^2 <= h^2
// of the FP32 forward decided which pairs are in the lists).
__device__ __forceinline__ double d_h2(const StepConst& c) { return (double)c.h * (double)c.h; }
__device__ __forceinline__ double d_poly6(const StepConst& c, double r2) {
    const double h2 = d_h2(c);
    if (r2 > h2) return 0.0;
    const double q = h2 - r2;
    return (double)c.poly6 * q * q * q;
}
__device__ __forceinline__ D3 d_poly6_grad(const StepConst& c, D3 r, double r2) {
    const double h2 = d_h2(c);
    if (r2 > h2) return dz3();
    const double q = h2 - r2;
    return scl(r, -6.0 * (double)c.poly6 * q * q);
}
__device__ __forceinline__ D3 d_spiky(const StepConst& c, D3 r, double r2) {
    if (r2 > d_h2(c) || r2 < 1e-10) return dz3();
    const double d = sqrt(r2);
    const double hd = (double)c.h - d;
    return scl(r, -(double)c.spiky * hd * hd / d);
}
// J(r) v for the spiky gradient field g(r) = phi(|r|) r of
// CS_ComputeLambdas.glsl:83-92, phi(d) = -K (h - d)^2 / d:
// J = phi I + phi'(d) / d r r^T (symmetric, even in r), zero where g is zero.
__device__ __forceinline__ D3 d_spiky_jvp(const StepConst& c, D3 r, double r2, D3 v) {
    if (r2 > d_h2(c) || r2 < 1e-10) return dz3();
    const double d = sqrt(r2);
    const double hd = (double)c.h - d;
    const double K = (double)c.spiky;
    const double phi = -K * hd * hd / d;
    const double dphi = K * (2.0 * hd / d + hd * hd / (d * d));
    return add(scl(v, phi), scl(r, dphi / d * dot(r, v)));
}
__device__ __forceinline__ double d_powr(double x, double n) {
    if (n >= 0.0 && n <= 64.0 && n == floor(n)) {
        uint32_t e = (uint32_t)n;
        double result = 1.0, base = x;
        while (e) {
            if (e & 1u) result *= base;
            e >>= 1u;
            if (e) base *= base;
        }
        return result;
    }
    return pow(x, n);
}
// s_corr (CS_ComputeDeltaP.glsl:94-104) and its gradient
// -k n (W/wq)^(n-1) / wq dW/dr.
__device__ __forceinline__ double d_scorr(const StepConst& c, double r2) {
    if (c.enable_scorr == 0u) return 0.0;
    const double w = d_poly6(c, r2);
    if (c.wq <= 1e-12f || w <= 0.0) return 0.0;
    return -(double)c.k_corr * d_powr(w / (double)c.wq, (double)c.n_corr);
}
__device__ __forceinline__ D3 d_scorr_grad(const StepConst& c, D3 r, double r2) {
    if (c.enable_scorr == 0u) return dz3();
    const double w = d_poly6(c, r2);
    if (c.wq <= 1e-12f || w <= 0.0 || c.n_corr == 0.0f) return dz3();
    const double p = d_powr(w / (double)c.wq, (double)c.n_corr - 1.0);
    return scl(d_poly6_grad(c, r, r2), -(double)c.k_corr * (double)c.n_corr * p / (double)c.wq);
}

__global__ __launch_bounds__(kAB) void k_f2d(uint32_t n, const float4* __restrict__ in, double4* __restrict__ out) {
    const uint32_t i = agid();
    if (i >= n) return;
    const float4 v = in[i];
    out[i] = make_double4(v.x, v.y, v.z, 0.0);
}

__global__ __launch_bounds__(kAB) void k_d2f(uint32_t n, const double4* __restrict__ in, float4* __restrict__ out) {
    const uint32_t i = agid();
    if (i >= n) return;
    const double4 v = in[i];
    out[i] = make_float4((float)v.x, (float)v.y, (float)v.z, 0.0f);
}

__global__ __launch_bounds__(kAB) void k_zero_f64(double* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * kAB + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Vorticity confinement, slot i's own terms: v_i += a (n x w_i), a = eps_v dt,
// n = eta / |eta|, eta = inv sum_j |w_j| g(p_i - p_j).
__global__ __launch_bounds__(kAB) void k_rev_vort(StepConst c, const float4* __restrict__ posS,
                                                  const float4* __restrict__ curl, const float4* __restrict__ eta_tape,
                                                  const uint32_t* __restrict__ perm, const double4* __restrict__ gvel,
                                                  const uint32_t* __restrict__ nbr, const uint32_t* __restrict__ cnt,
                                                  double4* __restrict__ gS_own, double4* __restrict__ gw_own,
                                                  double4* __restrict__ ep, double* __restrict__ es) {
    const uint32_t i = agid();
    if (i >= c.n) return;
    const uint32_t m = cnt[i];
    const float4 e4 = eta_tape[i];
    if (len(xyz(e4)) < 1e-6f) {  // the forward's test, in FP32
        for (uint32_t r = 0; r < m; ++r) {
            ep[(size_t)r * c.n + i] = make_double4(0.0, 0.0, 0.0, 0.0);
            es[(size_t)r * c.n + i] = 0.0;
        }
        gS_own[i] = make_double4(0.0, 0.0, 0.0, 0.0);
        gw_own[i] = make_double4(0.0, 0.0, 0.0, 0.0);
        return;
    }
    const D3 eta = d3(e4);
    const double el = len(eta);
    const double inv = (double)c.inv_rho0;
    const double a = (double)c.vort_eps * (double)c.sub_dt;
    const D3 nn = scl(eta, 1.0 / el);
    const D3 om = d3(curl[i]);
    const D3 vb = d3(gvel[perm[i]]);
    const D3 nb = scl(cross(om, vb), a);
    gw_own[i] = d4(scl(cross(vb, nn), a));
    const D3 etab = scl(sub(nb, scl(nn, dot(nn, nb))), 1.0 / el);
    const D3 pi = d3(posS[i]);
    D3 gs = dz3();
    for (uint32_t r = 0; r < m; ++r) {
        const size_t e = (size_t)r * c.n + i;
        const uint32_t j = nbr[e];
        const D3 d = sub(pi, d3(posS[j]));
        const double r2 = dot(d, d);
        const double mj = len(d3(curl[j]));
        const D3 rb = scl(d_spiky_jvp(c, d, r2, etab), inv * mj);
        gs = add(gs, rb);
        ep[e] = d4(rb);
        es[e] = inv * dot(d_spiky(c, d, r2), etab);
    }
    gS_own[i] = d4(gs);
}

// Vorticity, the terms of the pairs reaching j: position adjoint into the
// id-order state adjoint, |w_j| adjoint into w_j.
__global__ __launch_bounds__(kAB) void k_rev_vort_t(uint32_t n, const float4* __restrict__ curl,
                                                    const uint32_t* __restrict__ perm, const uint32_t* __restrict__ tin,
                                                    const uint32_t* __restrict__ tcnt, const double4* __restrict__ ep,
                                                    const double* __restrict__ es, const double4* __restrict__ gS_own,
                                                    const double4* __restrict__ gw_own, double4* __restrict__ gpos,
                                                    double4* __restrict__ gw) {
    const uint32_t j = agid();
    if (j >= n) return;
    D3 sp = dz3();
    double sm = 0.0;
    const uint32_t m = tcnt[j];
    for (uint32_t q = 0; q < m; ++q) {
        const uint32_t e = tin[(size_t)q * n + j];
        sp = add(sp, d3(ep[e]));
        sm += es[e];
    }
    const D3 g = sub(d3(gS_own[j]), sp);
    const uint32_t id = perm[j];
    gpos[id] = d4(add(d3(gpos[id]), g));
    const D3 wj = d3(curl[j]);
    const double mj = len(wj);
    D3 w = d3(gw_own[j]);
    if (mj > 0.0) w = add(w, scl(wj, sm / mj));
    gw[j] = d4(w);
}

// Integrate, slot k's own terms.  gpos / gvel: adjoint of the substep's output
// state (id order); gw: adjoint of the curl (null when no later pass reads it).
__global__ __launch_bounds__(kAB) void k_rev_integrate(StepConst c, const float4* __restrict__ xold,
                                                       const float4* __restrict__ vs, const float4* __restrict__ P,
                                                       const float4* __restrict__ visc_tape,
                                                       const uint32_t* __restrict__ perm,
                                                       const uint32_t* __restrict__ nbr,
                                                       const uint32_t* __restrict__ cnt,
                                                       const double4* __restrict__ gpos,
                                                       const double4* __restrict__ gvel,
                                                       const double4* __restrict__ gw, double4* __restrict__ gP_own,
                                                       double4* __restrict__ gvs_own, double4* __restrict__ gxold,
                                                       double* __restrict__ acc_visc, double4* __restrict__ ep,
                                                       double4* __restrict__ ev) {
    const uint32_t k = agid();
    if (k >= c.n) return;
    const uint32_t id = perm[k];
    // the forward's FP32 decisions: wall flags and the speed clamp
    V3 xp = xyz(P[k]);
    const Domain& dom = c.dom;
    bool cxh = false, cyh = false, czh = false;
    if (xp.x <= dom.bmin[0]) { xp.x = dom.bmin[0]; cxh = true; }
    if (xp.x >= dom.bmax[0]) { xp.x = dom.bmax[0]; cxh = true; }
    if (xp.y <= dom.bmin[1]) { xp.y = dom.bmin[1]; cyh = true; }
    if (xp.y >= dom.bmax[1]) { xp.y = dom.bmax[1]; cyh = true; }
    if (xp.z <= dom.bmin[2]) { xp.z = dom.bmin[2]; czh = true; }
    if (xp.z >= dom.bmax[2]) { xp.z = dom.bmax[2]; czh = true; }
    const bool do_visc = c.enable_visc == 1u;
    const bool do_vort = c.enable_vort == 1u;
    const float inv_dt_f = (c.sub_dt > 0.0f) ? (1.0f / c.sub_dt) : 0.0f;
    const V3 visc_f = xyz(visc_tape[k]);
    V3 vf = scl(sub(xp, xyz(xold[id])), inv_dt_f);
    if (do_visc) vf = add(vf, scl(visc_f, c.viscosity));
    const float sp2 = dot(vf, vf);
    const double inv_dt = (double)inv_dt_f;
    D3 gv = d3(gvel[id]);
    const double damp = (double)c.damping;
    if (cxh) gv.x = gv.x * -damp;
    if (cyh) gv.y = gv.y * -damp;
    if (czh) gv.z = gv.z * -damp;
    if (sp2 > 2500.0f) {
        const D3 v = dmk(vf.x, vf.y, vf.z);
        const double l = len(v);
        const D3 vh = scl(v, 1.0 / l);
        gv = scl(sub(gv, scl(vh, dot(vh, gv))), 50.0 / l);
    }
    const D3 visc = dmk(visc_f.x, visc_f.y, visc_f.z);
    D3 vb = dz3();
    if (do_visc) {
        acc_visc[id] += dot(gv, visc);
        vb = scl(gv, (double)c.viscosity);
    }
    const D3 gxp = add(d3(gpos[id]), scl(gv, inv_dt));
    gxold[id] = d4(scl(gv, -inv_dt));
    D3 gp = dmk(cxh ? 0.0 : gxp.x, cyh ? 0.0 : gxp.y, czh ? 0.0 : gxp.z);
    const D3 wb = (do_vort && gw) ? d3(gw[k]) : dz3();
    D3 gvi = dz3();
    const double inv = (double)c.inv_rho0;
    const uint32_t m = cnt[k];
    if (do_visc || do_vort) {
        const D3 p = d3(P[k]);
        const D3 vi = d3(vs[k]);
        for (uint32_t r = 0; r < m; ++r) {
            const size_t e = (size_t)r * c.n + k;
            const uint32_t j = nbr[e];
            const D3 d = sub(p, d3(P[j]));
            const double r2 = dot(d, d);
            const D3 dv = sub(d3(vs[j]), vi);
            D3 dvb = dz3();
            D3 rb = dz3();
            if (do_visc) {
                dvb = scl(vb, inv * d_poly6(c, r2));
                rb = scl(d_poly6_grad(c, d, r2), inv * dot(dv, vb));
            }
            if (do_vort) {
                dvb = add(dvb, scl(cross(d_spiky(c, d, r2), wb), inv));
                rb = add(rb, d_spiky_jvp(c, d, r2, scl(cross(wb, dv), inv)));
            }
            gvi = sub(gvi, dvb);
            gp = add(gp, rb);
            ev[e] = d4(dvb);
            ep[e] = d4(rb);
        }
    } else {
        for (uint32_t r = 0; r < m; ++r) {
            ep[(size_t)r * c.n + k] = make_double4(0.0, 0.0, 0.0, 0.0);
            ev[(size_t)r * c.n + k] = make_double4(0.0, 0.0, 0.0, 0.0);
        }
    }
    gP_own[k] = d4(gp);
    gvs_own[k] = d4(gvi);
}

__global__ __launch_bounds__(kAB) void k_rev_pairs_vec2(uint32_t n, const uint32_t* __restrict__ tin,
                                                        const uint32_t* __restrict__ tcnt,
                                                        const double4* __restrict__ ep, const double4* __restrict__ ev,
                                                        const double4* __restrict__ gP_own,
                                                        const double4* __restrict__ gvs_own, double4* __restrict__ gP,
                                                        double4* __restrict__ gvs) {
    const uint32_t j = agid();
    if (j >= n) return;
    D3 sp = dz3(), sv = dz3();
    const uint32_t m = tcnt[j];
    for (uint32_t q = 0; q < m; ++q) {
        const uint32_t e = tin[(size_t)q * n + j];
        sp = add(sp, d3(ep[e]));
        sv = add(sv, d3(ev[e]));
    }
    gP[j] = d4(sub(d3(gP_own[j]), sp));
    gvs[j] = d4(add(d3(gvs_own[j]), sv));
}

// Clamp-and-add of CS_ApplyDeltaP.glsl:80-106 and delta-p
// (CS_ComputeDeltaP.glsl:120-166) of one iteration, slot i's own terms.
// gP: adjoint of the corrected positions; out: adjoint of the positions the
// iteration started from (own terms) and of lambda.
__global__ __launch_bounds__(kAB) void k_rev_deltap(StepConst c, const float4* __restrict__ Pprev,
                                                    const float2* __restrict__ lr, const float4* __restrict__ dptape,
                                                    const uint32_t* __restrict__ nbr, const uint32_t* __restrict__ cnt,
                                                    const double4* __restrict__ gP, double4* __restrict__ gQ_own,
                                                    double* __restrict__ gl_own, double4* __restrict__ ep,
                                                    double* __restrict__ es) {
    const uint32_t i = agid();
    if (i >= c.n) return;
    const D3 g = d3(gP[i]);
    const V3 dpf = xyz(dptape[i]);
    const float maxdp = 0.5f * c.h;
    D3 gdp = g;
    if (dot(dpf, dpf) > maxdp * maxdp) {  // the forward's clamp decision
        const D3 dp = d3(dptape[i]);
        const double l = len(dp);
        const D3 dh = scl(dp, 1.0 / l);
        gdp = scl(sub(g, scl(dh, dot(dh, g))), (double)maxdp / l);
    }
    const double inv = (double)c.inv_rho0;
    const D3 xi = d3(Pprev[i]);
    const double li = (double)lr[i].x;
    D3 gq = g;
    double gl = 0.0;
    const uint32_t m = cnt[i];
    for (uint32_t r = 0; r < m; ++r) {
        const size_t e = (size_t)r * c.n + i;
        const uint32_t j = nbr[e];
        const D3 d = sub(xi, d3(Pprev[j]));
        const double r2 = dot(d, d);
        const double a = li + (double)lr[j].x + d_scorr(c, r2);
        const double t = inv * dot(d_spiky(c, d, r2), gdp);
        gl += t;
        const D3 rb = add(scl(d_spiky_jvp(c, d, r2, gdp), inv * a), scl(d_scorr_grad(c, d, r2), t));
        gq = add(gq, rb);
        ep[e] = d4(rb);
        es[e] = t;
    }
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
    const uint32_t j = agid();
    if (j >= n) return;
    D3 sp = dz3();
    double ss = 0.0;
    const uint32_t m = tcnt[j];
    for (uint32_t q = 0; q < m; ++q) {
        const uint32_t e = tin[(size_t)q * n + j];
        sp = add(sp, d3(ep[e]));
        ss += es[e];
    }
    gQ[j] = d4(sub(d3(gQ_own[j]), sp));
    gl[j] = gl_own[j] + ss;
}

// Lambda (CS_ComputeLambdas.glsl:108-162), slot i's own terms, added to gQ.
__global__ __launch_bounds__(kAB) void k_rev_lambda(StepConst c, const float4* __restrict__ Pprev,
                                                    const float2* __restrict__ lr, const float4* __restrict__ aux,
                                                    const uint32_t* __restrict__ perm,
                                                    const uint32_t* __restrict__ nbr, const uint32_t* __restrict__ cnt,
                                                    const double* __restrict__ gl, double4* __restrict__ gQ,
                                                    double* __restrict__ acc_eps, double4* __restrict__ ep) {
    const uint32_t i = agid();
    if (i >= c.n) return;
    const double lb = gl[i];
    const float rho_f = lr[i].y;
    const float4 a4 = aux[i];
    const D3 gi = d3(a4);
    const double D = (double)a4.w;
    // the forward's FP32 cohesion-floor decision, its clamped value
    const float craw_f = rho_f / c.rho0 - 1.0f;
    const bool active = craw_f > -c.cohesion;
    const double C = active ? (double)craw_f : -(double)c.cohesion;
    const double Cb = -lb / D;
    const double Db = lb * C / (D * D);
    acc_eps[perm[i]] += Db;
    const double inv = (double)c.inv_rho0;
    const double rhob = active ? Cb / (double)c.rho0 : 0.0;
    const double sb = 2.0 * inv * inv * Db;
    const D3 Gb = scl(gi, 2.0 * Db * inv);
    const D3 xi = d3(Pprev[i]);
    D3 acc = dz3();
    const uint32_t m = cnt[i];
    for (uint32_t r = 0; r < m; ++r) {
        const size_t e = (size_t)r * c.n + i;
        const uint32_t j = nbr[e];
        const D3 d = sub(xi, d3(Pprev[j]));
        const double r2 = dot(d, d);
        const D3 gw = d_spiky(c, d, r2);
        const D3 rb = add(scl(d_poly6_grad(c, d, r2), rhob), d_spiky_jvp(c, d, r2, add(scl(gw, sb), Gb)));
        acc = add(acc, rb);
        ep[e] = d4(rb);
    }
    gQ[i] = d4(add(d3(gQ[i]), acc));
}

__global__ __launch_bounds__(kAB) void k_rev_pairs_sub(uint32_t n, const uint32_t* __restrict__ tin,
                                                       const uint32_t* __restrict__ tcnt,
                                                       const double4* __restrict__ ep, double4* __restrict__ gQ) {
    const uint32_t j = agid();
    if (j >= n) return;
    D3 sp = dz3();
    const uint32_t m = tcnt[j];
    for (uint32_t q = 0; q < m; ++q) sp = add(sp, d3(ep[tin[(size_t)q * n + j]]));
    gQ[j] = d4(sub(d3(gQ[j]), sp));
}

// Predict (CS_PredictAndHash.glsl:88-130 without the interaction force):
// x* = x + v* dt, v* = v + g dt.
__global__ __launch_bounds__(kAB) void k_rev_predict(StepConst c, const uint32_t* __restrict__ perm,
                                                     const double4* __restrict__ gP0, const double4* __restrict__ gvs,
                                                     const double4* __restrict__ gxold, double4* __restrict__ gpos,
                                                     double4* __restrict__ gvel) {
    const uint32_t k = agid();
    if (k >= c.n) return;
    const uint32_t id = perm[k];
    const D3 gx = d3(gP0[k]);
    gpos[id] = d4(add(gx, d3(gxold[id])));
    gvel[id] = d4(add(d3(gvs[k]), scl(gx, (double)c.sub_dt)));
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
// of non-negative floats; a NaN orders above +inf).
__global__ __launch_bounds__(kAB) void k_absmax(uint32_t n, const double4* __restrict__ gpos,
                                                const double4* __restrict__ gvel, uint32_t* __restrict__ out) {
    const uint32_t i = agid();
    if (i >= n) return;
    const double4 p = gpos[i], v = gvel[i];
    const float m = (float)fmax(fmax(fabs(p.x), fabs(p.y)), fabs(p.z));
    const float q = (float)fmax(fmax(fabs(v.x), fabs(v.y)), fabs(v.z));
    uint32_t bp = __float_as_uint(m), bq = __float_as_uint(q);
    if (p.x != p.x || p.y != p.y || p.z != p.z) bp = 0x7fc00000u;
    if (v.x != v.x || v.y != v.y || v.z != v.z) bq = 0x7fc00000u;
    atomicMax(out, bp);
    atomicMax(out + 1, bq);
}

// Bitwise comparison of the replayed state with the forward's checkpoint.
__global__ __launch_bounds__(kAB) void k_compare(uint32_t n, const float4* __restrict__ a,
                                                 const float4* __restrict__ b, uint32_t* __restrict__ count) {
    const uint32_t i = agid();
    if (i >= n) return;
    const float4 x = a[i], y = b[i];
    const bool same = __float_as_uint(x.x) == __float_as_uint(y.x) && __float_as_uint(x.y) == __float_as_uint(y.y) &&
                      __float_as_uint(x.z) == __float_as_uint(y.z) && __float_as_uint(x.w) == __float_as_uint(y.w);
    if (!same) atomicAdd(count, 1u);
}

struct EdgeList {
    uint32_t* nbr = nullptr;   // kEdgeCap x n
    uint32_t* cnt = nullptr;   // n
    uint32_t* tin = nullptr;   // kInCap x n
    uint32_t* tcnt = nullptr;  // n
};

}  // namespace

struct fsi_adjoint {
    Core core;
    fsi_fluid_config cfg{};
    fsi_pbf_constants k{};
    uint32_t n = 0, frames = 0, S = 0, I = 0;
    bool forward_done = false;
    // checkpoints: frames + 1 states (the last is the forward's result)
    float4* ck_pos = nullptr;
    float4* ck_vel = nullptr;
    // tape of one frame
    float4* xold = nullptr;   // S x n, id order
    uint32_t* perm = nullptr; // S x n
    float4* vs = nullptr;     // S x n
    float4* P = nullptr;      // S x (I + 1) x n
    float2* lr = nullptr;     // S x I x n
    float4* aux = nullptr;    // S x I x n
    float4* dp = nullptr;     // S x I x n
    float4* visc = nullptr;   // S x n
    float4* eta = nullptr;    // n
    std::vector<EdgeList> lists;  // S x (I + 1) + 1
    uint32_t* list_mem = nullptr;
    // adjoint work (FP64)
    double4 *gpos = nullptr, *gvel = nullptr, *gxold = nullptr;
    double4 *gP = nullptr, *gQ = nullptr, *gown = nullptr, *gvs = nullptr, *gvs_own = nullptr;
    double4 *gw = nullptr, *gw_own = nullptr;
    double *gl = nullptr, *gl_own = nullptr, *acc_visc = nullptr, *acc_eps = nullptr;
    double4 *ep = nullptr, *ev = nullptr;
    double* es = nullptr;
    // [0] list overflow bits, [1] replayed records that differ from the
    // forward, [2] / [3] largest neighbour count / in-degree of any list
    uint32_t* flags = nullptr;
    uint32_t stats[4] = {0, 0, 0, 0};  // flags of the last backward
    uint32_t* norms = nullptr;          // frames x 2: max |adjoint| of pos, vel at each frame start
    std::vector<uint32_t> host_norms;
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
    dev_free(&a->ck_pos);
    dev_free(&a->ck_vel);
    dev_free(&a->xold);
    dev_free(&a->perm);
    dev_free(&a->vs);
    dev_free(&a->P);
    dev_free(&a->lr);
    dev_free(&a->aux);
    dev_free(&a->dp);
    dev_free(&a->visc);
    dev_free(&a->eta);
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

hipError_t alloc_all(fsi_adjoint* a) {
    const size_t n = a->n, S = a->S, I = a->I;
    hipError_t e = hipSuccess;
#define ALLOC(p, cnt) \
    if ((e = dev_alloc(&(p), (cnt))) != hipSuccess) return e;
    ALLOC(a->ck_pos, (a->frames + 1) * n);
    ALLOC(a->ck_vel, (a->frames + 1) * n);
    ALLOC(a->xold, S * n);
    ALLOC(a->perm, S * n);
    ALLOC(a->vs, S * n);
    ALLOC(a->P, S * (I + 1) * n);
    ALLOC(a->lr, S * I * n);
    ALLOC(a->aux, S * I * n);
    ALLOC(a->dp, S * I * n);
    ALLOC(a->visc, S * n);
    ALLOC(a->eta, n);
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
    ALLOC(a->flags, 4);
    ALLOC(a->norms, 2 * (size_t)a->frames);
    ALLOC(a->dred, 2);
#undef ALLOC
    return hipSuccess;
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

hipError_t build_list(fsi_adjoint* a, const StepConst& c, const EdgeList& L, const float4* X) {
    hipStream_t st = a->core.stream;
    const uint32_t n = a->n;
    hipLaunchKernelGGL(k_edges, agrid(n), dim3(kAB), 0, st, c, X, a->core.d.cells, L.nbr, L.cnt, a->flags);
    hipLaunchKernelGGL(k_zero_u32, agrid(n), dim3(kAB), 0, st, L.tcnt, (size_t)n);
    hipLaunchKernelGGL(k_transpose, agrid(n), dim3(kAB), 0, st, n, L.nbr, L.cnt, L.tin, L.tcnt, a->flags);
    hipLaunchKernelGGL(k_tsort, agrid(n), dim3(kAB), 0, st, n, L.tin, L.tcnt, a->flags);
    return hipGetLastError();
}

const EdgeList& list_of(const fsi_adjoint* a, uint32_t s, uint32_t k) { return a->lists[(size_t)s * (a->I + 1) + k]; }

// Replays one frame from the state in d.pos / d.vel, writing the tape.
int record_frame(fsi_adjoint* a, const StepConst& c) {
    DeviceState& d = a->core.d;
    hipStream_t st = a->core.stream;
    const size_t n = a->n;
    const uint32_t I = a->I;
    const int key_bits = log2u(a->cfg.hash_size);
    for (uint32_t s = 0; s < a->S; ++s) {
        FSI_HIP_TRY(hipMemcpyAsync(a->xold + s * n, d.pos, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
        FSI_HIP_TRY(launch_predict(c, d, st));
        FSI_HIP_TRY(radix_sort_pairs(d.keys, d.n, key_bits, a->core.sort_digit_bits, d.sort, &d.skeys, &d.perm, st));
        FSI_HIP_TRY(launch_cells(c, d, st));
        FSI_HIP_TRY(launch_reorder(c, d, st));
        uint32_t* perm = a->perm + s * n;
        float4* vs = a->vs + s * n;
        float4* Ps = a->P + (size_t)s * (I + 1) * n;
        FSI_HIP_TRY(hipMemcpyAsync(perm, d.perm, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
        FSI_HIP_TRY(hipMemcpyAsync(vs, d.vs, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
        FSI_HIP_TRY(hipMemcpyAsync(Ps, d.P, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
        for (uint32_t it = 0; it < I; ++it) {
            const EdgeList& L = list_of(a, s, it);
            FSI_HIP_TRY(build_list(a, c, L, d.P));
            const size_t t = ((size_t)s * I + it) * n;
            hipLaunchKernelGGL(k_rec_lambda, agrid(a->n), dim3(kAB), 0, st, c, d.P, L.nbr, L.cnt, d.PL, a->lr + t,
                               a->aux + t);
            hipLaunchKernelGGL(k_rec_deltap, agrid(a->n), dim3(kAB), 0, st, c, d.PL, L.nbr, L.cnt, d.P,
                               Ps + (it + 1) * n, a->dp + t);
        }
        const EdgeList& L = list_of(a, s, I);
        FSI_HIP_TRY(build_list(a, c, L, d.P));
        hipLaunchKernelGGL(k_rec_integrate, agrid(a->n), dim3(kAB), 0, st, c, d.P, d.vs, d.perm, d.pos, d.vel, L.nbr,
                           L.cnt, d.curl, d.posS, a->visc + s * n);
    }
    if (c.enable_vort) {
        const EdgeList& L = a->lists.back();
        FSI_HIP_TRY(build_list(a, c, L, d.posS));
        hipLaunchKernelGGL(k_rec_vorticity, agrid(a->n), dim3(kAB), 0, st, c, d.posS, d.curl, d.perm, d.vel, L.nbr,
                           L.cnt, a->eta);
    }
    FSI_HIP_TRY(hipGetLastError());
    return FSI_OK;
}

// Runs the recorded frame backwards: gpos / gvel (id order) hold the adjoint
// of the frame's final state on entry and of its initial state on exit.
int reverse_frame(fsi_adjoint* a, const StepConst& c) {
    DeviceState& d = a->core.d;
    hipStream_t st = a->core.stream;
    const uint32_t n = a->n;
    const size_t ns = n;
    const uint32_t I = a->I, S = a->S;
    const dim3 g = agrid(n), b(kAB);
    bool have_gw = false;
    if (c.enable_vort) {
        const EdgeList& L = a->lists.back();
        const uint32_t* perm = a->perm + (S - 1) * ns;
        hipLaunchKernelGGL(k_rev_vort, g, b, 0, st, c, d.posS, d.curl, a->eta, perm, a->gvel, L.nbr, L.cnt, a->gown,
                           a->gw_own, a->ep, a->es);
        hipLaunchKernelGGL(k_rev_vort_t, g, b, 0, st, n, d.curl, perm, L.tin, L.tcnt, a->ep, a->es, a->gown, a->gw_own,
                           a->gpos, a->gw);
        have_gw = true;
    }
    for (uint32_t s = S; s-- > 0;) {
        const uint32_t* perm = a->perm + s * ns;
        const float4* Ps = a->P + (size_t)s * (I + 1) * ns;
        const EdgeList& Li = list_of(a, s, I);
        hipLaunchKernelGGL(k_rev_integrate, g, b, 0, st, c, a->xold + s * ns, a->vs + s * ns, Ps + I * ns,
                           a->visc + s * ns, perm, Li.nbr, Li.cnt, a->gpos, a->gvel,
                           (have_gw && s == S - 1) ? a->gw : nullptr, a->gown, a->gvs_own, a->gxold, a->acc_visc,
                           a->ep, a->ev);
        hipLaunchKernelGGL(k_rev_pairs_vec2, g, b, 0, st, n, Li.tin, Li.tcnt, a->ep, a->ev, a->gown, a->gvs_own, a->gP,
                           a->gvs);
        for (uint32_t it = I; it-- > 0;) {
            const EdgeList& L = list_of(a, s, it);
            const size_t t = ((size_t)s * I + it) * ns;
            const float4* Pprev = Ps + it * ns;
            hipLaunchKernelGGL(k_rev_deltap, g, b, 0, st, c, Pprev, a->lr + t, a->dp + t, L.nbr, L.cnt, a->gP, a->gown,
                               a->gl_own, a->ep, a->es);
            hipLaunchKernelGGL(k_rev_pairs_vec_scalar, g, b, 0, st, n, L.tin, L.tcnt, a->ep, a->es, a->gown, a->gl_own,
                               a->gQ, a->gl);
            hipLaunchKernelGGL(k_rev_lambda, g, b, 0, st, c, Pprev, a->lr + t, a->aux + t, perm, L.nbr, L.cnt, a->gl,
                               a->gQ, a->acc_eps, a->ep);
            hipLaunchKernelGGL(k_rev_pairs_sub, g, b, 0, st, n, L.tin, L.tcnt, a->ep, a->gQ);
            std::swap(a->gP, a->gQ);
        }
        hipLaunchKernelGGL(k_rev_predict, g, b, 0, st, c, perm, a->gP, a->gvs, a->gxold, a->gpos, a->gvel);
    }
    FSI_HIP_TRY(hipGetLastError());
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
    if (e == hipSuccess) e = alloc_all(a);
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

int fsi_adj_forward(fsi_adjoint* a, const float* pos0, const float* vel0, float viscosity, float eps, float* pos_out,
                    float* vel_out, void* stream) {
    int rc = bind(a);
    if (rc) return rc;
    if (!pos0 || !vel0) return fail(FSI_ERR_INVALID_ARG, "NULL state");
    if (!std::isfinite(viscosity) || !std::isfinite(eps)) return fail(FSI_ERR_INVALID_ARG, "non-finite parameter");
    a->cfg.viscosity = viscosity;
    a->cfg.eps = eps;
    pack_constants(a->cfg, a->cfg.bounds_min, a->cfg.bounds_max, std::clamp(a->cfg.bound_damping, 0.0f, 1.0f), a->n,
                   a->k);
    const StepConst c = step_const(a);
    a->core.d.tiled = c.dom.dshift == kTileShift;
    FramePlan plan;
    plan.substeps = a->S;
    plan.max_iter = a->I;
    plan.key_bits = log2u(a->cfg.hash_size);
    plan.vorticity = a->cfg.enable_vorticity != 0;
    Core& core = a->core;
    DeviceState& d = core.d;
    hipStream_t st = core.stream;
    const size_t n = a->n;
    FSI_HIP_TRY(join_in(a, stream));
    FSI_HIP_TRY(hipMemcpyAsync(d.pos, pos0, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
    FSI_HIP_TRY(hipMemcpyAsync(d.vel, vel0, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
    for (uint32_t f = 0; f < a->frames; ++f) {
        FSI_HIP_TRY(hipMemcpyAsync(a->ck_pos + f * n, d.pos, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
        FSI_HIP_TRY(hipMemcpyAsync(a->ck_vel + f * n, d.vel, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
        if ((rc = core_frame(core, c, plan))) return rc;
    }
    FSI_HIP_TRY(hipMemcpyAsync(a->ck_pos + a->frames * n, d.pos, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
    FSI_HIP_TRY(hipMemcpyAsync(a->ck_vel + a->frames * n, d.vel, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
    if (pos_out) FSI_HIP_TRY(hipMemcpyAsync(pos_out, d.pos, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
    if (vel_out) FSI_HIP_TRY(hipMemcpyAsync(vel_out, d.vel, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
    FSI_HIP_TRY(join_out(a, stream));
    a->forward_done = true;
    return FSI_OK;
}

int fsi_adj_backward(fsi_adjoint* a, const float* grad_pos, const float* grad_vel, float* grad_pos0,
                     float* grad_vel0, double* grad_params, void* stream) {
    int rc = bind(a);
    if (rc) return rc;
    if (!a->forward_done) return fail(FSI_ERR_STATE, "backward before forward");
    const StepConst c = step_const(a);
    Core& core = a->core;
    DeviceState& d = core.d;
    hipStream_t st = core.stream;
    const size_t n = a->n;
    FSI_HIP_TRY(join_in(a, stream));
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(kAB), 0, st, a->flags, (size_t)4);
    hipLaunchKernelGGL(k_zero_u32, agrid(2 * a->frames), dim3(kAB), 0, st, a->norms, 2 * (size_t)a->frames);
    hipLaunchKernelGGL(k_zero_f64, agrid(a->n), dim3(kAB), 0, st, a->acc_visc, n);
    hipLaunchKernelGGL(k_zero_f64, agrid(a->n), dim3(kAB), 0, st, a->acc_eps, n);
    if (grad_pos)
        hipLaunchKernelGGL(k_f2d, agrid(a->n), dim3(kAB), 0, st, a->n, (const float4*)grad_pos, a->gpos);
    else
        hipLaunchKernelGGL(k_zero_f64, agrid(4 * a->n), dim3(kAB), 0, st, (double*)a->gpos, 4 * n);
    if (grad_vel)
        hipLaunchKernelGGL(k_f2d, agrid(a->n), dim3(kAB), 0, st, a->n, (const float4*)grad_vel, a->gvel);
    else
        hipLaunchKernelGGL(k_zero_f64, agrid(4 * a->n), dim3(kAB), 0, st, (double*)a->gvel, 4 * n);
    for (uint32_t f = a->frames; f-- > 0;) {
        FSI_HIP_TRY(hipMemcpyAsync(d.pos, a->ck_pos + f * n, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
        FSI_HIP_TRY(hipMemcpyAsync(d.vel, a->ck_vel + f * n, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
        if ((rc = record_frame(a, c))) return rc;
        hipLaunchKernelGGL(k_compare, agrid(a->n), dim3(kAB), 0, st, a->n, d.pos, a->ck_pos + (f + 1) * n, a->flags + 1);
        hipLaunchKernelGGL(k_compare, agrid(a->n), dim3(kAB), 0, st, a->n, d.vel, a->ck_vel + (f + 1) * n, a->flags + 1);
        if ((rc = reverse_frame(a, c))) return rc;
        hipLaunchKernelGGL(k_absmax, agrid(a->n), dim3(kAB), 0, st, a->n, a->gpos, a->gvel, a->norms + 2 * (size_t)f);
    }
    hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kAB), 0, st, a->n, a->acc_visc, a->acc_eps, a->dred);
    if (grad_pos0) hipLaunchKernelGGL(k_d2f, agrid(a->n), dim3(kAB), 0, st, a->n, a->gpos, (float4*)grad_pos0);
    if (grad_vel0) hipLaunchKernelGGL(k_d2f, agrid(a->n), dim3(kAB), 0, st, a->n, a->gvel, (float4*)grad_vel0);
    FSI_HIP_TRY(hipGetLastError());
    double red[2] = {0.0, 0.0};
    uint32_t* flags = a->stats;
    FSI_HIP_TRY(hipMemcpyAsync(red, a->dred, sizeof(red), hipMemcpyDeviceToHost, st));
    FSI_HIP_TRY(hipMemcpyAsync(flags, a->flags, sizeof(a->stats), hipMemcpyDeviceToHost, st));
    a->host_norms.resize(2 * (size_t)a->frames);
    FSI_HIP_TRY(hipMemcpyAsync(a->host_norms.data(), a->norms, sizeof(uint32_t) * a->host_norms.size(),
                               hipMemcpyDeviceToHost, st));
    FSI_HIP_TRY(join_out(a, stream));
    FSI_HIP_TRY(hipStreamSynchronize(st));
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
