// pbf_device.h — device-side constants and per-pair arithmetic of the PBF substep.
//
// Every expression keeps the operand order of the reference shaders
// (resources/shaders/compute/CS_*.glsl); the library is compiled with
// -ffp-contract=off and IEEE-rounded division / square root, so each FP32
// operation rounds exactly once, as in the shader source read literally.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsi {

// Per-scene state of a step: the box, the scene's slice of the cell table,
// its Morton bins and its SDF boundary.  The plain solver has one; the
// batched solver one per scene.
struct Domain {
    float bmin[3], bmax[3];  // padded AABB (PBFluids::setBounds)
    uint32_t hash_mask;      // hashSize - 1 (PBFluids.cpp:95)
    uint32_t hash_base;      // first entry of this scene's cell table
    // processing order: cells binned into a Morton grid of 2^axis_bits bins per
    // axis, bin = (cell - dcmin) >> dshift (clamped), numbered from bin_base
    int32_t dcmin[3];
    uint32_t dshift, axis_bits, bin_base;
    // SDF boundary (R32F grid, x fastest, voxel centres at origin + i*cell)
    uint32_t enable_sdf, sdf_mode;
    int32_t dims[3];
    float sdf_origin[3], sdf_cell, sdf_pad;
    const float* sdf_grid;
    // (minimum, maximum) of the grid over each block of kSdfBlock^3 voxels
    // and the voxels one step above it (sdf_bdims blocks per axis; (-inf,
    // +inf) for a block with a non-finite voxel), or nullptr: lets
    // sdf_project leave out the trilinear samples of a particle that no
    // projection can reach
    const float2* sdf_brange;
    int32_t sdf_bdims[3];
};

// Voxels per axis of an SDF block and the block count along an axis of d voxels.
constexpr int32_t kSdfBlockBits = 2, kSdfBlock = 1 << kSdfBlockBits;
__host__ __device__ inline int32_t sdf_blocks(int32_t d) { return (d + kSdfBlock - 1) >> kSdfBlockBits; }

// Constants of one step: the reference's FluidConfigUBO
// (include/fluid/PBFluids.h:21-61) plus the per-dispatch uniforms that
// PBFluids::step sets (src/fluid/PBFluids.cpp:432-504).  Passed by value.
struct StepConst {
    float g[3], sub_dt;
    float h, rho0, eps, wq;
    float k_corr, n_corr, viscosity, damping;
    float cohesion, irad, istr, w0_self;
    float poly6, spiky, inv_rho0, lod_max_dist;
    uint32_t n, enable_scorr, enable_visc;
    uint32_t min_lod, max_lod, enable_apbf, enable_vort;
    uint32_t interacting;
    float ipos[3], cam[3], vort_eps;
    float qscale;               // kQUnitsPerH / h: quantisation of the candidate pre-test
    uint32_t fast_div;          // the operands of the pair divisions lie where div_rn_normal is exact
    Domain dom;                 // the single scene of the plain solver
    const Domain* doms;         // batched: one per scene
    const uint16_t* scene_of;   // batched: scene of particle id i (= of sorted slot i)
};

// The domain of particle / slot i (the batched solver keeps every scene's
// particles in one contiguous id range, and its sorted slots in the same one).
template <bool B>
__device__ __forceinline__ const Domain& domain_of(const StepConst& c, uint32_t i) {
    if constexpr (B) return c.doms[c.scene_of[i]];
    else return c.dom;
}

// Runs f(dom) with the Domain of slot i.  Batched, the lanes of a wave take
// their scenes in turn: each pass serves the lanes whose scene is the first
// remaining lane's, so the Domain is uniform in it and its fields are read
// with scalar loads instead of occupying vector registers (a wave holds one
// scene, or two at a scene boundary).
template <bool B, class F>
__device__ __forceinline__ void with_domain(const StepConst& c, uint32_t i, F&& f) {
    if constexpr (B) {
        const uint32_t s = c.scene_of[i];
        for (;;) {
            const uint32_t s0 = __builtin_amdgcn_readfirstlane(s);
            if (s == s0) {
                f(c.doms[s0]);
                break;
            }
        }
    } else {
        f(c.dom);
    }
}

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 mk(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 scl(V3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 dvs(V3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ V3 xyz(float4 v) { return mk(v.x, v.y, v.z); }

// CS_PredictAndHash.glsl:77-83 — wrapping int32 products, XOR, mask.
__device__ __forceinline__ uint32_t cell_hash(int32_t cx, int32_t cy, int32_t cz, uint32_t mask) {
    const uint32_t n = ((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u);
    return n & mask;
}

// ivec3(floor(x / h)) — a true division (CS_PredictAndHash.glsl:125).
__device__ __forceinline__ void cell_of(V3 x, float h, int32_t& cx, int32_t& cy, int32_t& cz) {
    cx = (int32_t)floorf(x.x / h);
    cy = (int32_t)floorf(x.y / h);
    cz = (int32_t)floorf(x.z / h);
}

// Morton bin of a cell: (cell - dcmin) >> dshift per axis, clamped to
// [0, 2^axis_bits), bits interleaved (axis_bits <= 7).  Used only to order
// the work; results do not depend on it.
__device__ __forceinline__ uint32_t spread7(uint32_t v) {
    v &= 0x7fu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ __forceinline__ uint32_t clamp_bin(int32_t c, int32_t c0, uint32_t shift, uint32_t bits) {
    int32_t d = c - c0;
    d = d < 0 ? 0 : d;
    d >>= shift;
    const int32_t top = (1 << bits) - 1;
    return (uint32_t)(d > top ? top : d);
}
__device__ __forceinline__ uint32_t morton_bin(const Domain& d, int32_t cx, int32_t cy, int32_t cz) {
    const uint32_t b = d.axis_bits;
    return d.bin_base + (spread7(clamp_bin(cx, d.dcmin[0], d.dshift, b)) |
                         (spread7(clamp_bin(cy, d.dcmin[1], d.dshift, b)) << 1) |
                         (spread7(clamp_bin(cz, d.dcmin[2], d.dshift, b)) << 2));
}
constexpr uint32_t kOrderBins = 1u << 21;
// Inverse of spread7: the 7-bit coordinate held in every third bit.
__device__ __forceinline__ uint32_t compact7(uint32_t v) {
    v &= 0x09249249u;
    v = (v | (v >> 2)) & 0x030c30c3u;
    v = (v | (v >> 4)) & 0x0300f00fu;
    v = (v | (v >> 8)) & 0x7fu;
    return v;
}
// Bins of 2^kTileShift cells per axis are the tiles of the tiled gathers.
constexpr uint32_t kTileShift = 2;

// Poly6 (CS_ComputeLambdas.glsl:74-81) given r2.
__device__ __forceinline__ float poly6_r2(const StepConst& c, float r2) {
    const float h2 = c.h * c.h;
    if (r2 > h2) return 0.0f;
    const float diff = h2 - r2;
    return c.poly6 * diff * diff * diff;
}

// Correctly rounded square root of x >= 2^-96 (finite): the hardware
// estimate s (within 1 ulp) is replaced by its neighbour s -/+ 1 ulp when the
// residual x - s' * s, formed exactly by an FMA, shows that the neighbour is
// nearer.  This is the sequence the compiler emits for an IEEE sqrtf less the
// scaling of inputs below 2^-96 and the pass-through of 0 and infinity, so it
// returns the same bits on that range (checked for every float in it by the
// test library of this package).
__device__ __forceinline__ float sqrt_rn_normal(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float lo = __uint_as_float(__float_as_uint(s) - 1u);
    const float hi = __uint_as_float(__float_as_uint(s) + 1u);
    float r = __builtin_fmaf(-lo, s, x) <= 0.0f ? lo : s;
    r = __builtin_fmaf(-hi, s, x) > 0.0f ? hi : r;
    return r;
}

// ---- conservative candidate pre-test of the gathers ----
// Every gather tests the candidates of 27 buckets against the support
// r^2 <= h^2 (CS_ComputeLambdas.glsl:76,86); about 85% fail.  The scan first
// tests them on quantised positions, which costs fewer instructions and half
// the bytes, and hands only the candidates that pass to the exact FP32 test
// (in the reference's visit order, so the sums are unchanged).
//
// Quantisation: t = fl(x * s) with s = fl(1024 / h), q = floor(t) kept modulo
// 2^16 per axis (x | y << 16 in one word, z in the low half of another), so
// the offset of two particles is a 16-bit wrapping difference and its square
// sum a packed dot product (saturating).  While |t| < 2^22 (ulp(t) <= 1/2), q
// is within 1.25 units of t, and within 1.5 units for a neighbour j of such a
// particle (|t_j| < 2^23).  A pair that passes the exact FP32 test has
// r <= h (1 + 1e-6), i.e. |dq| <= 1024.01 + 2.75 sqrt(3) < 1029 units, so the
// pre-test |dq|^2 <= kQPass keeps it.  A particle at |t| >= 2^22 on some axis
// (or not finite) keeps every candidate: the exact test decides alone.
// Offsets that wrap (candidates 64 h away in aliased buckets) can only pass
// extra candidates to the exact test.
constexpr float kQUnitsPerH = 1024.0f;
constexpr int32_t kQPass = 1030 * 1030;

__device__ __forceinline__ uint32_t qaxis(float v, float s) {
    float t = v * s;
    t = fminf(fmaxf(t, -1073741824.0f), 1073741824.0f);  // +-2^30; NaN -> -2^30
    return (uint32_t)(int32_t)floorf(t);
}

__device__ __forceinline__ uint2 quantize(V3 x, float s) {
    return make_uint2((qaxis(x.x, s) & 0xffffu) | (qaxis(x.y, s) << 16), qaxis(x.z, s) & 0xffffu);
}

// The pre-test bound of a particle at x (see above).
__device__ __forceinline__ int32_t qpass_of(V3 x, float s) {
    const float lim = 4194304.0f;  // 2^22
    const bool in = fabsf(x.x * s) < lim && fabsf(x.y * s) < lim && fabsf(x.z * s) < lim;
    return in ? kQPass : 0x7fffffff;
}

typedef short s16x2 __attribute__((ext_vector_type(2)));

// |q_i - q_j|^2 in units^2, saturating at 2^31 - 1.
__device__ __forceinline__ int32_t qdist2(uint2 a, uint2 b) {
    const s16x2 dxy = __builtin_bit_cast(s16x2, a.x) - __builtin_bit_cast(s16x2, b.x);
    const s16x2 dz = __builtin_bit_cast(s16x2, a.y) - __builtin_bit_cast(s16x2, b.y);
    const int32_t s = __builtin_amdgcn_sdot2(dxy, dxy, 0, true);
    return __builtin_amdgcn_sdot2(dz, dz, s, true);
}

// ---- correctly rounded division without the range fix-ups ----
// This build's a / b (IEEE, -fhip-fp32-correctly-rounded-divide-sqrt) is
// v_div_scale of both operands, the reciprocal estimate and its Newton step,
// the quotient and two FMA residual corrections (the last in v_div_fmas),
// then v_div_fixup.  While a and b are normal, their exponents are less than
// 96 apart and the quotient is normal, v_div_scale returns both operands
// unchanged and clears VCC, v_div_fmas is then an FMA and v_div_fixup
// returns its input: div_rn_normal is the same sequence without those three
// steps and gives the same bits (the test library of this package compares
// it with a / b on every numerator of a wide range for many denominators).
// A zero numerator gives +0 where a / b gives a signed zero; the divisions
// below feed sums that start at +0, which a zero of either sign leaves +0.
// StepConst::fast_div holds the host's check that a step's operands stay in
// the range (pair_divisions_in_range); otherwise the kernels divide.
__device__ __forceinline__ float rcp_refined(float b) {
    const float r = __builtin_amdgcn_rcpf(b);
    return __builtin_fmaf(__builtin_fmaf(-b, r, 1.0f), r, r);
}
// a / b given r1 = rcp_refined(b)
__device__ __forceinline__ float div_rn_by(float a, float b, float r1) {
    const float q = a * r1;
    const float q1 = __builtin_fmaf(__builtin_fmaf(-b, q, a), r1, q);
    return __builtin_fmaf(__builtin_fmaf(-b, q1, a), r1, q1);
}
__device__ __forceinline__ float div_rn_normal(float a, float b) { return div_rn_by(a, b, rcp_refined(b)); }

// Spiky gradient (CS_ComputeLambdas.glsl:83-92).  r2 >= 1e-10 here, inside
// the range of sqrt_rn_normal.
__device__ __forceinline__ V3 spiky(const StepConst& c, V3 r, float r2) {
    const float h2 = c.h * c.h;
    if (r2 > h2 || r2 < 1e-10f) return mk(0.0f, 0.0f, 0.0f);
    const float dist = sqrt_rn_normal(r2);
    const float diff = c.h - dist;
    const float num = -c.spiky * diff * diff;
    const float scalar = c.fast_div != 0u ? div_rn_normal(num, dist) : num / dist;
    return scl(r, scalar);
}

// x^n: binary powering for integer n in [0, 64], powf otherwise.  For n = 4
// (the s_corr exponent of FluidConfig and of every workload here) the
// powering's products are x^2 = x x and (x^2)(x^2), then 1 x^4, which is
// x^4 itself: written out, without the loop the exponent drives (a loop of
// three selects per pair, since n is uniform but not a constant).
__device__ __forceinline__ float powr(float x, float n) {
    if (n == 4.0f) {
        const float x2 = x * x;
        return x2 * x2;
    }
    if (n >= 0.0f && n <= 64.0f && n == floorf(n)) {
        uint32_t e = (uint32_t)n;
        float result = 1.0f;
        float base = x;
        while (e) {
            if (e & 1u) result = result * base;
            e >>= 1u;
            if (e) base = base * base;
        }
        return result;
    }
    return powf(x, n);
}

// s_corr (CS_ComputeDeltaP.glsl:94-104) given r2.
__device__ __forceinline__ float scorr_r2(const StepConst& c, float r2) {
    if (c.enable_scorr == 0u) return 0.0f;
    const float w = poly6_r2(c, r2);
    if (c.wq <= 1e-12f || w <= 0.0f) return 0.0f;
    const float ratio = c.fast_div != 0u ? div_rn_by(w, c.wq, rcp_refined(c.wq)) : w / c.wq;
    if (ratio <= 0.0f) return 0.0f;
    return -c.k_corr * powr(ratio, c.n_corr);
}

// ---- SDF: GL_LINEAR + CLAMP_TO_EDGE emulation (Texture3D.cpp:25-32) ----

__device__ __forceinline__ void sdf_taps(float tc, int32_t size, int32_t& i0, int32_t& i1, float& a) {
    const float uu = tc * (float)size - 0.5f;
    float f = floorf(uu);
    a = uu - f;
    if (!(f >= -1.0f)) f = -1.0f;
    if (f > (float)size) f = (float)size;
    int32_t lo = (int32_t)f, hi = lo + 1;
    lo = lo < 0 ? 0 : (lo > size - 1 ? size - 1 : lo);
    hi = hi < 0 ? 0 : (hi > size - 1 ? size - 1 : hi);
    i0 = lo;
    i1 = hi;
}

__device__ __forceinline__ float lerp1(float t0, float t1, float a) { return (1.0f - a) * t0 + a * t1; }

// The eight voxels and weights of a trilinear sample at p.
struct SdfTaps {
    int32_t x0, x1, y0, y1, z0, z1;
    float ax, ay, az;
};
__device__ __forceinline__ SdfTaps sdf_locate(const Domain& c, V3 p) {
    const float tx = ((p.x - c.sdf_origin[0]) / c.sdf_cell + 0.5f) / (float)c.dims[0];
    const float ty = ((p.y - c.sdf_origin[1]) / c.sdf_cell + 0.5f) / (float)c.dims[1];
    const float tz = ((p.z - c.sdf_origin[2]) / c.sdf_cell + 0.5f) / (float)c.dims[2];
    SdfTaps t;
    sdf_taps(tx, c.dims[0], t.x0, t.x1, t.ax);
    sdf_taps(ty, c.dims[1], t.y0, t.y1, t.ay);
    sdf_taps(tz, c.dims[2], t.z0, t.z1, t.az);
    return t;
}
__device__ __forceinline__ float sdf_sample(const Domain& c, const SdfTaps& t) {
    const size_t W = (size_t)c.dims[0], WH = (size_t)c.dims[0] * (size_t)c.dims[1];
    const float* g = c.sdf_grid;
    const size_t r00 = (size_t)t.y0 * W + (size_t)t.z0 * WH, r10 = (size_t)t.y1 * W + (size_t)t.z0 * WH;
    const size_t r01 = (size_t)t.y0 * W + (size_t)t.z1 * WH, r11 = (size_t)t.y1 * W + (size_t)t.z1 * WH;
    const float c00 = lerp1(g[r00 + t.x0], g[r00 + t.x1], t.ax);
    const float c10 = lerp1(g[r10 + t.x0], g[r10 + t.x1], t.ax);
    const float c01 = lerp1(g[r01 + t.x0], g[r01 + t.x1], t.ax);
    const float c11 = lerp1(g[r11 + t.x0], g[r11 + t.x1], t.ax);
    const float c0 = lerp1(c00, c10, t.ay);
    const float c1 = lerp1(c01, c11, t.ay);
    return lerp1(c0, c1, t.az);
}

// querySDF (CS_Integrate.glsl:86-89).
__device__ __forceinline__ float sdf_query(const Domain& c, V3 p) { return sdf_sample(c, sdf_locate(c, p)); }

// sdfGradient (CS_ApplyDeltaP.glsl:58-65).
__device__ __forceinline__ V3 sdf_gradient(const Domain& c, V3 p) {
    const float e = c.sdf_cell;
    const float z = 0.0f;
    // one axis at a time (not unrolled): the six samples do not all hold
    // their registers at once
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
#pragma unroll 1
    for (int a = 0; a < 3; ++a) {
        const float ex = a == 0 ? e : z, ey = a == 1 ? e : z, ez = a == 2 ? e : z;
        const float ga = sdf_query(c, mk(p.x + ex, p.y + ey, p.z + ez)) - sdf_query(c, mk(p.x - ex, p.y - ey, p.z - ez));
        gx = a == 0 ? ga : gx;
        gy = a == 1 ? ga : gy;
        gz = a == 2 ? ga : gz;
    }
    return dvs(mk(gx, gy, gz), 2.0f * e);
}

// SDF projection shared by CS_ApplyDeltaP.glsl:108-121 and CS_Integrate.glsl:216-240.
// The shaders sample the gradient before testing the distance; its six
// samples only matter when the distance test passes, so they are taken
// after it (same results: a failed test returns false whatever the gradient).
//
// Before the sample, the range [m, M] of the voxels the sample reads (their
// block's) decides most particles: each lerp (1 - a) t0 + a t1 with a in
// [0, 1] rounds three times, so three levels of it give m - 9u|m| <= d <=
// M + 9u|M| (u = 2^-24) for normal values, and lb = m - |m| 2^-19 - 2^-120,
// ub = M + |M| 2^-19 + 2^-120 (rounded; the last term covers subnormal and
// flushed intermediates) lie outside that.  lb >= pad (obstacle: no
// projection unless d < pad) or ub <= -pad (container: none unless
// d > -pad) means the distance test fails whatever the eight voxels are:
// the same false as sampling them.
__device__ __forceinline__ bool sdf_out_of_reach(const Domain& c, const SdfTaps& tp) {
    if (c.sdf_brange == nullptr) return false;
    const int32_t b = (tp.x0 >> kSdfBlockBits) +
                      c.sdf_bdims[0] * ((tp.y0 >> kSdfBlockBits) + c.sdf_bdims[1] * (tp.z0 >> kSdfBlockBits));
    const float2 r = c.sdf_brange[b];
    if (c.sdf_mode == 1u) return r.y + fabsf(r.y) * 0x1p-19f + 0x1p-120f <= -c.sdf_pad;
    return r.x - fabsf(r.x) * 0x1p-19f - 0x1p-120f >= c.sdf_pad;
}
// The sampled part: the distance test, the gradient and the move.
__device__ __forceinline__ bool sdf_project_sampled(const Domain& c, const SdfTaps& tp, V3& p, V3& nrm) {
    const float pad = c.sdf_pad;
    const bool container = c.sdf_mode == 1u;
    const float d = sdf_sample(c, tp);
    if (container ? !(d > -pad) : !(d < pad)) return false;
    const V3 g = sdf_gradient(c, p);
    const float l = len(g);
    if (!(l > 1e-6f)) return false;
    if (container) {
        nrm = dvs(mk(-g.x, -g.y, -g.z), l);
        p = add(p, scl(nrm, d + pad));
    } else {
        nrm = dvs(g, l);
        p = add(p, scl(nrm, pad - d));
    }
    return true;
}
__device__ __forceinline__ bool sdf_project(const Domain& c, V3& p, V3& nrm) {
    const SdfTaps tp = sdf_locate(c, p);
    if (sdf_out_of_reach(c, tp)) return false;
    return sdf_project_sampled(c, tp, p, nrm);
}

}  // namespace fsi
