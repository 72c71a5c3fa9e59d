// selftest.hip — libfsi_selftest.so, a test-support library (not part of the
// C-ABI of include/fsi_pbf.h): device checks of arithmetic helpers of
// pbf_device.h that the parity tests cannot reach with ordinary inputs.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbf_device.h"

namespace {

constexpr int kSB = 256;

// Every float with bits in [lo, hi]: sqrt_rn_normal against the IEEE sqrtf
// of this build (compiled with correctly rounded division and square root).
__global__ __launch_bounds__(kSB) void k_sqrt_check(uint32_t lo, uint32_t hi, unsigned long long* __restrict__ bad,
                                                    uint32_t* __restrict__ first) {
    const uint64_t stride = (uint64_t)gridDim.x * kSB;
    uint32_t nbad = 0;
    uint32_t fb = 0xffffffffu;
    for (uint64_t b = (uint64_t)lo + (uint64_t)blockIdx.x * kSB + threadIdx.x; b <= (uint64_t)hi; b += stride) {
        const float x = __uint_as_float((uint32_t)b);
        const float want = sqrtf(x), got = fsi::sqrt_rn_normal(x);
        if (__float_as_uint(want) != __float_as_uint(got)) {
            ++nbad;
            fb = min(fb, (uint32_t)b);
        }
    }
    if (nbad) {
        atomicAdd(bad, (unsigned long long)nbad);
        atomicMin(first, fb);
    }
}

// Every numerator with bits in [lo, hi] (and its negation) against each
// denominator b[blockIdx.y]: div_rn_normal and div_rn_by against the IEEE
// division of this build.
__global__ __launch_bounds__(kSB) void k_div_check(uint32_t lo, uint32_t hi, const float* __restrict__ bs,
                                                   unsigned long long* __restrict__ bad, uint32_t* __restrict__ first) {
    const float b = bs[blockIdx.y];
    const float r1 = fsi::rcp_refined(b);
    const uint64_t stride = (uint64_t)gridDim.x * kSB;
    uint32_t nbad = 0;
    uint32_t fb = 0xffffffffu;
    for (uint64_t u = (uint64_t)lo + (uint64_t)blockIdx.x * kSB + threadIdx.x; u <= (uint64_t)hi; u += stride) {
#pragma unroll
        for (uint32_t sign = 0; sign < 2; ++sign) {
            const float a = __uint_as_float((uint32_t)u | (sign << 31));
            const uint32_t want = __float_as_uint(a / b);
            if (__float_as_uint(fsi::div_rn_normal(a, b)) != want || __float_as_uint(fsi::div_rn_by(a, b, r1)) != want) {
                ++nbad;
                fb = min(fb, (uint32_t)u);
            }
        }
    }
    if (nbad) {
        atomicAdd(bad, (unsigned long long)nbad);
        atomicMin(first, fb);
    }
}

}  // namespace

extern "C" {

// Counts the floats with bit patterns in [lo_bits, hi_bits] on which
// sqrt_rn_normal and sqrtf differ; first_bad receives the smallest such
// pattern (0xffffffff if none).  Returns 0 or a hipError_t.
int fsi_selftest_sqrt(uint32_t lo_bits, uint32_t hi_bits, unsigned long long* mismatches, uint32_t* first_bad) {
    if (!mismatches || !first_bad || lo_bits > hi_bits) return (int)hipErrorInvalidValue;
    unsigned long long* d_bad = nullptr;
    uint32_t* d_first = nullptr;
    hipError_t e = hipMalloc((void**)&d_bad, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc((void**)&d_first, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(d_bad, 0, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(d_first, 0xff, sizeof(uint32_t));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_sqrt_check, dim3(8192), dim3(kSB), 0, 0, lo_bits, hi_bits, d_bad, d_first);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(mismatches, d_bad, sizeof(unsigned long long), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(first_bad, d_first, sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (d_bad) (void)hipFree(d_bad);
    if (d_first) (void)hipFree(d_first);
    return (int)e;
}

// Counts the numerators with bit patterns in [lo_bits, hi_bits], of either
// sign, on which div_rn_normal (or div_rn_by) differs from the IEEE division
// by one of the nb denominators b (host array); first_bad receives the
// smallest such pattern (0xffffffff if none).  Returns 0 or a hipError_t.
int fsi_selftest_div(uint32_t lo_bits, uint32_t hi_bits, const float* b, uint32_t nb, unsigned long long* mismatches,
                     uint32_t* first_bad) {
    if (!b || nb == 0 || nb > 65535u || !mismatches || !first_bad || lo_bits > hi_bits)
        return (int)hipErrorInvalidValue;
    unsigned long long* d_bad = nullptr;
    uint32_t* d_first = nullptr;
    float* d_b = nullptr;
    hipError_t e = hipMalloc((void**)&d_bad, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc((void**)&d_first, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&d_b, sizeof(float) * nb);
    if (e == hipSuccess) e = hipMemcpy(d_b, b, sizeof(float) * nb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(d_bad, 0, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(d_first, 0xff, sizeof(uint32_t));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_div_check, dim3(2048, nb), dim3(kSB), 0, 0, lo_bits, hi_bits, d_b, d_bad, d_first);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(mismatches, d_bad, sizeof(unsigned long long), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(first_bad, d_first, sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (d_bad) (void)hipFree(d_bad);
    if (d_first) (void)hipFree(d_first);
    if (d_b) (void)hipFree(d_b);
    return (int)e;
}

}  // extern "C"
