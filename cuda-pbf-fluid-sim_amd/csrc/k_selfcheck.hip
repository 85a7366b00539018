// k_selfcheck.hip — device self-checks of the pair arithmetic's unscaled IEEE
// sequences (pbf_math.hpp), for a maintainer bringing the library up on a new
// compiler or device.  No reference counterpart: the reference calls sqrtf and
// divides.  These launchers allocate, synchronise and read back; they are not part
// of the simulation step.
#include "pbf_common.hpp"
#include "pbf_hip.h"
#include "pbf_math.hpp"
#include "sorted_path.hpp"

namespace pbfx {
namespace {

constexpr int kBS = 256;

// out[0] square roots that differ from sqrtf over every finite float >= 2^-96;
// out[1] inputs in [+0, 2^-96) whose unscaled root is NaN, negative or above 2^-47;
// out[2] quotients that differ from a / b over the sampled pairs;
// out[3] zero numerators (either sign) whose quotient is not +0;
// out[4] max_nonneg results that differ from fmaxf over the sampled pairs.
enum { kSqrtDiff, kSqrtTiny, kDivDiff, kDivZero, kMaxDiff, kCounts };

__device__ __forceinline__ uint64_t mix64(uint64_t z) {  // splitmix64 finaliser
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(kBS) void k_check_sqrt(uint32_t lo, uint32_t hi, unsigned long long* out) {
    const uint32_t tiny = 0x0F800000u;  // 2^-96
    const float cap = 0x1p-47f;
    unsigned long long diff = 0, bad = 0;
    for (uint64_t u = lo + (uint64_t)blockIdx.x * kBS + threadIdx.x; u < hi; u += (uint64_t)gridDim.x * kBS) {
        const float x = __uint_as_float((uint32_t)u);
        const float r = sqrt_unscaled(x);
        if ((uint32_t)u >= tiny) {
            diff += __float_as_uint(r) != __float_as_uint(sqrtf(x));
        } else {
            bad += !(r >= 0.f && r <= cap) || (__float_as_uint(r) >> 31);
        }
    }
    if (diff) atomicAdd(out + kSqrtDiff, diff);
    if (bad) atomicAdd(out + kSqrtTiny, bad);
}

// Pairs with |a| in [2^-100, 2^101), b in [2^-100, 2^101) and a's exponent minus b's in
// [-100, 90]: the region div_unscaled_range admits, and beyond it on the low side.
__global__ __launch_bounds__(kBS) void k_check_div(uint64_t n, uint64_t seed, unsigned long long* out) {
    unsigned long long diff = 0, zero = 0, mx = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBS) {
        const uint64_t r0 = mix64(seed + 2 * i), r1 = mix64(seed + 2 * i + 1);
        const int ea = (int)(r0 % 201u) - 100;                      // [-100, 100]
        const int lo = max(-100, ea - 90), hi = min(100, ea + 100);  // eb range
        const int eb = lo + (int)((r0 >> 32) % (uint64_t)(hi - lo + 1));
        // mantissas: uniform, or (one pair in four) b's mantissa all ones / a a small
        // multiple of b, the cases nearest a rounding boundary
        uint32_t ma = (uint32_t)r1 & 0x7FFFFFu, mb = (uint32_t)(r1 >> 23) & 0x7FFFFFu;
        const uint32_t kind = (uint32_t)(r1 >> 60);
        if (kind == 0) mb = 0x7FFFFFu;
        if (kind == 1) mb = 0x7FFFFFu - (uint32_t)((r1 >> 46) & 0xFu);
        const float b = __uint_as_float(((uint32_t)(eb + 127) << 23) | mb);
        float a = __uint_as_float(((uint32_t)(ea + 127) << 23) | ma);
        if (kind == 2) a = b * (float)(1u + ((r1 >> 46) & 0x3Fu));  // exact small multiple (or its rounding)
        if (kind == 2 && !(fabsf(a) >= 0x1p-100f && fabsf(a) < 0x1p101f)) a = b;
        if ((r0 >> 63) & 1u) a = -a;
        const float y = rcp_refined(b);
        diff += __float_as_uint(div_unscaled(a, b, y)) != __float_as_uint(a / b);
        zero += __float_as_uint(div_unscaled(0.0f, b, y)) != 0u;
        zero += __float_as_uint(div_unscaled(-0.0f, b, y)) != 0u;
        const float p = __uint_as_float((uint32_t)r0 & 0x7F7FFFFFu), q = __uint_as_float((uint32_t)r1 & 0x7F7FFFFFu);
        mx += __float_as_uint(max_nonneg(p, q)) != __float_as_uint(fmaxf(p, q));
    }
    if (diff) atomicAdd(out + kDivDiff, diff);
    if (zero) atomicAdd(out + kDivZero, zero);
    if (mx) atomicAdd(out + kMaxDiff, mx);
}

__global__ void k_pair_unscaled(sim::DeviceParams dp, int* out) {
    const PairConsts c = make_pair_consts(dp);
    *out = c.unscaled ? 1 : 0;
}

}  // namespace
}  // namespace pbfx

extern "C" int pbfDebugPairUnscaled(sim::DeviceParams dp) {
    int* d = nullptr;
    int h = -1;
    if (hipMalloc(&d, sizeof(int)) != hipSuccess) {
        pbfx::set_error((int)hipErrorOutOfMemory, "pbfDebugPairUnscaled");
        return -1;
    }
    pbfx::k_pair_unscaled<<<1, 1>>>(dp, d);
    const hipError_t e = hipMemcpy(&h, d, sizeof(int), hipMemcpyDeviceToHost);
    PBFX_CHECK(hipFree(d), "pbfDebugPairUnscaled");
    if (e != hipSuccess) {
        pbfx::set_error((int)e, "pbfDebugPairUnscaled");
        return -1;
    }
    return h;
}

extern "C" int pbfDebugScorrRatio(sim::DeviceParams dp, float* out12) {
    if (!out12) return -1;
    pbfx::RatioCoef* d = nullptr;
    uint32_t* bad = nullptr;
    if (hipMalloc(&d, sizeof(pbfx::RatioCoef)) != hipSuccess || hipMalloc(&bad, sizeof(uint32_t)) != hipSuccess) {
        if (d) PBFX_CHECK(hipFree(d), "pbfDebugScorrRatio");
        pbfx::set_error((int)hipErrorOutOfMemory, "pbfDebugScorrRatio");
        return -1;
    }
    pbfx::launch_ratio_check(dp, d, bad, nullptr);
    hipError_t e = hipGetLastError();
    pbfx::RatioCoef h{};
    if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost);
    PBFX_CHECK(hipFree(d), "pbfDebugScorrRatio");
    PBFX_CHECK(hipFree(bad), "pbfDebugScorrRatio");
    if (e != hipSuccess) {
        pbfx::set_error((int)e, "pbfDebugScorrRatio");
        return -1;
    }
    const float v[12] = {h.zh, h.zl, h.ok, h.w_q, h.poly6, h.h2, h.neg_k, h.smin, h.rsq_ok, h.h, h.eps, h.neg3_spiky};
    for (int i = 0; i < 12; ++i) out12[i] = v[i];
    return 0;
}

extern "C" int pbfDebugCheckUnscaledMath(uint64_t divSamples, uint64_t seed, unsigned long long* out5) {
    if (!out5) return -1;
    unsigned long long* d = nullptr;
    if (hipMalloc(&d, pbfx::kCounts * sizeof(unsigned long long)) != hipSuccess) {
        pbfx::set_error((int)hipErrorOutOfMemory, "pbfDebugCheckUnscaledMath");
        return -1;
    }
    hipError_t e = hipMemset(d, 0, pbfx::kCounts * sizeof(unsigned long long));
    if (e == hipSuccess) {
        // every finite float from +0 up (bit patterns 0 .. 0x7F7FFFFF)
        pbfx::k_check_sqrt<<<8192, pbfx::kBS>>>(0u, 0x7F800000u, d);
        if (divSamples) pbfx::k_check_div<<<8192, pbfx::kBS>>>(divSamples, seed, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out5, d, pbfx::kCounts * sizeof(unsigned long long), hipMemcpyDeviceToHost);
    PBFX_CHECK(hipFree(d), "pbfDebugCheckUnscaledMath");
    if (e != hipSuccess) {
        pbfx::set_error((int)e, "pbfDebugCheckUnscaledMath");
        return -1;
    }
    return 0;
}
