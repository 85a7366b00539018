// pbf_math.hpp — per-pair PBF arithmetic shared by every neighbour kernel.
//
// Expression association follows the reference (pbf_lambda.cu, pbf_delta_apply.cu,
// xsph.cu).  The code is built without implicit FMA contraction, with IEEE division
// and square root; in the neighbour-pair expressions a product feeding an add is one
// explicit fused multiply-add, in the fixed association below (the reference is
// built with nvcc's default contraction, which fuses such pairs too), and the CPU
// oracle forms the same fmaf calls, so every kernel and the oracle produce the same
// bits:
//   r2       = fma(rz, rz, fma(ry, ry, rx * rx))                  pbf_lambda.cu:68
//   density  = fma(((m * poly6) * hr2) * hr2, hr2, density)       pbf_lambda.cu:73-74
//   coeff    = ((-3 * spiky) * (t * t)) / max(r, eps)             pbf_lambda.cu:76-79
//   sumGrad2 += fma(gz, gz, fma(gy, gy, gx * gx))                 pbf_lambda.cu:80-82
//   lambda   = -(rho/rho0 - 1) / max(denom, eps)                  pbf_lambda.cu:86-107
//   s_corr   = -k * (W / W(dq))^n, floored                        pbf_delta_apply.cu:77-84
//   delta    = fma(lambda_i + lambda_j + s_corr, grad, delta)     pbf_delta_apply.cu:86-89
//   dv       = fma(v_j - v_i, W, dv)                              xsph.cu:66-72
// The per-particle expressions (lambda from its sums, the scaled and clamped
// correction, the XSPH update) keep separate roundings.
#pragma once

#include <type_traits>

#include "pbf_common.hpp"

namespace pbfx {

// s_corr's pow(W / W(dq), n) (the reference calls powf, pbf_delta_apply.cu:81; CUDA
// documents 2 ulp for it).  Exponents 1 to 4 are float products by squaring:
// x, x*x, x*(x*x), (x*x)*(x*x), within 1.5 ulp of the power (two roundings, the first
// doubled by the squaring).  Other integral exponents up to 16 are the same binary
// exponentiation in double, rounded once to float; the rest use powf.  The CPU oracle
// evaluates the same products in the same order, so both sides produce the same bits.
__device__ __forceinline__ float pow_scorr(float x, float n, int ni) {
    switch (ni) {
        case 1: return x;
        case 2: return x * x;
        case 3: return x * (x * x);
        case 4: {
            const float x2 = x * x;
            return x2 * x2;
        }
        default: break;
    }
    if (ni > 0) {
        double b = (double)x, r = 1.0;
        int e = ni;
        while (true) {
            if (e & 1) r *= b;
            e >>= 1;
            if (!e) break;
            b *= b;
        }
        return (float)r;
    }
    return powf(x, n);
}

// Squared distance of a pair, r2 = fma(rz, rz, fma(ry, ry, rx * rx)): the one form
// every neighbour test and pair expression uses (kernels and oracle alike).
__device__ __forceinline__ float pair_r2(float rx, float ry, float rz) {
    return __builtin_fmaf(rz, rz, __builtin_fmaf(ry, ry, rx * rx));
}

// The integral exponent of pow_scorr (0 when n is not an integer in [1, 16]).
__device__ __forceinline__ int scorr_int_exponent(float n) {
    return (n >= 1.0f && n <= 16.0f && (float)(int)n == n) ? (int)n : 0;
}

// ------------------------------------------------ unscaled IEEE sequences
// The compiler's IEEE sqrtf and division on gfx950 wrap a core sequence in scaling
// steps for operands near the ends of the exponent range.  Where the pair arithmetic
// can prove its operands stay clear of those ends (pair_unscaled_ok), the core
// sequence alone returns the same bits in fewer instructions.

// Square root from the hardware reciprocal square root y = rsq(x) and one correction
// (Markstein's form): g = x*y, then s = g + (x - g*g) * (y/2) with both steps fused.
// Six instructions against nine for the raw v_sqrt_f32 and a choice between its one-ulp
// neighbours.  The device self-check (pbfDebugCheckUnscaledMath) compares it with
// sqrtf for every finite float >= 2^-96 and finds no difference on gfx950: correctly
// rounded there.  The guard keeps x = 0 finite (rsq(0) is inf); for 0 <= x < 2^-96 the
// result is non-negative and at most 2^-47 (also checked for every such input).
__device__ __forceinline__ float sqrt_unscaled(float x) {
    const float y = __builtin_amdgcn_rsqf(fmaxf(x, 0x1p-96f));
    const float g = x * y;
    const float half_y = 0.5f * y;
    const float r = __builtin_fmaf(-g, g, x);
    return __builtin_fmaf(r, half_y, g);
}

// Reciprocal of b refined by one Newton step: the first half of the division below,
// which depends on b alone (so a constant divisor pays for it once).
__device__ __forceinline__ float rcp_refined(float b) {
    const float y = __builtin_amdgcn_rcpf(b);
    return __builtin_fmaf(__builtin_fmaf(-b, y, 1.0f), y, y);
}

// a / b from y = rcp_refined(b): a product and two Newton corrections of the quotient
// (v_div_scale and v_div_fixup dropped, v_div_fmas as a plain fma).  Those steps are
// the identity when b and a / b are normal, 1 / b is normal, |a| >= 2^-103 and a's
// exponent exceeds b's by less than 96; there this is the correctly rounded quotient.
// a = +-0 gives +0 (IEEE: a zero with the sign of a / b), which no sum in the pair
// arithmetic can tell apart: every accumulator starts at +0.
__device__ __forceinline__ float div_unscaled(float a, float b, float y) {
    float q = a * y;
    float e = __builtin_fmaf(-b, q, a);
    q = __builtin_fmaf(e, y, q);
    e = __builtin_fmaf(-b, q, a);
    return __builtin_fmaf(e, y, q);
}

// max(a, b) for a, b >= +0 and not NaN, whose bit patterns order as their values
// (equal to fmaxf there, without its NaN-quieting step).
__device__ __forceinline__ float max_nonneg(float a, float b) {
    return __uint_as_float(max(__float_as_uint(a), __float_as_uint(b)));
}

// Constants hoisted out of the neighbour loops (each is the first product the
// reference's left-to-right expression forms, so hoisting does not change bits).
struct PairConsts {
    float h, h2, poly6, m_poly6, neg3_spiky, grad_r_eps;
    float w_q, neg_scorr_k, scorr_n, scorr_min;
    int scorr_enable, scorr_ni;
    float rcp_w_q;  // rcp_refined(w_q)
    float rcp_eps;  // rcp_refined(grad_r_eps)
    bool unscaled;  // pair_unscaled_ok: the pair arithmetic uses the unscaled sequences
    // the checked shortcuts, taken together when both were checked for these constants:
    // the two-instruction s_corr ratio (scorr_ratio2) and the spiky coefficient from the
    // square root's rsq (spiky_coef_rsq)
    float rq_hi = 0.f, rq_lo = 0.f;
    bool ratio2 = false;
};

// s_corr's ratio W / w_q with the divisor fixed for a launch, from the reciprocal's two
// parts zh = RN(1 / w_q) and zl = RN(1 / w_q - zh): two instructions against
// div_unscaled's five.  It is not the correctly rounded quotient for every divisor and
// dividend (with the benchmark's w_q it differs below W = 3e-26, where W^4 underflows),
// so it is used only where the simulator has checked, on the device, that the s_corr
// term it gives equals the division's for every float W from +0 to the largest W a pair
// can have (launch_ratio_check).
__device__ __forceinline__ float scorr_ratio2(float w, float zh, float zl) {
    return __builtin_fmaf(w, zh, w * zl);
}
// The fixed form's s_corr term from its ratio: -k (ratio^2)^2, floored at scorr_min.
__device__ __forceinline__ float scorr_fixed(float ratio, const PairConsts& c) {
    const float x2 = ratio * ratio;  // pow_scorr's n = 4 product
    return fmaxf(c.neg_scorr_k * (x2 * x2), c.scorr_min);
}

// The spiky gradient's coefficient neg3_spiky (h - r)^2 / max(r, eps), r = sqrt(r2), in
// the unscaled sequences (reference: pbf_lambda.cu:76-79, pbf_delta_apply.cu:71-74).
__device__ __forceinline__ float spiky_coef(float r2, const PairConsts& c) {
    const float r = sqrt_unscaled(r2);  // +0 or positive: r2 is a sum of squares
    const float t = c.h - r;
    const float r_safe = max_nonneg(r, c.grad_r_eps);
    return div_unscaled(c.neg3_spiky * (t * t), r_safe, rcp_refined(r_safe));
}
// The same coefficient with the reciprocal's Newton step started from the square root's
// own rsq instead of a v_rcp: min(rsq(r2), 1 / eps) is within a few ulps of
// 1 / max(r, eps) (r within half an ulp of sqrt(r2)), one transcendental fewer a pair.
// A different start can change the refined reciprocal's last bit, so it is used only
// where the simulator has checked, on the device, that it gives spiky_coef's bits for
// every float r2 from +0 to h2 (launch_ratio_check).  A NaN r2 gives NaN in both.
__device__ __forceinline__ float spiky_coef_rsq(float r2, const PairConsts& c) {
    const float y = __builtin_amdgcn_rsqf(fmaxf(r2, 0x1p-96f));
    const float g = r2 * y;  // sqrt_unscaled's steps, sharing y
    const float half_y = 0.5f * y;
    const float r = __builtin_fmaf(__builtin_fmaf(-g, g, r2), half_y, g);
    const float t = c.h - r;
    const float r_safe = max_nonneg(r, c.grad_r_eps);
    // both positive and finite (y <= 2^48, eps >= 2^-40 in the unscaled form): their
    // bit patterns order as their values
    const float y0 = __uint_as_float(min(__float_as_uint(y), __float_as_uint(c.rcp_eps)));
    const float yr = __builtin_fmaf(__builtin_fmaf(-r_safe, y0, 1.0f), y0, y0);
    return div_unscaled(c.neg3_spiky * (t * t), r_safe, yr);
}

// Quotients a / b with a in {0} or [amin, amax] and b in [bmin, bmax] (all positive)
// meet div_unscaled's conditions, with margins of 2^6 and more on every bound.
__device__ __forceinline__ bool div_unscaled_range(double amin, double amax, double bmin, double bmax) {
    return amin >= 0x1p-100 && amax <= 0x1p100 && bmin >= 0x1p-100 && bmax <= 0x1p100 && amax / bmin <= 0x1p90 &&
           amin / bmax >= 0x1p-100;
}

// Whether every square root and division of lambda_pair and delta_pair is exact in
// its unscaled form, from bounds on their operands:
//  - r = sqrt(r2), r2 in [0, h2]: below 2^-96 the unscaled result is at most 2^-47,
//    and r enters the sums only through t = h - r (== h for r < ulp(h) / 2) and
//    max(r, eps) (== eps for r < eps), so h >= 2^-16 and eps >= 2^-40 make it moot;
//  - coeff = (neg3_spiky * t^2) / max(r, eps): t = h - r is 0 or at least h * 2^-25
//    in magnitude (h - r is exact when r >= h / 2, a multiple of half an ulp of h),
//    and r <= h (1 + 2^-23);
//  - ratio = poly6 * hr2^3 / w_q: hr2 = h2 - r2 is 0 or at least h2 * 2^-25 for
//    the same reason.
__device__ __forceinline__ bool pair_unscaled_ok(const PairConsts& c) {
    const double h = c.h, eps = c.grad_r_eps, k = fabs((double)c.neg3_spiky);
    if (!(h >= 0x1p-16 && h <= 0x1p16 && eps >= 0x1p-40 && eps <= h)) return false;
    const double tmin = h * 0x1p-26, tmax = h * (1.0 + 0x1p-20);
    if (!div_unscaled_range(k * tmin * tmin, k * tmax * tmax, eps, tmax)) return false;
    if (c.scorr_enable && c.w_q > 0.f) {
        const double p = c.poly6, lo = (double)c.h2 * 0x1p-26, hi = (double)c.h2 * (1.0 + 0x1p-20);
        if (!div_unscaled_range(p * lo * lo * lo, p * hi * hi * hi, c.w_q, c.w_q)) return false;
    }
    return true;
}

// The common setting: unscaled sequences, s_corr on with W(dq) > 0, exponent 4 and a
// floor below zero.  In it the pair functions' uniform branches are fixed at compile
// time (their kFixed form): the same expressions, without a dozen scalar branches per
// pair in the kernels' hot loops.
__device__ __forceinline__ bool pair_fixed(const PairConsts& c) {
    return c.unscaled && c.scorr_enable && c.w_q > 0.f && c.scorr_ni == 4 && c.scorr_min < 0.f;
}

// f(std::true_type{}) in the common setting, f(std::false_type{}) otherwise: a kernel
// instantiates its walk for both forms and takes one per launch.
template <class F>
__device__ __forceinline__ void with_pair_form(const PairConsts& c, F&& f) {
    if (pair_fixed(c))
        f(std::true_type{});
    else
        f(std::false_type{});
}

// The delta kernels' forms: the common setting with the checked shortcuts (the
// two-instruction s_corr ratio and the spiky coefficient's rsq start), the common
// setting with the full sequences, and the general form.
template <bool kF, bool kR>
struct DeltaForm {
    static constexpr bool kFixed = kF, kRatio2 = kR;
};
template <class F>
__device__ __forceinline__ void with_delta_form(const PairConsts& c, F&& f) {
    if (pair_fixed(c)) {
        if (c.ratio2)
            f(DeltaForm<true, true>{});
        else
            f(DeltaForm<true, false>{});
    } else {
        f(DeltaForm<false, false>{});
    }
}

__device__ __forceinline__ PairConsts make_pair_consts(const sim::DeviceParams& dp) {
    PairConsts c;
    c.h = dp.kernel.h;
    c.h2 = dp.kernel.h2;
    c.poly6 = dp.kernel.poly6;
    c.m_poly6 = dp.particleMass * dp.kernel.poly6;
    c.neg3_spiky = -3.0f * dp.kernel.spiky;
    c.grad_r_eps = dp.pbf.grad_r_eps;
    const float dq = dp.pbf.scorr_dq_h * dp.kernel.h;
    const float hr2_q = dp.kernel.h2 - dq * dq;
    c.w_q = (hr2_q > 0.f) ? (dp.kernel.poly6 * hr2_q * hr2_q * hr2_q) : dp.pbf.wq_min;
    c.neg_scorr_k = -dp.pbf.scorr_k;
    c.scorr_n = dp.pbf.scorr_n;
    c.scorr_min = dp.pbf.scorr_min;
    c.scorr_enable = dp.pbf.scorr_enable;
    c.scorr_ni = scorr_int_exponent(dp.pbf.scorr_n);
    c.rcp_w_q = rcp_refined(c.w_q);
    c.rcp_eps = rcp_refined(c.grad_r_eps);
    c.unscaled = pair_unscaled_ok(c);
    return c;
}

struct LambdaAcc {
    float density = 0.f, gx = 0.f, gy = 0.f, gz = 0.f, sumGrad2 = 0.f;
    int n = 0;
};

// One in-range neighbour (range already tested: r2 <= h2, or NaN) of the lambda sum.  kFixed: the
// caller has established pair_fixed(c).
template <bool kFixed = false>
__device__ __forceinline__ void lambda_pair(LambdaAcc& a, float rx, float ry, float rz, float r2,
                                            const PairConsts& c) {
    a.n++;
    const float hr2 = c.h2 - r2;
    a.density = __builtin_fmaf(c.m_poly6 * hr2 * hr2, hr2, a.density);
    float coeff;
    if (kFixed || c.unscaled) {
        coeff = spiky_coef(r2, c);
    } else {
        const float r = sqrtf(r2);
        const float r_safe = fmaxf(r, c.grad_r_eps);
        const float t = c.h - r;
        coeff = c.neg3_spiky * (t * t) / r_safe;
    }
    const float ax = coeff * rx, ay = coeff * ry, az = coeff * rz;
    a.gx += ax;
    a.gy += ay;
    a.gz += az;
    a.sumGrad2 += __builtin_fmaf(az, az, __builtin_fmaf(ay, ay, ax * ax));
}

// Lambda from the accumulated sums (reference: pbf_lambda.cu:85-107).
__device__ __forceinline__ float lambda_finish(const LambdaAcc& a, const sim::DeviceParams& dp) {
    const float gradSum2 = a.gx * a.gx + a.gy * a.gy + a.gz * a.gz;
    const float scale = (dp.restDensity > 0.f) ? (dp.particleMass / dp.restDensity) : 0.f;
    float denom = (scale * scale) * (a.sumGrad2 + gradSum2) + dp.pbf.lambda_denom_eps;
    if (dp.pbf.compliance > 0.f && dp.dt > 0.f) denom += dp.pbf.compliance / (dp.dt * dp.dt);
    denom = fmaxf(denom, dp.pbf.lambda_denom_eps);
    if (a.n < 2 || (a.sumGrad2 + gradSum2) < 1e-12f) return 0.0f;
    const float C = (dp.restDensity > 0.f) ? (a.density / dp.restDensity - 1.f) : 0.f;
    float lam = -C / denom;
    if (!isfinite(lam)) lam = 0.f;
    if (dp.pbf.enable_lambda_clamp) {
        if (fabsf(lam) > dp.pbf.lambda_max_abs) lam = copysignf(dp.pbf.lambda_max_abs, lam);
    }
    return lam;
}

// Coefficients of the reverse pass at the branch lambda_finish took: with
// lambda = -C / den, dlambda = b0 dC' + e0 dden' where C' = sum W (so dC = scale dC')
// and den' = sum |g|^2 + |sum g|^2 (den = scale^2 den' + const):
//   b0 = -scale / den,  e0 = C scale^2 / den^2.
// Both are zero where lambda is gated, non-finite or clamped; e0 is zero where den
// was floored at lambda_denom_eps.
__device__ __forceinline__ void lambda_adjoint_coefs(const LambdaAcc& a, const sim::DeviceParams& dp, float& b0,
                                                     float& e0) {
    b0 = 0.f;
    e0 = 0.f;
    const float gradSum2 = a.gx * a.gx + a.gy * a.gy + a.gz * a.gz;
    const float scale = (dp.restDensity > 0.f) ? (dp.particleMass / dp.restDensity) : 0.f;
    float den = (scale * scale) * (a.sumGrad2 + gradSum2) + dp.pbf.lambda_denom_eps;
    if (dp.pbf.compliance > 0.f && dp.dt > 0.f) den += dp.pbf.compliance / (dp.dt * dp.dt);
    const bool floored = !(den > dp.pbf.lambda_denom_eps);
    den = fmaxf(den, dp.pbf.lambda_denom_eps);
    if (a.n < 2 || (a.sumGrad2 + gradSum2) < 1e-12f) return;
    const float C = (dp.restDensity > 0.f) ? (a.density / dp.restDensity - 1.f) : 0.f;
    const float lam = -C / den;
    if (!isfinite(lam)) return;
    if (dp.pbf.enable_lambda_clamp && fabsf(lam) > dp.pbf.lambda_max_abs) return;
    b0 = -scale / den;
    e0 = floored ? 0.f : C * (scale * scale) / (den * den);
}

// One in-range neighbour of the position-correction sum.  kFixed: the caller has
// established pair_fixed(c); kRatio2 (with kFixed): also c.ratio2 (both checked shortcuts).
template <bool kFixed = false, bool kRatio2 = false>
__device__ __forceinline__ void delta_pair(float& dx, float& dy, float& dz, float rx, float ry, float rz, float r2,
                                           float lam_i, float lam_j, const PairConsts& c) {
    float coeff;
    if (kFixed || c.unscaled) {
        coeff = kRatio2 ? spiky_coef_rsq(r2, c) : spiky_coef(r2, c);
    } else {
        const float r = sqrtf(r2);
        const float t = c.h - r;
        const float r_safe = fmaxf(r, c.grad_r_eps);
        coeff = c.neg3_spiky * (t * t) / r_safe;
    }
    const float ax = coeff * rx, ay = coeff * ry, az = coeff * rz;
    float scorr = 0.0f;
    if (kFixed) {
        const float hr2 = c.h2 - r2;
        const float w = c.poly6 * hr2 * hr2 * hr2;
        scorr = scorr_fixed(kRatio2 ? scorr_ratio2(w, c.rq_hi, c.rq_lo) : div_unscaled(w, c.w_q, c.rcp_w_q), c);
    } else if (c.scorr_enable) {
        const float hr2 = c.h2 - r2;
        const float w = c.poly6 * hr2 * hr2 * hr2;
        const float ratio = (c.w_q > 0.f) ? (c.unscaled ? div_unscaled(w, c.w_q, c.rcp_w_q) : (w / c.w_q)) : 0.0f;
        scorr = c.neg_scorr_k * pow_scorr(ratio, c.scorr_n, c.scorr_ni);
        if (c.scorr_min < 0.f) scorr = fmaxf(scorr, c.scorr_min);
    }
    const float lij = (lam_i + lam_j + scorr);
    dx = __builtin_fmaf(lij, ax, dx);
    dy = __builtin_fmaf(lij, ay, dy);
    dz = __builtin_fmaf(lij, az, dz);
}

// One in-range neighbour of the XSPH sum: dv += (v_j - v_i) W (fused).
__device__ __forceinline__ void xsph_pair(float& sx, float& sy, float& sz, float4 vj, float4 vi, float w) {
    sx = __builtin_fmaf(vj.x - vi.x, w, sx);
    sy = __builtin_fmaf(vj.y - vi.y, w, sy);
    sz = __builtin_fmaf(vj.z - vi.z, w, sz);
}

// Scale and relax of the summed correction, before the clamp
// (reference: pbf_delta_apply.cu:93-100).
__device__ __forceinline__ float3 delta_scaled(float dx, float dy, float dz, const sim::DeviceParams& dp) {
    const float invRest = (dp.restDensity > 0.f) ? (1.0f / dp.restDensity) : 0.f;
    float ex = invRest * dx, ey = invRest * dy, ez = invRest * dz;
    if (dp.pbf.enable_relax) {
        ex *= dp.pbf.relax_omega;
        ey *= dp.pbf.relax_omega;
        ez *= dp.pbf.relax_omega;
    }
    return make_float3(ex, ey, ez);
}

// Displacement clamp of a scaled correction (reference: pbf_delta_apply.cu:101-110).
__device__ __forceinline__ float3 delta_clamp(float3 e, const sim::DeviceParams& dp) {
    if (dp.pbf.enable_disp_clamp) {
        const float maxDisp = dp.pbf.disp_clamp_max_h * dp.kernel.h;
        const float maxDisp2 = maxDisp * maxDisp;
        const float len2 = e.x * e.x + e.y * e.y + e.z * e.z;
        if (len2 > maxDisp2) {
            const float len = sqrtf(len2);
            const float s = maxDisp / (len + 1e-20f);
            e.x *= s;
            e.y *= s;
            e.z *= s;
        }
    }
    return e;
}

__device__ __forceinline__ float3 delta_finish(float dx, float dy, float dz, const sim::DeviceParams& dp) {
    return delta_clamp(delta_scaled(dx, dy, dz, dp), dp);
}

// XSPH neighbour-count gate (reference: xsph.cu:80-88).
__device__ __forceinline__ float xsph_gate(int n, const sim::DeviceParams& dp) {
    float gate = 1.0f;
    if (dp.pbf.xsph_gate_enable) {
        const int nMin = max(0, dp.pbf.xsph_n_min);
        const int nMax = max(nMin + 1, dp.pbf.xsph_n_max);
        const float t = (float(n) - float(nMin)) / float(nMax - nMin);
        gate = fminf(1.f, fmaxf(0.f, t));
    }
    return gate;
}

// XSPH update of v_i from the summed (v_j - v_i) W (reference: xsph.cu:80-93).
__device__ __forceinline__ float3 xsph_finish(float vx, float vy, float vz, float sx, float sy, float sz, int n,
                                              const sim::DeviceParams& dp, float xsph_c) {
    const float gate = xsph_gate(n, dp);
    const float mor = (dp.restDensity > 0.f) ? (dp.particleMass / dp.restDensity) : 0.f;
    const float cg = xsph_c * gate;
    return make_float3(vx + cg * mor * sx, vy + cg * mor * sy, vz + cg * mor * sz);
}

// --------------------------------------------------------- row enumeration
// For a neighbour row (dz, dy) the cells x-reach..x+reach (clipped to the grid)
// hold consecutive keys, so their sorted slots form one contiguous span that
// visits them in the reference's dx-then-slot order.

// Dense provider over start/end arrays with sentinel empties.
struct DenseRows {
    const uint32_t* __restrict__ start;
    const uint32_t* __restrict__ end;
    __device__ __forceinline__ void span(uint32_t k0, uint32_t k1, uint32_t& lo, uint32_t& hi) const {
        lo = 0xFFFFFFFFu;
        hi = 0u;
        for (uint32_t c = k0; c <= k1; ++c) {
            const uint32_t b = start[c], e = end[c];
            if (b == 0xFFFFFFFFu || b >= e) continue;
            lo = min(lo, b);
            hi = max(hi, e);
        }
        if (lo == 0xFFFFFFFFu) lo = hi = 0;
    }
};

// Compact provider: one lower_bound over the unique keys per row
// (reference: cuda_grid_utils.cuh:34-67 searches once per cell).
struct CompactRows {
    const uint32_t* __restrict__ uniq;
    const uint32_t* __restrict__ off;
    uint32_t M;
    __device__ __forceinline__ void span(uint32_t k0, uint32_t k1, uint32_t& lo, uint32_t& hi) const {
        uint32_t a = 0, b = M;
        while (a < b) {
            const uint32_t mid = (a + b) >> 1;
            if (uniq[mid] < k0) a = mid + 1; else b = mid;
        }
        uint32_t e = a;
        while (e < M && uniq[e] <= k1) ++e;
        if (e > a) { lo = off[a]; hi = off[e]; } else { lo = hi = 0; }
    }
};

}  // namespace pbfx
