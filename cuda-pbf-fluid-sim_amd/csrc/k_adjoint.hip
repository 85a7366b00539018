// k_adjoint.hip — hand-written reverse pass of one recorded PBF step.
//
// Forward step (sorted order, per particle i; reference: phases.cpp:30-106,
// post_ops.cpp:25-71):
//   P0 = clamp(X + (V + g dt) dt)
//   for it: lambda_i = -C_i / den_i,  C_i = rho_i/rho0 - 1,  rho_i = m sum_j W6(r_ij),
//           den_i = (m/rho0)^2 (sum_j |g_ij|^2 + |sum_j g_ij|^2) + eps (+ compliance/dt^2)
//           u_i = sum_j (lambda_i + lambda_j + scorr_ij) g_ij,   D_i = clampDisp(omega/rho0 * u_i)
//           P_{it+1} = P_it + D
//   Pf = clamp(P_I);  Vn = (Pf - X) / dt
//   Vout_i = Vn_i + kappa_i sum_j (Vn_j - Vn_i) W6(Pf_i - Pf_j),  kappa_i = c gate_i m / rho0
// with g_ij = f(r) r_ij the spiky gradient, f(r) = -3 spiky (h - r)^2 / max(r, eps).
//
// Reverse, in the order below; each kernel gathers over its own neighbours, so
// every particle's gradient is written by one lane (no atomics, deterministic).
// The forward pass, while recording, also kept what the reverse pass would
// otherwise walk the neighbours again for: the lambda coefficients (sum of spiky
// gradients, b0, e0, the slot where the neighbour cap cut the walk), the scaled
// correction before the displacement clamp, and the XSPH gate.
//   gather   dL/dPf, dL/dVout to sorted order
//   xsph     dL/dVn_k = gVo_k (1 - kappa_k sum_j W_kj) + sum_j kappa_j W_kj gVo_j
//            dL/dPf_k += sum_j ((kappa_k gVo_k - kappa_j gVo_j) . (Vn_j - Vn_k)) dW6_kj
//            dL/dc    += gVo_k . (gate_k m/rho0 sum_j (Vn_j - Vn_k) W_kj)
//   velocity dL/dP_I = finalMask (gPf + gVn / dt);   dL/dX = -gVn / dt
//            and a of the last iteration (below)
//   per iteration, last first:
//     delta   dL/dlambda_k = sum_j g_kj . (a_k - a_j)
//             dL/dP_k += sum_j T_kj (a_k - a_j),  T_kj = c_kj H_kj + g_kj (x) dscorr_kj
//             (H the Hessian of the spiky kernel; T is even in r_kj, hence symmetric)
//             then beta_k = gLam_k b0_k, eps_k = gLam_k e0_k
//     lambda  dL/dP_k += sum_{j in Ncap(k)} [beta_k dW6_kj + 2 eps_k H_kj (g_kj + G_k)]
//                      + sum_{j : k in Ncap(j)} [beta_j dW6_kj + 2 eps_j H_kj (g_kj - G_j)]
//             (Ncap: the first maxNeighbors in-range neighbours in walk order), then
//             a_i = (omega/rho0) J_i^T gP_i of the iteration before (J: Jacobian of the
//             displacement clamp at the recorded scaled correction)
//   predict  dL/dX += predMask gP0,  dL/dV = dt predMask gP0
// Walk order is ascending slot order: rows go dz-major then dy, every row is one
// x-line of cells, and slots are sorted by the row-major cell key, so the lines
// (and the slots) of later rows come later.  "j in Ncap(k)" is therefore
// j <= cut(k), with cut(k) the slot at which k's walk reached the cap (~0 if never).
#include "adjoint.hpp"
#include "pbf_common.hpp"
#include "pbf_hip.h"
#include "pbf_math.hpp"
#include "walk.hpp"

namespace pbfx {
namespace {

using namespace walk;

__device__ __forceinline__ float3 f3(const float4& v) { return make_float3(v.x, v.y, v.z); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 add3(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 mul3(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float4 mk4(float3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

__device__ __forceinline__ float3 mask3(float3 g, uint32_t bits) {
    return make_float3((bits & 1u) ? 0.f : g.x, (bits & 2u) ? 0.f : g.y, (bits & 4u) ? 0.f : g.z);
}

// Spiky gradient factor and its Hessian applied to a vector: g = f r,
// H v = f v + (df/dr / |r|) (r . v) r.
struct Spiky {
    float f, dfr;  // f(r) and (df/dr)/|r|
};
// One division per pair: 1 / r, correctly rounded, then products.  kFixed: the
// caller has established pair_fixed(c) (so the unscaled sequences apply).
template <bool kFixed>
__device__ __forceinline__ Spiky spiky_of(float r2, const PairConsts& c) {
    const bool unscaled = kFixed || c.unscaled;
    const float r = unscaled ? sqrt_unscaled(r2) : sqrtf(r2);
    const float t = c.h - r;
    Spiky s;
    if (r > c.grad_r_eps) {
        // r in (eps, h]: inside div_unscaled's range whenever the pair arithmetic is
        const float ir = unscaled ? div_unscaled(1.0f, r, rcp_refined(r)) : 1.0f / r;
        const float t2 = t * t;
        s.f = c.neg3_spiky * t2 * ir;
        // f = -3 spiky t^2 / r  ->  df/dr = 3 spiky (2 t / r + t^2 / r^2)
        s.dfr = (-c.neg3_spiky) * (2.0f * t * ir + t2 * ir * ir) * ir;
    } else {
        s.f = c.neg3_spiky * (t * t) / c.grad_r_eps;
        // r below the clamp: f depends on r only through t^2.  The Hessian's second
        // term, dfr (r . v) r, is (df/dr) r times a unit-vector product: relative to f v
        // it is 2 r / t, below float32 resolution once r <= h 2^-30, where it is dropped
        // (there (df/dr) / r would overflow for the near-coincident pairs of a collapsed
        // pile-up and meet a zero component of r as inf * 0).
        s.dfr = (r > c.h * 0x1p-30f) ? ((-c.neg3_spiky) * 2.0f * t / c.grad_r_eps) / r : 0.f;
    }
    return s;
}
// (W / Wq)^(n - 1), the power in s_corr's derivative: for integral n the product in
// double that pow_scorr forms, otherwise powf.
__device__ __forceinline__ float pow_scorr_deriv(float x, const PairConsts& c) {
    if (c.scorr_ni == 1) return 1.0f;
    if (c.scorr_ni > 1) return pow_scorr(x, c.scorr_n - 1.0f, c.scorr_ni - 1);
    return powf(x, c.scorr_n - 1.0f);
}

__device__ __forceinline__ float3 hess_apply(const Spiky& s, float3 r, float3 v) {
    return add3(mul3(v, s.f), mul3(r, s.dfr * dot3(r, v)));
}

// Launch shape: the three neighbour kernels below deal their 256-slot tiles to the XCDs
// in contiguous runs (xcd_tile), as the forward kernels do, so neighbouring tiles share
// an L2; dL/dc partials are stored per tile, so their sum keeps slot order.
//
// Occupancy: the three neighbour kernels below are held to 72 VGPRs (seven waves per
// SIMD, amdgpu_waves_per_eu(7)); left alone they take 80-89 (five or six waves), and
// the extra waves outweigh the few spills (8% faster reverse pass on 4.2M particles).
//
// The neighbour kernels below gather over every in-range neighbour j != k of their
// slot with the forward kernels' wave walk (walk.hpp): the same candidates, the same
// r^2 <= h^2 test and the same ascending-slot order, with the pair work deferred
// through the per-lane FIFO so a wave runs it with most lanes busy.  Lanes past N
// take part in the walk's wave-level steps with no rows.

// gVo carries the slot's recorded XSPH factor gm = gate m / rho0 in w, so the XSPH
// adjoint gathers a neighbour's gradient and factor with one load.
// gmax (renormalised form): the step's running maxima (one per scene, kGmaxStride
// words apart), cleared here before k_adj_predict of the same step raises them.
__global__ __launch_bounds__(kBS) void k_adj_gather(const uint32_t* __restrict__ idxS, const float4* __restrict__ gX,
                                                    const float4* __restrict__ gV, const float* __restrict__ gm,
                                                    float4* __restrict__ gPf, float4* __restrict__ gVo, uint32_t N,
                                                    uint32_t* __restrict__ gmax, uint32_t scenes) {
    const uint32_t s = blockIdx.x * kBS + threadIdx.x;
    if (gmax && s == 0)
        for (uint32_t q = 0; q < scenes; ++q) gmax[q * kGmaxStride] = 0u;
    if (s >= N) return;
    const uint32_t i = idxS[s];
    gPf[s] = gX[i];
    gVo[s] = mk4(f3(gV[i]), gm[s]);
}

// a = s J^T g for the recorded scaled correction e (s = omega / rho0, or 1 / rho0
// without relaxation).  J = maxDisp / |e| (I - e^ e^T) where the clamp was active.
__device__ __forceinline__ float3 disp_adjoint(float3 g, const float4& e4, const sim::DeviceParams& dp) {
    const float invRest = (dp.restDensity > 0.f) ? (1.0f / dp.restDensity) : 0.f;
    const float sc = dp.pbf.enable_relax ? invRest * dp.pbf.relax_omega : invRest;
    if (dp.pbf.enable_disp_clamp) {
        const float3 e = f3(e4);
        const float maxDisp = dp.pbf.disp_clamp_max_h * dp.kernel.h;
        const float len2 = dot3(e, e);
        if (len2 > maxDisp * maxDisp) {
            const float len = sqrtf(len2);
            const float3 eh = mul3(e, 1.0f / len);
            g = mul3(sub3(g, mul3(eh, dot3(eh, g))), maxDisp / (len + 1e-20f));
        }
    }
    return mul3(g, sc);
}

// XSPH coefficient of a sorted slot: the scene's in an ensemble.
__device__ __forceinline__ float xsph_c_of(uint32_t s, const sim::DeviceParams& dp, const Scenes& sc,
                                          const uint32_t* sceneOfSlot) {
    return sc.params ? sc.params[sceneOfSlot[s]].xsph_c : dp.xsph_c;
}

__global__ __launch_bounds__(kBS) __attribute__((amdgpu_waves_per_eu(7))) void k_adj_xsph(const float4* __restrict__ Pf, const float4* __restrict__ Vn,
                                                  const uint32_t* __restrict__ segOf, SpanTable st,
                                                  sim::DeviceParams dp, Scenes sc,
                                                  const uint32_t* __restrict__ sceneOfSlot, uint32_t N,
                                                  const float4* __restrict__ gVo,
                                                  float4* __restrict__ gVn, float4* __restrict__ gPf,
                                                  double* __restrict__ partial, double* __restrict__ gcSlot) {
    __shared__ double sh[kBS];
    __shared__ uint32_t sQ[kQD * kBS];
    __shared__ float4 sBuf[kBS];
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t s = tile * kBS + threadIdx.x;
    const bool valid = s < N;
    const uint32_t ss = valid ? s : N - 1;
    double gcl = 0.0;
    {
        const float h2 = dp.kernel.h2, poly6 = dp.kernel.poly6;
        // neighbours share the slot's scene: rows never leave a scene's key range
        const float c = xsph_c_of(ss, dp, sc, sceneOfSlot);
        const float4 gk4 = gVo[ss];
        const float3 pk = f3(Pf[ss]), vk = f3(Vn[ss]), gk = f3(gk4);
        const float gmk = gk4.w;
        const float kk = c * gmk;
        float sumW = 0.f;
        float3 accV = make_float3(0.f, 0.f, 0.f), accP = accV, dv = accV;
        auto dist2 = [&](const float4& pj) {
            const float3 r = sub3(pk, f3(pj));
            return pair_r2(r.x, r.y, r.z);
        };
        auto apply = [&](uint32_t j) {
            const float3 r = sub3(pk, f3(Pf[j]));
            const float r2 = pair_r2(r.x, r.y, r.z);
            const float hr2 = h2 - r2;
            const float W = poly6 * hr2 * hr2 * hr2;
            const float3 vj = f3(Vn[j]);
            const float3 dvj = sub3(vj, vk);
            dv = add3(dv, mul3(dvj, W));
            if (c != 0.f) {
                const float3 dW = mul3(r, -6.0f * poly6 * hr2 * hr2);
                const float4 gj4 = gVo[j];
                const float3 gj = f3(gj4);
                const float kj = c * gj4.w;
                sumW += W;
                accV = add3(accV, mul3(gj, kj * W));
                accP = add3(accP, mul3(dW, dot3(sub3(mul3(gk, kk), mul3(gj, kj)), dvj)));
            }
        };
        neighbor_walk<kBS>(valid, s, SpanRows(st, segOf[ss]), SortedFetch{SlotBuf(Pf, N + kSlotPad)},
                                  sBuf + (threadIdx.x & ~63u), sQ + threadIdx.x, 0, h2, dist2, apply);
        if (valid) {
            gVn[s] = mk4(add3(mul3(gk, 1.0f - kk * sumW), accV), 0.f);
            gPf[s] = mk4(add3(f3(gPf[s]), accP), 0.f);
            gcl = (double)gmk * (double)dot3(gk, dv);
            if (gcSlot) gcSlot[s] = gcl;
        }
    }
    if (gcSlot) return;  // ensembles reduce per scene (k_adj_scene_sum)
    sh[threadIdx.x] = gcl;
    __syncthreads();
    for (int w = kBS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[tile] = sh[0];
}

// log2Scale (renormalised form): the incoming gradients are 2^-E of the true ones (E
// per scene), so the step's dL/dc is scaled back by 2^E (exact in double).
__global__ void k_adj_sum(const double* __restrict__ partial, uint32_t n, double* __restrict__ accum,
                          const int32_t* __restrict__ log2Scale) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double t = 0.0;
    for (uint32_t b = 0; b < n; ++b) t += partial[b];
    *accum += log2Scale ? ldexp(t, *log2Scale) : t;
}

// One block per scene: dL/dc of the scene's slots, summed in a fixed order.
__global__ __launch_bounds__(kBS) void k_adj_scene_sum(const double* __restrict__ gcSlot, Scenes sc,
                                                       double* __restrict__ accum,
                                                       const int32_t* __restrict__ log2Scale) {
    __shared__ double sh[kBS];
    const SceneParams& sp = sc.params[blockIdx.x];
    double t = 0.0;
    for (uint32_t i = threadIdx.x; i < sp.numParticles; i += kBS) t += gcSlot[sp.particleBase + i];
    sh[threadIdx.x] = t;
    __syncthreads();
    for (int w = kBS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) accum[blockIdx.x] += log2Scale ? ldexp(sh[0], log2Scale[blockIdx.x]) : sh[0];
}

__global__ __launch_bounds__(kBS) void k_adj_velocity(const uint32_t* __restrict__ idxS,
                                                      const uint32_t* __restrict__ finalClamp,
                                                      const float4* __restrict__ gPf, const float4* __restrict__ gVn,
                                                      const float4* __restrict__ U, sim::DeviceParams dp, float invDt,
                                                      uint32_t N, float4* __restrict__ gP, float4* __restrict__ A,
                                                      float4* __restrict__ gX) {
    const uint32_t s = blockIdx.x * kBS + threadIdx.x;
    if (s >= N) return;
    const float3 gv = f3(gVn[s]);
    const float3 g = mask3(add3(f3(gPf[s]), mul3(gv, invDt)), finalClamp[s]);
    gP[s] = mk4(g, 0.f);
    A[s] = mk4(disp_adjoint(g, U[s], dp), 0.f);
    gX[idxS[s]] = mk4(mul3(gv, -invDt), 0.f);
}

__global__ __launch_bounds__(kBS) __attribute__((amdgpu_waves_per_eu(7))) void k_adj_delta(const float4* __restrict__ P, const uint32_t* __restrict__ segOf,
                                                   SpanTable st, sim::DeviceParams dp, uint32_t N,
                                                   const float4* __restrict__ A, const float4* __restrict__ lamA,
                                                   const uint2* __restrict__ lamB, float4* __restrict__ gP,
                                                   float4* __restrict__ lamQ, uint2* __restrict__ lamE) {
    __shared__ uint32_t sQ[kQD * kBS];
    __shared__ float4 sBuf[kBS];
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t s = tile * kBS + threadIdx.x;
    const bool valid = s < N;
    const uint32_t ss = valid ? s : N - 1;
    const PairConsts c = make_pair_consts(dp);
    const float inv_w_q = (c.w_q > 0.f) ? 1.0f / c.w_q : 0.f;
    const float4 pk4 = P[ss];
    const float3 pk = f3(pk4), ak = f3(A[ss]);
    float gl = 0.f;
    float3 acc = make_float3(0.f, 0.f, 0.f);
    auto dist2 = [&](const float4& pj) {
        const float3 r = sub3(pk, f3(pj));
        return pair_r2(r.x, r.y, r.z);
    };
    with_pair_form(c, [&](auto form) {
        constexpr bool kFixed = decltype(form)::value;
        auto apply = [&](uint32_t j) {
            const float4 pj4 = P[j];
            const float3 r = sub3(pk, f3(pj4));
            const float r2 = pair_r2(r.x, r.y, r.z);
            const Spiky sp = spiky_of<kFixed>(r2, c);
            const float3 g = mul3(r, sp.f);
            const float3 da = sub3(ak, f3(A[j]));
            gl += dot3(g, da);
            // c_kj and its gradient (s_corr, reference: pbf_delta_apply.cu:77-84)
            float scorr = 0.f, dsr = 0.f;  // dscorr = dsr * r
            if (kFixed) {
                // the general branch below in the common setting (exponent 4, a floor)
                const float hr2 = c.h2 - r2;
                const float W = c.poly6 * hr2 * hr2 * hr2;
                const float ratio = div_unscaled(W, c.w_q, c.rcp_w_q);
                const float x2 = ratio * ratio;
                const float raw = c.neg_scorr_k * (x2 * x2);
                scorr = raw;
                if (raw < c.scorr_min) {
                    scorr = c.scorr_min;
                } else if (ratio > 0.f) {
                    const float pn1 = ratio * (ratio * ratio);
                    dsr = c.neg_scorr_k * c.scorr_n * pn1 * inv_w_q * (-6.0f * c.poly6 * hr2 * hr2);
                }
            } else if (c.scorr_enable) {
                const float hr2 = c.h2 - r2;
                const float W = c.poly6 * hr2 * hr2 * hr2;
                // the forward's ratio and s_corr, bit for bit
                const float ratio =
                    (c.w_q > 0.f) ? (c.unscaled ? div_unscaled(W, c.w_q, c.rcp_w_q) : (W / c.w_q)) : 0.0f;
                const float raw = c.neg_scorr_k * pow_scorr(ratio, c.scorr_n, c.scorr_ni);
                scorr = raw;
                if (c.scorr_min < 0.f && raw < c.scorr_min) {
                    scorr = c.scorr_min;
                } else if (c.w_q > 0.f && ratio > 0.f) {
                    // d/dr of -k (W/Wq)^n: -k n (W/Wq)^(n-1) / Wq * dW6, dW6 = -6 poly6 hr2^2 r
                    const float pn1 = pow_scorr_deriv(ratio, c);
                    dsr = c.neg_scorr_k * c.scorr_n * pn1 * inv_w_q * (-6.0f * c.poly6 * hr2 * hr2);
                }
            }
            const float ckj = pk4.w + pj4.w + scorr;
            acc = add3(acc, add3(mul3(hess_apply(sp, r, da), ckj), mul3(g, dsr * dot3(r, da))));
        };
        neighbor_walk<kBS>(valid, s, SpanRows(st, segOf[ss]), SortedFetch{SlotBuf(P, N + kSlotPad)},
                                  sBuf + (threadIdx.x & ~63u), sQ + threadIdx.x, 0, c.h2, dist2, apply);
    });
    if (!valid) return;
    gP[s] = mk4(add3(f3(gP[s]), acc), 0.f);
    const float4 la = lamA[s];
    const uint2 lb = lamB[s];
    lamQ[s] = make_float4(la.x, la.y, la.z, gl * la.w);
    lamE[s] = make_uint2(__float_as_uint(gl * __uint_as_float(lb.x)), lb.y);
}

// Lambda terms of dL/dP; then a of the iteration before (U = its recorded scaled
// correction), or nothing after the first iteration (U == nullptr).
__global__ __launch_bounds__(kBS) __attribute__((amdgpu_waves_per_eu(7))) void k_adj_lambda(const float4* __restrict__ P, const uint32_t* __restrict__ segOf,
                                                    SpanTable st, sim::DeviceParams dp, uint32_t N,
                                                    const float4* __restrict__ lamQ, const uint2* __restrict__ lamE,
                                                    const float4* __restrict__ U, float4* __restrict__ gP,
                                                    float4* __restrict__ A) {
    __shared__ uint32_t sQ[kQD * kBS];
    __shared__ float4 sBuf[kBS];
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t s = tile * kBS + threadIdx.x;
    const bool valid = s < N;
    const uint32_t ss = valid ? s : N - 1;
    const PairConsts c = make_pair_consts(dp);
    const float3 pk = f3(P[ss]);
    const float4 qk = lamQ[ss];
    const uint2 ek = lamE[ss];
    const float3 Gk = f3(qk);
    const float bk = qk.w, epk = __uint_as_float(ek.x);
    const uint32_t cutK = ek.y;
    const float poly6 = dp.kernel.poly6;
    const bool own = bk != 0.f || epk != 0.f;
    float3 acc = make_float3(0.f, 0.f, 0.f);
    auto dist2 = [&](const float4& pj) {
        const float3 r = sub3(pk, f3(pj));
        return pair_r2(r.x, r.y, r.z);
    };
    with_pair_form(c, [&](auto form) {
        auto apply = [&](uint32_t j) {
            const float3 r = sub3(pk, f3(P[j]));
            const float r2 = pair_r2(r.x, r.y, r.z);
            const float4 qj = lamQ[j];
            const uint2 ej = lamE[j];
            const float epj = __uint_as_float(ej.x);
            const bool inK = own && j <= cutK;                          // j in Ncap(k)
            const bool inJ = (qj.w != 0.f || epj != 0.f) && s <= ej.y;  // k in Ncap(j)
            if (!inK && !inJ) return;
            const float hr2 = c.h2 - r2;
            const float3 dW6 = mul3(r, -6.0f * poly6 * hr2 * hr2);
            const Spiky sp = spiky_of<decltype(form)::value>(r2, c);
            const float3 g = mul3(r, sp.f);
            if (inK) acc = add3(acc, add3(mul3(dW6, bk), mul3(hess_apply(sp, r, add3(g, Gk)), 2.0f * epk)));
            if (inJ) acc = add3(acc, add3(mul3(dW6, qj.w), mul3(hess_apply(sp, r, sub3(g, f3(qj))), 2.0f * epj)));
        };
        neighbor_walk<kBS>(valid, s, SpanRows(st, segOf[ss]), SortedFetch{SlotBuf(P, N + kSlotPad)},
                                  sBuf + (threadIdx.x & ~63u), sQ + threadIdx.x, 0, c.h2, dist2, apply);
    });
    if (!valid) return;
    const float3 g = add3(f3(gP[s]), acc);
    gP[s] = mk4(g, 0.f);
    if (U) A[s] = mk4(disp_adjoint(g, U[s], dp), 0.f);
}

// |x| as ordered bits: for non-negative floats the unsigned order of the bits is the
// order of the values, with NaN above +inf.
__device__ __forceinline__ uint32_t abs_bits(float x) { return __float_as_uint(x) & 0x7FFFFFFFu; }

// gmax (renormalised form): raised to the largest |component| this kernel writes, per
// scene.  A scene's particles fill one contiguous index range, the same for particle
// order and sorted slots, so sceneOf (particle order) also gives the scene of a slot.
__global__ __launch_bounds__(kBS) void k_adj_predict(const uint32_t* __restrict__ idxS,
                                                     const uint32_t* __restrict__ predClamp,
                                                     const float4* __restrict__ gP, float dt, uint32_t N,
                                                     float4* __restrict__ gX, float4* __restrict__ gV,
                                                     uint32_t* __restrict__ gmax,
                                                     const uint32_t* __restrict__ sceneOf) {
    const uint32_t s = blockIdx.x * kBS + threadIdx.x;
    uint32_t m = 0u;
    if (s < N) {
        const uint32_t i = idxS[s];
        const float3 g = mask3(f3(gP[s]), predClamp[i]);
        const float3 x = add3(f3(gX[i]), g), v = mul3(g, dt);
        gX[i] = mk4(x, 0.f);
        gV[i] = mk4(v, 0.f);
        m = max(max(max(abs_bits(x.x), abs_bits(x.y)), max(abs_bits(x.z), abs_bits(v.x))),
                max(abs_bits(v.y), abs_bits(v.z)));
    }
    if (!gmax) return;
    // lanes past N join the reductions with the wave's first scene and m = 0
    const uint32_t first = __builtin_amdgcn_readfirstlane(sceneOf && s < N ? sceneOf[s] : 0u);
    const uint32_t q = (sceneOf && s < N) ? sceneOf[s] : first;
    if (__all(q == first)) {
        m = wave_max_u32(m);
        if ((threadIdx.x & 63u) == 0 && m) atomicMax(gmax + first * kGmaxStride, m);
    } else if (m) {
        atomicMax(gmax + q * kGmaxStride, m);  // a wave across a scene boundary
    }
}

// Renormalised form: scale dL/dX_t and dL/dV_t of each scene by 2^-e so their largest
// magnitude lies in [0.5, 1), and add e to the scene's exponent the caller keeps.  e is
// the frexp exponent of the scene's maximum (k_adj_predict), limited to [-126, 126] so
// 2^-e is a normal float and every product is exact; a zero or non-finite maximum
// leaves the scene's gradients as they are.  Scenes are independent (no neighbour
// crosses a scene), so each keeps its own exponent and a small scene's gradients never
// sink below float32 beside a large one's.
__global__ __launch_bounds__(kBS) void k_adj_renorm(float4* __restrict__ gX, float4* __restrict__ gV, uint32_t N,
                                                    const uint32_t* __restrict__ gmax,
                                                    int32_t* __restrict__ log2Scale, Scenes sc) {
    const uint32_t i = blockIdx.x * kBS + threadIdx.x;
    if (i >= N) return;
    const uint32_t q = sc.params ? sc.sceneOf[i] : 0u;
    const uint32_t mb = gmax[q * kGmaxStride];
    if (mb == 0u || mb >= 0x7F800000u) return;
    const int e = min(126, max(-126, (int)(mb >> 23) - 126));
    if (e == 0) return;
    if (i == (sc.params ? sc.params[q].particleBase : 0u)) log2Scale[q] += e;
    const float f = __uint_as_float((uint32_t)(127 - e) << 23);
    gX[i] = mk4(mul3(f3(gX[i]), f), 0.f);
    gV[i] = mk4(mul3(f3(gV[i]), f), 0.f);
}

}  // namespace

void adjoint_step(const StepTape& t, const SpanTable& stIn, const sim::SimParams& p, const Scenes& sc,
                  const float4* gXin, const float4* gVin, float4* gXout, float4* gVout, double* gcAccum,
                  AdjointScratch& w, hipStream_t s, int32_t* log2Scale) {
    const uint32_t N = p.numParticles;
    if (N == 0) return;
    const int I = p.solverIters;
    if (I < 1 || I > kMaxTapeIters) {
        set_error(PBFX_E_INVALID_ARGUMENT, "adjoint_step: 1..8 solver iterations");
        return;
    }
    const uint32_t B = ceil_div(N, kBS);
    if (B > w.blocks) {
        set_error(PBFX_E_CAPACITY, "adjoint_step: scratch smaller than the particle count");
        return;
    }
    const sim::DeviceParams dp = sim::MakeDeviceParams(p);
    SpanTable st = stIn;
    st.spans = t.spans;
    const float4* Pf = t.P[I];
    if (log2Scale && !w.gmax) {
        set_error(PBFX_E_INVALID_ARGUMENT, "adjoint_step: renormalised form without its scratch");
        return;
    }
    uint32_t* gmax = log2Scale ? w.gmax : nullptr;
    const bool ens = sc.params != nullptr;
    k_adj_gather<<<B, kBS, 0, s>>>(t.idxS, gXin, gVin, t.xsphGm, w.gPf, w.gVo, N, gmax, ens ? sc.count : 1u);
    if (ens && (!t.sceneOfSlot || !w.gcSlot)) {
        set_error(PBFX_E_INVALID_ARGUMENT, "adjoint_step: ensemble tape without scene slots");
        return;
    }
    k_adj_xsph<<<B, kBS, 0, s>>>(Pf, t.Vs, t.segOf, st, dp, sc, t.sceneOfSlot, N, w.gVo, w.gVn, w.gPf,
                                 w.partial, ens ? w.gcSlot : nullptr);
    if (ens)
        k_adj_scene_sum<<<sc.count, kBS, 0, s>>>(w.gcSlot, sc, gcAccum, log2Scale);
    else
        k_adj_sum<<<1, 1, 0, s>>>(w.partial, B, gcAccum, log2Scale);
    k_adj_velocity<<<B, kBS, 0, s>>>(t.idxS, t.finalClamp, w.gPf, w.gVn, t.dispU[I - 1], dp, 1.0f / p.dt, N, w.gP,
                                     w.a, gXout);
    for (int it = I - 1; it >= 0; --it) {
        const float4* Pit = t.P[it];
        k_adj_delta<<<B, kBS, 0, s>>>(Pit, t.segOf, st, dp, N, w.a, t.lamA[it], t.lamB[it], w.gP, w.lamQ, w.lamE);
        k_adj_lambda<<<B, kBS, 0, s>>>(Pit, t.segOf, st, dp, N, w.lamQ, w.lamE, it > 0 ? t.dispU[it - 1] : nullptr,
                                       w.gP, w.a);
    }
    k_adj_predict<<<B, kBS, 0, s>>>(t.idxS, t.predClamp, w.gP, p.dt, N, gXout, gVout, gmax,
                                    ens ? sc.sceneOf : nullptr);
    if (log2Scale) k_adj_renorm<<<B, kBS, 0, s>>>(gXout, gVout, N, gmax, log2Scale, sc);
    check_launch("adjoint_step");
}

}  // namespace pbfx
