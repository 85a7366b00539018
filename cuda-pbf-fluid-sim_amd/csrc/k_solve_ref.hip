// k_solve_ref.hip — drop-in lambda / delta-apply / XSPH launchers over the
// reference's data layout: float4 AoS state in original particle order, reached
// through the sorted index array.
//
// One lane per sorted slot.  Each of the (2*reach+1)^2 neighbour rows is one
// contiguous span of sorted slots (its cells have consecutive keys), found with
// at most 2*(2*reach+1) range loads (dense) or one lower_bound (compact), and
// walked in the reference's z -> y -> x, ascending-slot order, so sums are
// formed in the same order as the reference's per-cell loops.  The walk is the
// cell-reordered kernels' wave walk (walk.hpp: unrolled per-lane candidate loads,
// in-range pairs deferred through a per-lane FIFO), with each candidate's position
// read through the sorted index array.
// The headless pipeline uses the cell-reordered kernels in k_solve_sorted.hip;
// these exist so the sim/ phase API can be served unchanged.
#include "pbf_common.hpp"
#include "pbf_hip.h"
#include "pbf_math.hpp"
#include "walk.hpp"

namespace pbfx {
namespace {

using namespace walk;

// Neighbour rows of the cell `key`, computed from the cell lookup the launcher is
// handed (dense cellStart/cellEnd or the compact unique keys): row r = (dz, dy) is one
// contiguous span of sorted slots because its cells have consecutive keys.
template <class Lookup>
struct CellRows {
    Lookup lookup;
    sim::GridBounds g;
    int reach, side, cx, cy, cz, x0, x1;
    // the drop-in launchers take no tuning: lanes walk their own spans, and a row
    // is streamed through LDS only when every lane of the wave shares it
    uint32_t streamRatio = 1, chunkBits = 0;
    __device__ __forceinline__ CellRows(const Lookup& l, const sim::GridBounds& grid, int R, uint32_t key)
        : lookup(l), g(grid), reach(R), side(2 * R + 1) {
        decode_key(key, g.dim, cx, cy, cz);
        x0 = max(0, cx - R);
        x1 = min(g.dim.x - 1, cx + R);
    }
    __device__ __forceinline__ int count() const { return side * side; }
    __device__ __forceinline__ void span(int r, uint32_t& lo, uint32_t& hi) const {
        const int z = cz + r / side - reach, y = cy + r % side - reach;
        lo = hi = 0u;
        if (z < 0 || z >= g.dim.z || y < 0 || y >= g.dim.y || x0 > x1) return;
        const uint32_t rowBase = (uint32_t)((z * g.dim.y + y) * g.dim.x);
        lookup.span(rowBase + (uint32_t)x0, rowBase + (uint32_t)x1, lo, hi);
    }
};

// Positions in original particle order, reached through the sorted index array
// (which has no slots past the last particle); a particle's id is its original index.
struct IndexedFetch {
    static constexpr bool kPadded = false;
    const float4* P;
    const uint32_t* idxS;
    __device__ __forceinline__ uint32_t id(uint32_t k) const { return idxS[k]; }
    __device__ __forceinline__ float4 at(uint32_t j) const { return P[j]; }
};

// Every kernel below runs one lane per sorted slot and walks the slot's neighbours
// with the shared wave walk; lanes past N take part in the walk's wave-level steps
// with no rows.  The neighbour kernels deal their 256-slot tiles to the XCDs in
// contiguous runs (xcd_tile), so a tile and its neighbours' tiles share an L2: 12%
// faster drop-in steps on the 1M dam break.
template <class Lookup>
__global__ __launch_bounds__(kBS) void k_lambda_ref(float* __restrict__ lambda, const float4* __restrict__ P,
                                                    const uint32_t* __restrict__ idxS,
                                                    const uint32_t* __restrict__ keyS, Lookup lookup,
                                                    sim::DeviceParams dp, uint32_t N) {
    __shared__ uint32_t sQ[kQD * kBS];
    __shared__ float4 sBuf[kBS];
    const uint32_t s = xcd_tile(blockIdx.x, gridDim.x) * kBS + threadIdx.x;
    const bool valid = s < N;
    const uint32_t ss = valid ? s : N - 1;
    const uint32_t i = idxS[ss];
    const float4 pi = P[i];
    const PairConsts c = make_pair_consts(dp);
    LambdaAcc a;
    auto dist2 = [&](const float4& pj) {
        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        return rx * rx + ry * ry + rz * rz;
    };
    const CellRows<Lookup> rows(lookup, dp.grid, reach_of(c.h, dp.grid.cellSize), keyS[ss]);
    with_pair_form(c, [&](auto form) {
        auto apply = [&](uint32_t j) {
            const float4 pj = P[j];
            const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
            lambda_pair<decltype(form)::value>(a, rx, ry, rz, rx * rx + ry * ry + rz * rz, c);
        };
        neighbor_walk<false, kBS>(valid, s, rows, CandMask{}, IndexedFetch{P, idxS}, sBuf + (threadIdx.x & ~63u),
                                  sQ + threadIdx.x, dp.maxNeighbors, c.h2, dist2, apply);
    });
    if (valid) lambda[i] = lambda_finish(a, dp);
}

// Held to 72 VGPRs (seven waves per SIMD); at its natural 77 it ran six, and the
// extra wave made the drop-in Δ pass 7% faster on the 1M dam break.
template <class Lookup>
__global__ __launch_bounds__(kBS) __attribute__((amdgpu_waves_per_eu(7))) void k_delta_ref(float4* __restrict__ delta, const float4* __restrict__ P,
                                                   const float* __restrict__ lambda,
                                                   const uint32_t* __restrict__ idxS,
                                                   const uint32_t* __restrict__ keyS, Lookup lookup,
                                                   sim::DeviceParams dp, uint32_t N) {
    __shared__ uint32_t sQ[kQD * kBS];
    __shared__ float4 sBuf[kBS];
    const uint32_t s = xcd_tile(blockIdx.x, gridDim.x) * kBS + threadIdx.x;
    const bool valid = s < N;
    const uint32_t ss = valid ? s : N - 1;
    const uint32_t i = idxS[ss];
    const float4 pi = P[i];
    const float lam_i = lambda[i];
    const PairConsts c = make_pair_consts(dp);
    float dx = 0.f, dy = 0.f, dz = 0.f;
    auto dist2 = [&](const float4& pj) {
        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        return rx * rx + ry * ry + rz * rz;
    };
    const CellRows<Lookup> rows(lookup, dp.grid, reach_of(c.h, dp.grid.cellSize), keyS[ss]);
    with_pair_form(c, [&](auto form) {
        auto apply = [&](uint32_t j) {
            const float4 pj = P[j];
            const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
            delta_pair<decltype(form)::value>(dx, dy, dz, rx, ry, rz, rx * rx + ry * ry + rz * rz, lam_i, lambda[j],
                                              c);
        };
        neighbor_walk<false, kBS>(valid, s, rows, CandMask{}, IndexedFetch{P, idxS}, sBuf + (threadIdx.x & ~63u),
                                  sQ + threadIdx.x, 0, c.h2, dist2, apply);
    });
    if (!valid) return;
    const float3 d = delta_finish(dx, dy, dz, dp);
    delta[i] = make_float4(d.x, d.y, d.z, 0.0f);
}

// P[i] += delta[i] over sorted slots (reference: pbf_delta_apply.cu:223-235).
__global__ __launch_bounds__(kBS) void k_apply_delta(float4* __restrict__ P, const float4* __restrict__ delta,
                                                     const uint32_t* __restrict__ idxS, uint32_t N) {
    const uint32_t s = blockIdx.x * kBS + threadIdx.x;
    if (s >= N) return;
    const uint32_t i = idxS[s];
    const float4 d = delta[i];
    const float4 b = P[i];
    P[i] = make_float4(b.x + d.x, b.y + d.y, b.z + d.z, b.w);
}

template <class Lookup>
__global__ __launch_bounds__(kBS) void k_xsph_ref(float4* __restrict__ Vout, const float4* __restrict__ Vin,
                                                  const float4* __restrict__ P, const uint32_t* __restrict__ idxS,
                                                  const uint32_t* __restrict__ keyS, Lookup lookup,
                                                  sim::DeviceParams dp, uint32_t N) {
    __shared__ uint32_t sQ[kQD * kBS];
    __shared__ float4 sBuf[kBS];
    const uint32_t s = xcd_tile(blockIdx.x, gridDim.x) * kBS + threadIdx.x;
    const bool valid = s < N;
    const uint32_t ss = valid ? s : N - 1;
    const uint32_t i = idxS[ss];
    const float4 pi = P[i];
    const float4 vi = Vin[i];
    const float h2 = dp.kernel.h2, poly6 = dp.kernel.poly6;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    int n = 0;
    auto dist2 = [&](const float4& pj) {
        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        return rx * rx + ry * ry + rz * rz;
    };
    auto apply = [&](uint32_t j) {
        const float4 pj = P[j];
        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        const float hr2 = h2 - (rx * rx + ry * ry + rz * rz);
        const float w = poly6 * hr2 * hr2 * hr2;
        const float4 vj = Vin[j];
        sx += (vj.x - vi.x) * w;
        sy += (vj.y - vi.y) * w;
        sz += (vj.z - vi.z) * w;
        ++n;
    };
    const CellRows<Lookup> rows(lookup, dp.grid, reach_of(dp.kernel.h, dp.grid.cellSize), keyS[ss]);
    neighbor_walk<false, kBS>(valid, s, rows, CandMask{}, IndexedFetch{P, idxS}, sBuf + (threadIdx.x & ~63u),
                         sQ + threadIdx.x, 0, h2, dist2, apply);
    if (!valid) return;
    const float3 v = xsph_finish(vi.x, vi.y, vi.z, sx, sy, sz, n, dp, dp.xsph_c);
    Vout[i] = make_float4(v.x, v.y, v.z, 0.0f);
}

// Compact variants read M on the device (the count is produced on the stream).
struct CompactRowsDev {
    const uint32_t* __restrict__ uniq;
    const uint32_t* __restrict__ off;
    const uint32_t* __restrict__ count;
    __device__ __forceinline__ void span(uint32_t k0, uint32_t k1, uint32_t& lo, uint32_t& hi) const {
        CompactRows r{uniq, off, *count};
        r.span(k0, k1, lo, hi);
    }
};

bool check_dims(const sim::DeviceParams& dp, const char* where) {
    if (dp.grid.dim.x <= 0 || dp.grid.dim.y <= 0 || dp.grid.dim.z <= 0) {
        set_error(PBFX_E_INVALID_ARGUMENT, where);
        return false;
    }
    return true;
}

}  // namespace
}  // namespace pbfx

using namespace pbfx;

extern "C" void LaunchLambda(float* lambda, const float4* pos_pred, const uint32_t* indicesSorted,
                             const uint32_t* keysSorted, const uint32_t* cellStart, const uint32_t* cellEnd,
                             sim::DeviceParams dp, uint32_t N, hipStream_t s) {
    if (N == 0 || !check_dims(dp, "LaunchLambda: empty grid")) return;
    k_lambda_ref<DenseRows><<<ceil_div(N, kBS), kBS, 0, s>>>(lambda, pos_pred, indicesSorted, keysSorted,
                                                            DenseRows{cellStart, cellEnd}, dp, N);
    check_launch("LaunchLambda");
}

extern "C" void LaunchLambdaCompact(float* lambda, const float4* pos_pred, const uint32_t* indicesSorted,
                                    const uint32_t* keysSorted, const uint32_t* uniqueKeys, const uint32_t* offsets,
                                    const uint32_t* compactCount, sim::DeviceParams dp, uint32_t N, hipStream_t s) {
    if (N == 0 || !check_dims(dp, "LaunchLambdaCompact: empty grid")) return;
    k_lambda_ref<CompactRowsDev><<<ceil_div(N, kBS), kBS, 0, s>>>(
        lambda, pos_pred, indicesSorted, keysSorted, CompactRowsDev{uniqueKeys, offsets, compactCount}, dp, N);
    check_launch("LaunchLambdaCompact");
}

extern "C" void LaunchDeltaApply(float4* pos_pred, float4* delta, const float* lambda, const uint32_t* indicesSorted,
                                 const uint32_t* keysSorted, const uint32_t* cellStart, const uint32_t* cellEnd,
                                 sim::DeviceParams dp, uint32_t N, hipStream_t s) {
    if (N == 0 || !check_dims(dp, "LaunchDeltaApply: empty grid")) return;
    k_delta_ref<DenseRows><<<ceil_div(N, kBS), kBS, 0, s>>>(delta, pos_pred, lambda, indicesSorted, keysSorted,
                                                           DenseRows{cellStart, cellEnd}, dp, N);
    k_apply_delta<<<ceil_div(N, kBS), kBS, 0, s>>>(pos_pred, delta, indicesSorted, N);
    check_launch("LaunchDeltaApply");
}

extern "C" void LaunchDeltaApplyCompact(float4* pos_pred, float4* delta, const float* lambda,
                                        const uint32_t* indicesSorted, const uint32_t* keysSorted,
                                        const uint32_t* uniqueKeys, const uint32_t* offsets,
                                        const uint32_t* compactCount, sim::DeviceParams dp, uint32_t N,
                                        hipStream_t s) {
    if (N == 0 || !check_dims(dp, "LaunchDeltaApplyCompact: empty grid")) return;
    k_delta_ref<CompactRowsDev><<<ceil_div(N, kBS), kBS, 0, s>>>(delta, pos_pred, lambda, indicesSorted,
                                                                keysSorted,
                                                                CompactRowsDev{uniqueKeys, offsets, compactCount},
                                                                dp, N);
    k_apply_delta<<<ceil_div(N, kBS), kBS, 0, s>>>(pos_pred, delta, indicesSorted, N);
    check_launch("LaunchDeltaApplyCompact");
}

extern "C" void LaunchXSPH(float4* vel_out, const float4* vel_in, const float4* pos_pred,
                           const uint32_t* indicesSorted, const uint32_t* keysSorted, const uint32_t* cellStart,
                           const uint32_t* cellEnd, sim::DeviceParams dp, uint32_t N, hipStream_t s) {
    if (N == 0 || dp.xsph_c <= 0.f) return;
    if (vel_out == vel_in) { set_error(PBFX_E_ALIAS, "LaunchXSPH: vel_out aliases vel_in"); return; }
    if (!check_dims(dp, "LaunchXSPH: empty grid")) return;
    k_xsph_ref<DenseRows><<<ceil_div(N, kBS), kBS, 0, s>>>(vel_out, vel_in, pos_pred, indicesSorted, keysSorted,
                                                          DenseRows{cellStart, cellEnd}, dp, N);
    check_launch("LaunchXSPH");
}

extern "C" void LaunchXSPHCompact(float4* vel_out, const float4* vel_in, const float4* pos_pred,
                                  const uint32_t* indicesSorted, const uint32_t* keysSorted,
                                  const uint32_t* uniqueKeys, const uint32_t* offsets, const uint32_t* compactCount,
                                  sim::DeviceParams dp, uint32_t N, hipStream_t s) {
    if (N == 0 || dp.xsph_c <= 0.f) return;
    if (vel_out == vel_in) { set_error(PBFX_E_ALIAS, "LaunchXSPHCompact: vel_out aliases vel_in"); return; }
    if (!check_dims(dp, "LaunchXSPHCompact: empty grid")) return;
    k_xsph_ref<CompactRowsDev><<<ceil_div(N, kBS), kBS, 0, s>>>(vel_out, vel_in, pos_pred, indicesSorted,
                                                               keysSorted,
                                                               CompactRowsDev{uniqueKeys, offsets, compactCount},
                                                               dp, N);
    check_launch("LaunchXSPHCompact");
}
