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
// read through the sorted index array, or, when staging buffers are reserved for the
// launch stream (EnsureDropInStaging), from a cell-sorted copy the launcher gathers
// first.
// The headless pipeline uses the cell-reordered kernels in k_solve_sorted.hip;
// these exist so the sim/ phase API can be served unchanged.
#include <mutex>
#include <unordered_map>

#include "grid.hpp"
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
// Lanes per block of the drop-in neighbour kernels (the staging and apply kernels run
// kBS): on the collapsed 1M dam break, blocks of 128 lanes (8 KB of pair queue each)
// ran the neighbour walks faster than blocks of 256.
constexpr int kWalkBS = 128;

struct IndexedFetch {
    static constexpr bool kPadded = false;
    const float4* P;
    const uint32_t* idxS;
    __device__ __forceinline__ uint32_t id(uint32_t k) const { return idxS[k]; }
    __device__ __forceinline__ float4 at(uint32_t j) const { return P[j]; }
};

// Every kernel below runs one lane per sorted slot and walks the slot's neighbours
// with the shared wave walk (the staged XSPH with the queue-free inline walk); lanes
// past N take part in the walk's wave-level steps with no rows.  The neighbour kernels
// run kWalkBS lanes a block and deal their tiles to the XCDs in contiguous runs
// (xcd_tile), so a tile and its neighbours' tiles share an L2: 12% faster drop-in
// steps on the 1M dam break.
//
// kStaged: the launcher has gathered the inputs into cell-sorted staging buffers (Ps:
// position with lambda or 0 in w; Vs: velocity), so the walk reads candidates there
// in slot order (SortedFetch, queueing slots) instead of through the sorted index
// array.  The pairs, their order and their arithmetic are the same either way.
template <class Lookup, bool kStaged>
__global__ __launch_bounds__(kWalkBS) void k_lambda_ref(float* __restrict__ lambda, const float4* __restrict__ P,
                                                    const float4* __restrict__ Ps,
                                                    const uint32_t* __restrict__ idxS,
                                                    const uint32_t* __restrict__ keyS, Lookup lookup,
                                                    sim::DeviceParams dp, uint32_t N) {
    __shared__ uint32_t sQ[kQD * kWalkBS];
    __shared__ float4 sBuf[kWalkBS];
    const uint32_t s = xcd_tile(blockIdx.x, gridDim.x) * kWalkBS + threadIdx.x;
    const bool valid = s < N;
    const uint32_t ss = valid ? s : N - 1;
    const uint32_t i = idxS[ss];
    const float4 pi = kStaged ? Ps[ss] : P[i];
    const PairConsts c = make_pair_consts(dp);
    LambdaAcc a;
    auto dist2 = [&](const float4& pj) {
        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        return pair_r2(rx, ry, rz);
    };
    const CellRows<Lookup> rows(lookup, dp.grid, reach_of(c.h, dp.grid.cellSize), keyS[ss]);
    auto walk = [&](const auto& fetch) {
        with_pair_form(c, [&](auto form) {
            auto apply = [&](uint32_t j) {
                const float4 pj = fetch.at(j);
                const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
                lambda_pair<decltype(form)::value>(a, rx, ry, rz, pair_r2(rx, ry, rz), c);
            };
            neighbor_walk<kWalkBS>(valid, s, rows, fetch, sBuf + (threadIdx.x & ~63u),
                                      sQ + threadIdx.x, dp.maxNeighbors, c.h2, dist2, apply);
        });
    };
    if (kStaged)
        walk(SortedFetch{SlotBuf(Ps, N + kSlotPad)});
    else
        walk(IndexedFetch{P, idxS});
    if (valid) lambda[i] = lambda_finish(a, dp);
}

// Held to 72 VGPRs (seven waves per SIMD); at its natural 77 it ran six, and the
// extra wave made the drop-in Δ pass 7% faster on the 1M dam break.
template <class Lookup, bool kStaged>
__global__ __launch_bounds__(kWalkBS) __attribute__((amdgpu_waves_per_eu(7))) void k_delta_ref(
    float4* __restrict__ delta, const float4* __restrict__ P, const float4* __restrict__ Ps,
    const float* __restrict__ lambda, const uint32_t* __restrict__ idxS, const uint32_t* __restrict__ keyS,
    Lookup lookup, sim::DeviceParams dp, uint32_t N) {
    __shared__ uint32_t sQ[kQD * kWalkBS];
    __shared__ float4 sBuf[kWalkBS];
    const uint32_t s = xcd_tile(blockIdx.x, gridDim.x) * kWalkBS + threadIdx.x;
    const bool valid = s < N;
    const uint32_t ss = valid ? s : N - 1;
    const uint32_t i = idxS[ss];
    const float4 pi = kStaged ? Ps[ss] : P[i];
    const float lam_i = kStaged ? pi.w : lambda[i];
    const PairConsts c = make_pair_consts(dp);
    float dx = 0.f, dy = 0.f, dz = 0.f;
    auto dist2 = [&](const float4& pj) {
        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        return pair_r2(rx, ry, rz);
    };
    const CellRows<Lookup> rows(lookup, dp.grid, reach_of(c.h, dp.grid.cellSize), keyS[ss]);
    auto walk = [&](const auto& fetch) {
        with_pair_form(c, [&](auto form) {
            auto apply = [&](uint32_t j) {
                const float4 pj = fetch.at(j);
                const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
                const float lam_j = kStaged ? pj.w : lambda[j];
                delta_pair<decltype(form)::value>(dx, dy, dz, rx, ry, rz, pair_r2(rx, ry, rz), lam_i,
                                                  lam_j, c);
            };
            neighbor_walk<kWalkBS>(valid, s, rows, fetch, sBuf + (threadIdx.x & ~63u),
                                      sQ + threadIdx.x, 0, c.h2, dist2, apply);
        });
    };
    if (kStaged)
        walk(SortedFetch{SlotBuf(Ps, N + kSlotPad)});
    else
        walk(IndexedFetch{P, idxS});
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

template <class Lookup, bool kStaged>
__global__ __launch_bounds__(kWalkBS) void k_xsph_ref(float4* __restrict__ Vout, const float4* __restrict__ Vin,
                                                  const float4* __restrict__ P, const float4* __restrict__ Ps,
                                                  const float4* __restrict__ Vs, const uint32_t* __restrict__ idxS,
                                                  const uint32_t* __restrict__ keyS, Lookup lookup,
                                                  sim::DeviceParams dp, uint32_t N) {
    const uint32_t s = xcd_tile(blockIdx.x, gridDim.x) * kWalkBS + threadIdx.x;
    const bool valid = s < N;
    const uint32_t ss = valid ? s : N - 1;
    const uint32_t i = idxS[ss];
    const float4 pi = kStaged ? Ps[ss] : P[i];
    const float4 vi = kStaged ? Vs[ss] : Vin[i];
    const float h2 = dp.kernel.h2, poly6 = dp.kernel.poly6;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    int n = 0;
    auto dist2 = [&](const float4& pj) {
        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        return pair_r2(rx, ry, rz);
    };
    const CellRows<Lookup> rows(lookup, dp.grid, reach_of(dp.kernel.h, dp.grid.cellSize), keyS[ss]);
    if constexpr (kStaged) {
        // the staged copies are in slot order: XSPH's walk evaluates every candidate in
        // place, as the cell-reordered kernel does (no queue)
        const SlotBuf PS(Ps, N + kSlotPad), VS(Vs, N + kSlotPad);
        inline_walk(
            valid, s, rows, h2, [&](uint32_t k, uint32_t u) { return PS.at(k, u); }, dist2,
            [&](uint32_t k, uint32_t u, bool in) { return VS.at(in ? k : kSlotBufBeyond, u); },
            [&](const float4&, const float4& vj, float r2, bool in) {
                const float hr2 = h2 - r2;
                const float w = poly6 * hr2 * hr2 * hr2;
                float tx = sx, ty = sy, tz = sz;
                xsph_pair(tx, ty, tz, vj, vi, w);
                sx = in ? tx : sx;
                sy = in ? ty : sy;
                sz = in ? tz : sz;
                n += in ? 1 : 0;
            });
    } else {
        __shared__ uint32_t sQ[kQD * kWalkBS];
        __shared__ float4 sBuf[kWalkBS];
        auto apply = [&](uint32_t j) {
            const float4 pj = P[j];
            const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
            const float hr2 = h2 - pair_r2(rx, ry, rz);
            const float w = poly6 * hr2 * hr2 * hr2;
            const float4 vj = Vin[j];
            xsph_pair(sx, sy, sz, vj, vi, w);
            ++n;
        };
        neighbor_walk<kWalkBS>(valid, s, rows, IndexedFetch{P, idxS}, sBuf + (threadIdx.x & ~63u),
                                  sQ + threadIdx.x, 0, h2, dist2, apply);
    }
    if (!valid) return;
    const float3 v = xsph_finish(vi.x, vi.y, vi.z, sx, sy, sz, n, dp, dp.xsph_c);
    Vout[i] = make_float4(v.x, v.y, v.z, 0.0f);
}

// Staged XSPH, two particles of one cell a lane (the slot pairs of build_slot_pairs):
// they share the cell's neighbour rows, so a round's candidate positions are read once
// for both and a candidate's velocity once when it is a neighbour of either; each
// particle's sum takes its own neighbours in walk order, as in k_xsph_ref.  Blocks
// past the pair count return at once; the XCD-aware order is taken over the others.
template <class Lookup>
__global__ __launch_bounds__(kWalkBS) void k_xsph_ref_pairs(float4* __restrict__ Vout, const float4* __restrict__ Ps,
                                                        const float4* __restrict__ Vs,
                                                        const uint32_t* __restrict__ idxS,
                                                        const uint32_t* __restrict__ keyS, Lookup lookup,
                                                        sim::DeviceParams dp, uint32_t N,
                                                        const uint2* __restrict__ pairs,
                                                        const uint32_t* __restrict__ pairCount) {
    const uint32_t np = *pairCount;
    const uint32_t nb = (np + (uint32_t)kWalkBS - 1u) / (uint32_t)kWalkBS;
    if (blockIdx.x >= nb) return;
    const uint32_t p = xcd_tile(blockIdx.x, nb) * kWalkBS + threadIdx.x;
    if (p >= np) return;
    const uint2 ab = pairs[p];
    const uint32_t a = ab.x;
    const bool hasB = ab.y != 0xFFFFFFFFu;
    const uint32_t b = hasB ? ab.y : a;
    const float4 pa = Ps[a], pb = Ps[b], va = Vs[a], vb = Vs[b];
    const float h2 = dp.kernel.h2, poly6 = dp.kernel.poly6;
    float sax = 0.f, say = 0.f, saz = 0.f, sbx = 0.f, sby = 0.f, sbz = 0.f;
    int na = 0, nbr = 0;
    const CellRows<Lookup> rows(lookup, dp.grid, reach_of(dp.kernel.h, dp.grid.cellSize), keyS[a]);
    const SlotBuf PS(Ps, N + kSlotPad), VS(Vs, N + kSlotPad);
    inline_walk_pair(
        a, b, hasB, rows, h2, [&](uint32_t k, uint32_t u) { return PS.at(k, u); },
        [&](const float4& pj) { return pair_r2(pa.x - pj.x, pa.y - pj.y, pa.z - pj.z); },
        [&](const float4& pj) { return pair_r2(pb.x - pj.x, pb.y - pj.y, pb.z - pj.z); },
        [&](uint32_t k, uint32_t u, bool in) { return VS.at(in ? k : kSlotBufBeyond, u); },
        [&](const float4& vj, float r2, bool in) {
            const float hr2 = h2 - r2;
            const float w = poly6 * hr2 * hr2 * hr2;
            float tx = sax, ty = say, tz = saz;
            xsph_pair(tx, ty, tz, vj, va, w);
            sax = in ? tx : sax;
            say = in ? ty : say;
            saz = in ? tz : saz;
            na += in ? 1 : 0;
        },
        [&](const float4& vj, float r2, bool in) {
            const float hr2 = h2 - r2;
            const float w = poly6 * hr2 * hr2 * hr2;
            float tx = sbx, ty = sby, tz = sbz;
            xsph_pair(tx, ty, tz, vj, vb, w);
            sbx = in ? tx : sbx;
            sby = in ? ty : sby;
            sbz = in ? tz : sbz;
            nbr += in ? 1 : 0;
        });
    const float3 u = xsph_finish(va.x, va.y, va.z, sax, say, saz, na, dp, dp.xsph_c);
    Vout[idxS[a]] = make_float4(u.x, u.y, u.z, 0.0f);
    if (hasB) {
        const float3 v = xsph_finish(vb.x, vb.y, vb.z, sbx, sby, sbz, nbr, dp, dp.xsph_c);
        Vout[idxS[b]] = make_float4(v.x, v.y, v.z, 0.0f);
    }
}

// Staging gather: Ps[s] = (P[i].xyz, lambda[i] or 0), Vs[s] = V[i] (when V), i = idxS[s].
__global__ __launch_bounds__(kBS) void k_stage(float4* __restrict__ Ps, float4* __restrict__ Vs,
                                               const float4* __restrict__ P, const float* __restrict__ lambda,
                                               const float4* __restrict__ V, const uint32_t* __restrict__ idxS,
                                               uint32_t N) {
    const uint32_t s = blockIdx.x * kBS + threadIdx.x;
    if (s >= N) return;
    const uint32_t i = idxS[s];
    const float4 p = P[i];
    Ps[s] = make_float4(p.x, p.y, p.z, lambda ? lambda[i] : 0.0f);
    if (V) Vs[s] = V[i];
}

// Staging buffers of one stream, sized by EnsureDropInStaging: N + kSlotPad slots each,
// the pad zeroed once (the walk's padded loads may read it); and what the staged XSPH
// builds its slot pairs in (the cells of the sorted keys, then their slots two at a
// time).
struct Staging {
    float4* P = nullptr;
    float4* V = nullptr;
    uint32_t* uniq = nullptr;   // N cell keys
    uint32_t* off = nullptr;    // N + 1 cell offsets
    uint32_t* count = nullptr;  // cells
    uint32_t* segOf = nullptr;  // N: cell of each slot
    uint2* pairs = nullptr;     // N slot pairs
    uint32_t* pairCount = nullptr;
    uint32_t* tileCounts = nullptr;   // compact_scratch_elems(N) each
    uint32_t* tileOffsets = nullptr;
    uint32_t capacity = 0;
};

void free_staging(Staging& st) {
    for (void* p : {(void*)st.P, (void*)st.V, (void*)st.uniq, (void*)st.off, (void*)st.count, (void*)st.segOf,
                    (void*)st.pairs, (void*)st.pairCount, (void*)st.tileCounts, (void*)st.tileOffsets})
        if (p) (void)hipFree(p);
    st = Staging{};
}
std::mutex g_stage_mu;
std::unordered_map<hipStream_t, Staging> g_stage;

// The stream's staging buffers when they hold N slots, else nullptr (read through the
// sorted index array).
bool staging_for(hipStream_t s, uint32_t N, Staging& out) {
    std::lock_guard<std::mutex> lk(g_stage_mu);
    const auto it = g_stage.find(s);
    if (it == g_stage.end() || it->second.capacity < N) return false;
    out = it->second;
    return true;
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

namespace pbfx {
namespace {

template <class Lookup>
void launch_lambda(float* lambda, const float4* P, const uint32_t* idxS, const uint32_t* keyS, const Lookup& lk,
                   const sim::DeviceParams& dp, uint32_t N, hipStream_t s) {
    Staging st;
    if (staging_for(s, N, st)) {
        k_stage<<<ceil_div(N, kBS), kBS, 0, s>>>(st.P, nullptr, P, nullptr, nullptr, idxS, N);
        k_lambda_ref<Lookup, true><<<ceil_div(N, kWalkBS), kWalkBS, 0, s>>>(lambda, P, st.P, idxS, keyS, lk, dp, N);
    } else {
        k_lambda_ref<Lookup, false><<<ceil_div(N, kWalkBS), kWalkBS, 0, s>>>(lambda, P, nullptr, idxS, keyS, lk, dp,
                                                                             N);
    }
}

template <class Lookup>
void launch_delta(float4* P, float4* delta, const float* lambda, const uint32_t* idxS, const uint32_t* keyS,
                  const Lookup& lk, const sim::DeviceParams& dp, uint32_t N, hipStream_t s) {
    Staging st;
    if (staging_for(s, N, st)) {
        k_stage<<<ceil_div(N, kBS), kBS, 0, s>>>(st.P, nullptr, P, lambda, nullptr, idxS, N);
        k_delta_ref<Lookup, true><<<ceil_div(N, kWalkBS), kWalkBS, 0, s>>>(delta, P, st.P, lambda, idxS, keyS, lk,
                                                                           dp, N);
    } else {
        k_delta_ref<Lookup, false><<<ceil_div(N, kWalkBS), kWalkBS, 0, s>>>(delta, P, nullptr, lambda, idxS, keyS,
                                                                            lk, dp, N);
    }
    k_apply_delta<<<ceil_div(N, kBS), kBS, 0, s>>>(P, delta, idxS, N);
}

template <class Lookup>
void launch_xsph(float4* Vout, const float4* Vin, const float4* P, const uint32_t* idxS, const uint32_t* keyS,
                 const Lookup& lk, const sim::DeviceParams& dp, uint32_t N, hipStream_t s) {
    Staging st;
    if (staging_for(s, N, st)) {
        k_stage<<<ceil_div(N, kBS), kBS, 0, s>>>(st.P, st.V, P, nullptr, Vin, idxS, N);
        compact_segments(st.uniq, st.off, st.count, st.segOf, keyS, N, st.tileCounts, st.tileOffsets, s);
        build_slot_pairs(st.pairs, st.pairCount, st.segOf, st.off, N, st.tileCounts, st.tileOffsets, s);
        k_xsph_ref_pairs<Lookup><<<ceil_div(N, kWalkBS), kWalkBS, 0, s>>>(Vout, st.P, st.V, idxS, keyS, lk, dp, N, st.pairs,
                                                                  st.pairCount);
    } else {
        k_xsph_ref<Lookup, false><<<ceil_div(N, kWalkBS), kWalkBS, 0, s>>>(Vout, Vin, P, nullptr, nullptr, idxS, keyS, lk,
                                                                    dp, N);
    }
}

}  // namespace
}  // namespace pbfx

extern "C" bool EnsureDropInStaging(uint32_t N, hipStream_t stream) {
    std::lock_guard<std::mutex> lk(g_stage_mu);
    Staging& st = g_stage[stream];
    if (N <= st.capacity) return reserve_stream_compact_scratch(N, stream);
    free_staging(st);
    const size_t slots = (size_t)N + kSlotPad;
    const size_t tiles = compact_scratch_elems(N);
    // the pad is cleared on the stream the launchers use, then waited for
    if (hipMalloc(&st.P, sizeof(float4) * slots) != hipSuccess ||
        hipMalloc(&st.V, sizeof(float4) * slots) != hipSuccess ||
        hipMalloc(&st.uniq, sizeof(uint32_t) * N) != hipSuccess ||
        hipMalloc(&st.off, sizeof(uint32_t) * ((size_t)N + 1)) != hipSuccess ||
        hipMalloc(&st.count, sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&st.segOf, sizeof(uint32_t) * N) != hipSuccess ||
        hipMalloc(&st.pairs, sizeof(uint2) * N) != hipSuccess ||
        hipMalloc(&st.pairCount, sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&st.tileCounts, sizeof(uint32_t) * tiles) != hipSuccess ||
        hipMalloc(&st.tileOffsets, sizeof(uint32_t) * tiles) != hipSuccess ||
        hipMemsetAsync(st.P, 0, sizeof(float4) * slots, stream) != hipSuccess ||
        hipMemsetAsync(st.V, 0, sizeof(float4) * slots, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
        free_staging(st);
        g_stage.erase(stream);
        set_error(PBFX_E_CAPACITY, "EnsureDropInStaging: allocation failed");
        return false;
    }
    st.capacity = N;
    return reserve_stream_compact_scratch(N, stream);
}

extern "C" void ReleaseDropInStaging(hipStream_t stream) {
    release_stream_compact_scratch(stream);
    std::lock_guard<std::mutex> lk(g_stage_mu);
    const auto it = g_stage.find(stream);
    if (it == g_stage.end()) return;
    free_staging(it->second);
    g_stage.erase(it);
}

extern "C" void LaunchLambda(float* lambda, const float4* pos_pred, const uint32_t* indicesSorted,
                             const uint32_t* keysSorted, const uint32_t* cellStart, const uint32_t* cellEnd,
                             sim::DeviceParams dp, uint32_t N, hipStream_t s) {
    if (N == 0 || !check_dims(dp, "LaunchLambda: empty grid")) return;
    launch_lambda(lambda, pos_pred, indicesSorted, keysSorted, DenseRows{cellStart, cellEnd}, dp, N, s);
    check_launch("LaunchLambda");
}

extern "C" void LaunchLambdaCompact(float* lambda, const float4* pos_pred, const uint32_t* indicesSorted,
                                    const uint32_t* keysSorted, const uint32_t* uniqueKeys, const uint32_t* offsets,
                                    const uint32_t* compactCount, sim::DeviceParams dp, uint32_t N, hipStream_t s) {
    if (N == 0 || !check_dims(dp, "LaunchLambdaCompact: empty grid")) return;
    launch_lambda(lambda, pos_pred, indicesSorted, keysSorted, CompactRowsDev{uniqueKeys, offsets, compactCount}, dp,
                  N, s);
    check_launch("LaunchLambdaCompact");
}

extern "C" void LaunchDeltaApply(float4* pos_pred, float4* delta, const float* lambda, const uint32_t* indicesSorted,
                                 const uint32_t* keysSorted, const uint32_t* cellStart, const uint32_t* cellEnd,
                                 sim::DeviceParams dp, uint32_t N, hipStream_t s) {
    if (N == 0 || !check_dims(dp, "LaunchDeltaApply: empty grid")) return;
    launch_delta(pos_pred, delta, lambda, indicesSorted, keysSorted, DenseRows{cellStart, cellEnd}, dp, N, s);
    check_launch("LaunchDeltaApply");
}

extern "C" void LaunchDeltaApplyCompact(float4* pos_pred, float4* delta, const float* lambda,
                                        const uint32_t* indicesSorted, const uint32_t* keysSorted,
                                        const uint32_t* uniqueKeys, const uint32_t* offsets,
                                        const uint32_t* compactCount, sim::DeviceParams dp, uint32_t N,
                                        hipStream_t s) {
    if (N == 0 || !check_dims(dp, "LaunchDeltaApplyCompact: empty grid")) return;
    launch_delta(pos_pred, delta, lambda, indicesSorted, keysSorted, CompactRowsDev{uniqueKeys, offsets, compactCount},
                 dp, N, s);
    check_launch("LaunchDeltaApplyCompact");
}

extern "C" void LaunchXSPH(float4* vel_out, const float4* vel_in, const float4* pos_pred,
                           const uint32_t* indicesSorted, const uint32_t* keysSorted, const uint32_t* cellStart,
                           const uint32_t* cellEnd, sim::DeviceParams dp, uint32_t N, hipStream_t s) {
    if (N == 0 || dp.xsph_c <= 0.f) return;
    if (vel_out == vel_in) { set_error(PBFX_E_ALIAS, "LaunchXSPH: vel_out aliases vel_in"); return; }
    if (!check_dims(dp, "LaunchXSPH: empty grid")) return;
    launch_xsph(vel_out, vel_in, pos_pred, indicesSorted, keysSorted, DenseRows{cellStart, cellEnd}, dp, N, s);
    check_launch("LaunchXSPH");
}

extern "C" void LaunchXSPHCompact(float4* vel_out, const float4* vel_in, const float4* pos_pred,
                                  const uint32_t* indicesSorted, const uint32_t* keysSorted,
                                  const uint32_t* uniqueKeys, const uint32_t* offsets, const uint32_t* compactCount,
                                  sim::DeviceParams dp, uint32_t N, hipStream_t s) {
    if (N == 0 || dp.xsph_c <= 0.f) return;
    if (vel_out == vel_in) { set_error(PBFX_E_ALIAS, "LaunchXSPHCompact: vel_out aliases vel_in"); return; }
    if (!check_dims(dp, "LaunchXSPHCompact: empty grid")) return;
    launch_xsph(vel_out, vel_in, pos_pred, indicesSorted, keysSorted,
                CompactRowsDev{uniqueKeys, offsets, compactCount}, dp, N, s);
    check_launch("LaunchXSPHCompact");
}
