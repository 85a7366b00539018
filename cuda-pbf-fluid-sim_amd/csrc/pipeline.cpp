// pipeline.cpp — headless phase pipeline, dispatcher registry and Simulator.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "grid.hpp"
#include "pbf_common.hpp"
#include "pbf_hip.h"
#include "radix_sort.hpp"
#include "sim_host.h"
#include "scenes.hpp"
#include "stats.hpp"

namespace sim {

using pbfx::set_error;

namespace {

// Event marks recorded after every launch while a profiled (eager) step runs.  The
// events come from the simulator's pool, created once: creating one per launch kept
// the host behind the short grid-build launches at the start of a step.
struct LaunchTimer {
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t>* pool = nullptr;
    std::vector<const char*> names;
    std::vector<hipEvent_t> events;
    void mark(const char* n) {
        if (pool->size() <= events.size()) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) return;
            pool->push_back(e);
        }
        hipEvent_t e = (*pool)[events.size()];
        (void)hipEventRecord(e, stream);
        names.push_back(n);
        events.push_back(e);
    }
};
thread_local LaunchTimer* t_timer = nullptr;
inline void mark(const char* n) {
    if (t_timer) t_timer->mark(n);
}

template <class T>
bool dalloc(T*& p, size_t count) {
    p = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * count) != hipSuccess) {
        p = nullptr;
        set_error(PBFX_E_CAPACITY, "device allocation failed");
        return false;
    }
    return true;
}

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Waits for the null stream, where the blocking hipMemset calls of the allocation
// routines run (they may return before the memory is cleared).
bool null_stream_sync() { return hipStreamSynchronize(nullptr) == hipSuccess; }

int reach_for(const SimParams& p) {
    const float cs = p.grid.cellSize;
    if (!(cs > 0.f)) return 1;
    const float q = std::ceil(p.kernel.h / cs);
    return std::max(1, q >= 1e6f ? 1000000 : (int)q);
}

__global__ void k_copy4(float4* __restrict__ dst, const float4* __restrict__ src, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

void copy4(float4* dst, const float4* src, uint32_t n, hipStream_t s) {
    if (n == 0) return;
    k_copy4<<<pbfx::ceil_div(n, 256), 256, 0, s>>>(dst, src, n);
    pbfx::check_launch("copy4");
}

// ------------------------------------------------------------ reference path
// Phase bodies restate sim/phases.cpp:33-105 and sim/post_ops.cpp:25-71 over the
// drop-in launchers.
void RefIntegrate(DeviceBuffers& b, GridBuffers&, const SimParams& p, hipStream_t s) {
    LaunchIntegratePred(b.d_pos, b.d_vel, b.d_pos_pred, p.gravity, p.dt, p.numParticles, s);
    mark("integrate");
    LaunchBoundary(b.d_pos_pred, b.d_vel, p.grid, 0.0f, p.numParticles, s);
    mark("boundary");
}

void RefGridBuild(DeviceBuffers& b, GridBuffers& g, const SimParams& p, hipStream_t s) {
    LaunchHashKeys(g.d_cellKeys, g.d_indices, b.d_pos_pred, p.grid, p.numParticles, s);
    mark("hash");
    LaunchSortPairs(g.d_sortTemp, g.sortTempBytes, g.d_cellKeys, g.d_cellKeys_sorted, g.d_indices,
                    g.d_indices_sorted, p.numParticles, s);
    mark("sort");
    if (g.kind == GridKind::Dense) {
        LaunchCellRanges(g.d_cellStart, g.d_cellEnd, g.d_cellKeys_sorted, p.numParticles, g.numCells, s);
    } else {
        LaunchCellRangesCompact(g.d_cellUniqueKeys, g.d_cellOffsets, g.d_compactCount, g.d_cellKeys_sorted,
                                p.numParticles, s);
    }
    mark("cell_ranges");
}

void RefSolve(DeviceBuffers& b, GridBuffers& g, const SimParams& p, hipStream_t s) {
    for (int it = 0; it < p.solverIters; ++it) {
        const DeviceParams dp = MakeDeviceParams(p);
        if (g.kind == GridKind::Dense) {
            LaunchLambda(b.d_lambda, b.d_pos_pred, g.d_indices_sorted, g.d_cellKeys_sorted, g.d_cellStart,
                         g.d_cellEnd, dp, p.numParticles, s);
            mark("lambda");
            LaunchDeltaApply(b.d_pos_pred, b.d_delta, b.d_lambda, g.d_indices_sorted, g.d_cellKeys_sorted,
                             g.d_cellStart, g.d_cellEnd, dp, p.numParticles, s);
        } else {
            LaunchLambdaCompact(b.d_lambda, b.d_pos_pred, g.d_indices_sorted, g.d_cellKeys_sorted,
                                g.d_cellUniqueKeys, g.d_cellOffsets, g.d_compactCount, dp, p.numParticles, s);
            mark("lambda");
            LaunchDeltaApplyCompact(b.d_pos_pred, b.d_delta, b.d_lambda, g.d_indices_sorted, g.d_cellKeys_sorted,
                                    g.d_cellUniqueKeys, g.d_cellOffsets, g.d_compactCount, dp, p.numParticles, s);
        }
        mark("delta_apply");
    }
    LaunchBoundary(b.d_pos_pred, b.d_vel, p.grid, 0.0f, p.numParticles, s);
    mark("boundary");
}

void RefVelocity(DeviceBuffers& b, GridBuffers&, const SimParams& p, hipStream_t s) {
    LaunchVelocity(b.d_vel, b.d_pos, b.d_pos_pred, 1.0f / p.dt, p.numParticles, s);
    mark("velocity");
}

void RefPost(DeviceBuffers& b, GridBuffers& g, const SimParams& p, hipStream_t s) {
    if (p.xsph_c > 0.f && p.numParticles > 0) {
        const DeviceParams dp = MakeDeviceParams(p);
        copy4(b.d_vel_snapshot, b.d_vel, p.numParticles, s);
        if (g.kind == GridKind::Dense) {
            LaunchXSPH(b.d_vel, b.d_vel_snapshot, b.d_pos_pred, g.d_indices_sorted, g.d_cellKeys_sorted,
                       g.d_cellStart, g.d_cellEnd, dp, p.numParticles, s);
        } else {
            LaunchXSPHCompact(b.d_vel, b.d_vel_snapshot, b.d_pos_pred, g.d_indices_sorted, g.d_cellKeys_sorted,
                              g.d_cellUniqueKeys, g.d_cellOffsets, g.d_compactCount, dp, p.numParticles, s);
        }
        mark("xsph");
    }
    LaunchBoundary(b.d_pos_pred, b.d_vel, p.grid, p.boundaryRestitution, p.numParticles, s);
    mark("boundary");
}

// ------------------------------------------------------------ reordered path

void SortedPredict(DeviceBuffers& b, GridBuffers& g, const SimParams& p, hipStream_t s) {
    pbfx::launch_predict_hash(b.d_pos, b.d_vel, b.d_pos_pred, g.d_cellKeys,
                              g.kind == GridKind::Hashed ? g.table.epoch : nullptr, p, g.scenes, s,
                              b.rec.predClamp, g.d_stepFlags);
    mark("predict_hash");
}

void SortedGridBuild(DeviceBuffers& b, GridBuffers& g, const SimParams& p, hipStream_t s) {
    const uint32_t N = p.numParticles;
    pbfx::radix_sort_pairs(g.d_sortTemp, g.sortTempBytes, g.d_cellKeys, g.d_cellKeys_sorted, nullptr,
                           g.d_indices_sorted, N, 0, g.keyBits, s, g.sortItems, g.sortDigitBits);
    mark("sort");
    pbfx::compact_segments(g.d_cellUniqueKeys, g.d_cellOffsets, g.d_compactCount, g.d_segOf, g.d_cellKeys_sorted, N,
                           g.d_blockCounts, g.d_blockOffsets, s);
    mark("compact");
    pbfx::CellLookup lk = pbfx::CellLookup::Compact;
    if (g.kind == GridKind::Dense) {
        LaunchCellRanges(g.d_cellStart, g.d_cellEnd, g.d_cellKeys_sorted, N, g.numCells, s);
        lk = pbfx::CellLookup::Dense;
    } else if (g.kind == GridKind::Hashed) {
        pbfx::launch_hash_insert(g.table, g.d_cellUniqueKeys, g.d_cellOffsets, g.d_compactCount, N, s);
        lk = pbfx::CellLookup::Hashed;
    }
    if (lk != pbfx::CellLookup::Compact) mark("cell_lookup");
    pbfx::launch_spans_gather(lk, b.d_pos_pred, g.d_indices_sorted, b.rec.P[0] ? b.rec.P[0] : b.d_pos_sorted[0], N,
                              g.d_cellUniqueKeys,
                              g.d_cellOffsets, g.d_compactCount, g.d_cellStart, g.d_cellEnd, g.table, g.spans,
                              p.grid, g.scenes, s);
    mark("spans_gather");
    if (g.laneOrder) {
        pbfx::launch_lane_order(b.rec.P[0] ? b.rec.P[0] : b.d_pos_sorted[0], N, p.grid, g.scenes, g.d_laneOrder, s);
        mark("lane_order");
    }
    // XSPH's slot pairs (the compaction's tile scratch is free again)
    pbfx::build_slot_pairs(g.d_slotPairs, g.d_slotPairCount, g.d_segOf, g.d_cellOffsets, N, g.d_blockCounts,
                           g.d_blockOffsets, s);
    mark("slot_pairs");
}

void SortedSolve(DeviceBuffers& b, GridBuffers& g, const SimParams& p, hipStream_t s) {
    const uint32_t N = p.numParticles;
    const DeviceParams dp = MakeDeviceParams(p);
    const float invDt = 1.0f / p.dt;
    if (p.solverIters <= 0) {
        pbfx::launch_finalize_sorted(b.d_pos_sorted[0], b.d_pos, g.d_indices_sorted, b.d_vel_sorted, p.grid, invDt,
                                     N, g.scenes, s, g.d_stepFlags);
        mark("finalize");
        return;
    }
    // The candidate mask, built from the predicted positions before the first iteration,
    // serves every lambda pass (they stop at the neighbour cap, so skipping the candidates
    // outside the skin pays).  The delta passes, which take every neighbour, walk every
    // candidate: at the collapsed state that measured faster than walking the mask.
    pbfx::CandMask mk = g.mask;
    mk.rs2 = pbfx::mask_skin_radius2(p);
    const bool useMask = g.candidateMask && g.mask.words != nullptr && mk.rs2 > 0.f;
    const bool rec = b.rec.P[0] != nullptr;
    const uint32_t* perm = g.laneOrder ? g.d_laneOrder : nullptr;
    if (useMask) {
        pbfx::launch_mask_build(rec ? b.rec.P[0] : b.d_pos_sorted[0], g.d_segOf, g.spans, mk, N, g.d_slotPairs,
                                g.d_slotPairCount, s);
        mark("mask_build");
    }
    int cur = 0;
    for (int it = 0; it < p.solverIters; ++it) {
        float4* Pa = rec ? b.rec.P[it] : b.d_pos_sorted[cur];
        float4* Pb = rec ? b.rec.P[it + 1] : b.d_pos_sorted[cur ^ 1];
        const bool last = it == p.solverIters - 1;
        pbfx::launch_lambda_sorted(Pa, g.d_segOf, g.spans, mk, useMask, dp, N, s,
                                   rec ? b.rec.lamA[it] : nullptr, rec ? b.rec.lamB[it] : nullptr, perm,
                                   g.d_stepFlags);
        mark("lambda");
        pbfx::launch_delta_sorted(Pa, Pb, g.d_segOf, g.spans, dp, N, last, b.d_pos, g.d_indices_sorted,
                                  rec ? b.rec.Vs : b.d_vel_sorted, invDt, g.scenes, s,
                                  last ? b.rec.finalClamp : nullptr, rec ? b.rec.dispU[it] : nullptr, perm,
                                  g.d_stepFlags, g.d_ratioCoef);
        mark(last ? "delta_last" : "delta");
        cur ^= 1;
    }
}

void SortedPost(DeviceBuffers& b, GridBuffers& g, const SimParams& p, hipStream_t s) {
    const int fin = p.solverIters > 0 ? (p.solverIters & 1) : 0;
    // XSPH walks its candidates directly, evaluating each in place, two particles of a
    // cell a lane: each of its pairs also gathers a velocity, and the scattered loads of
    // the candidate-mask walk cost more than the consecutive candidates the mask would skip
    const bool rec = b.rec.P[0] != nullptr;
    pbfx::launch_xsph_scatter(rec ? b.rec.P[p.solverIters] : b.d_pos_sorted[fin], rec ? b.rec.Vs : b.d_vel_sorted,
                              g.d_segOf, g.spans, MakeDeviceParams(p), p.numParticles, g.d_slotPairs,
                              g.d_slotPairCount, g.d_indices_sorted, b.d_pos_pred, b.d_vel, g.scenes, s,
                              rec ? b.rec.xsphGm : nullptr, g.d_stepFlags);
    mark("xsph_scatter");
}

// Phase objects look their launch function up in the dispatcher.
class DispatchedPhase : public IPhase {
public:
    DispatchedPhase(PhaseType t, const char* n) : m_type(t), m_name(n) {}
    PhaseType type() const override { return m_type; }
    const char* name() const override { return m_name; }
    void run(SimulationContext& ctx, const SimParams& p, hipStream_t s) override {
        if (p.numParticles == 0) return;
        KernelLaunchFn fn = ctx.dispatcher->find(KernelVariant{ctx.grid->kind, ctx.path}, m_type);
        if (!fn) {
            set_error(PBFX_E_INVALID_ARGUMENT, "no kernel registered for phase");
            return;
        }
        fn(*ctx.bufs, *ctx.grid, p, s);
    }

private:
    PhaseType m_type;
    const char* m_name;
};

}  // namespace

// ------------------------------------------------------------------ buffers
bool DeviceBuffers::allocate(uint32_t cap) {
    release();
    if (cap == 0) cap = 1;
    bool ok = dalloc(d_pos_curr, cap) && dalloc(d_pos_next, cap) && dalloc(d_vel, cap) && dalloc(d_lambda, cap) &&
              dalloc(d_delta, cap) && dalloc(d_vel_snapshot, cap) &&
              dalloc(d_pos_sorted[0], (size_t)cap + pbfx::kSlotPad) &&
              dalloc(d_pos_sorted[1], (size_t)cap + pbfx::kSlotPad) &&
              dalloc(d_vel_sorted, (size_t)cap + pbfx::kSlotPad);
    if (!ok) {
        release();
        return false;
    }
    capacity = cap;
    d_pos = d_pos_curr;
    d_pos_pred = d_pos_next;
    ok = hipMemset(d_pos_curr, 0, sizeof(float4) * cap) == hipSuccess &&
         hipMemset(d_pos_next, 0, sizeof(float4) * cap) == hipSuccess &&
         hipMemset(d_vel, 0, sizeof(float4) * cap) == hipSuccess &&
         hipMemset(d_pos_sorted[0], 0, sizeof(float4) * ((size_t)cap + pbfx::kSlotPad)) == hipSuccess &&
         hipMemset(d_pos_sorted[1], 0, sizeof(float4) * ((size_t)cap + pbfx::kSlotPad)) == hipSuccess &&
         hipMemset(d_vel_sorted, 0, sizeof(float4) * ((size_t)cap + pbfx::kSlotPad)) == hipSuccess;
    // hipMemset of device memory may return before it lands, and the simulator's
    // stream does not wait for the null stream
    return null_stream_sync() && ok;
}

void DeviceBuffers::release() {
    if (externalPos) {
        d_pos_curr = d_pos_next = nullptr;
        externalPos = false;
    }
    dfree(d_pos_curr); dfree(d_pos_next); dfree(d_vel); dfree(d_lambda); dfree(d_delta); dfree(d_vel_snapshot);
    dfree(d_pos_sorted[0]); dfree(d_pos_sorted[1]); dfree(d_vel_sorted);
    d_pos = d_pos_pred = nullptr;
    capacity = 0;
}

void DeviceBuffers::adoptExternalPositions(float4* a, float4* b) {
    if (!externalPos) {
        dfree(d_pos_curr);
        dfree(d_pos_next);
    }
    d_pos_curr = a;
    d_pos_next = b;
    d_pos = d_pos_curr;
    d_pos_pred = d_pos_next;
    externalPos = true;
}

void DeviceBuffers::swapPositionPingPong() {
    std::swap(d_pos_curr, d_pos_next);
    d_pos = d_pos_curr;
    d_pos_pred = d_pos_next;
}

bool GridBuffers::allocate(uint32_t cap, uint32_t nCells, GridKind k, uint32_t tLog2, int reach, bool reordered) {
    release();
    if (cap == 0) cap = 1;
    // The walks index the span table (rows x capacity entries) with 32-bit integers; a cell
    // size far below h (reach = ceil(h / cellSize) cells a side) would need more rows.
    const uint64_t rows = (2 * (uint64_t)std::max(reach, 1) + 1) * (2 * (uint64_t)std::max(reach, 1) + 1);
    if (rows * cap >= (uint64_t(1) << 32) || rows > 0x7FFFFFFFull) {
        set_error(PBFX_E_CAPACITY, "grid: neighbour rows x particles exceed 2^32 (cell size too small for h)");
        return false;
    }
    kind = k;
    capacity = cap;
    numCells = nCells;
    keyBits = reordered ? pbfx::key_bits_for(nCells) : 32;
    bool ok = dalloc(d_cellKeys, cap) && dalloc(d_cellKeys_sorted, cap) && dalloc(d_indices, cap) &&
              dalloc(d_indices_sorted, cap);
    sortTempBytes = pbfx::radix_sort_temp_bytes(cap, 0, 32);
    ok = ok && (hipMalloc(&d_sortTemp, sortTempBytes) == hipSuccess);
    if (k == GridKind::Dense) ok = ok && dalloc(d_cellStart, nCells) && dalloc(d_cellEnd, nCells);
    ok = ok && dalloc(d_cellUniqueKeys, cap) && dalloc(d_cellOffsets, (size_t)cap + 1) && dalloc(d_compactCount, 1);
    // segment index, row spans and compaction scratch: used by the reordered step and,
    // on both paths, by the density statistics
    const uint32_t tiles = pbfx::compact_scratch_elems(cap);
    spans.reach = reach;
    spans.rows = (int)rows;
    spans.stride = cap;
    ok = ok && dalloc(d_segOf, cap) && dalloc(d_blockCounts, tiles) && dalloc(d_blockOffsets, tiles) &&
         dalloc(spans.spans, (size_t)spans.rows * cap);
    if (reordered) {
        mask.stride = cap;
        ok = ok && dalloc(mask.words, (size_t)pbfx::kCandMaskWords * cap) && dalloc(d_laneOrder, cap) &&
             dalloc(d_slotPairs, cap) && dalloc(d_slotPairCount, 1) && dalloc(d_stepFlags, 1) &&
             hipMemset(d_stepFlags, 0xFF, sizeof(uint32_t)) == hipSuccess && dalloc(d_ratioCoef, 1) &&
             dalloc(d_ratioBad, 1) && hipMemset(d_ratioCoef, 0, sizeof(pbfx::RatioCoef)) == hipSuccess;
        if (k == GridKind::Hashed) {
            tableLog2 = tLog2;
            const size_t slots = size_t(1) << tLog2;
            table.mask = (uint32_t)(slots - 1);
            // a key space larger than the table gets hashed homes (placement only: every
            // lookup compares full keys, so results never depend on it)
            table.scatter = (uint64_t)nCells > (uint64_t)slots ? 1u : 0u;
            ok = ok && dalloc(table.slots, slots) && dalloc(table.epoch, 1) && dalloc(table.overflow, 1);
            ok = ok && hipMemset(table.slots, 0, sizeof(pbfx::HashSlot) * slots) == hipSuccess &&
                 hipMemset(table.epoch, 0, sizeof(uint32_t)) == hipSuccess &&
                 hipMemset(table.overflow, 0, sizeof(uint32_t)) == hipSuccess;
        }
    }
    if (!ok) {
        release();
        return false;
    }
    ok = hipMemset(d_compactCount, 0, sizeof(uint32_t)) == hipSuccess;
    return null_stream_sync() && ok;
}

void GridBuffers::release() {
    dfree(d_cellKeys); dfree(d_cellKeys_sorted); dfree(d_indices); dfree(d_indices_sorted);
    dfree(d_cellStart); dfree(d_cellEnd); dfree(d_cellUniqueKeys); dfree(d_cellOffsets); dfree(d_compactCount);
    dfree(d_segOf); dfree(d_blockCounts); dfree(d_blockOffsets); dfree(d_laneOrder);
    dfree(d_slotPairs); dfree(d_slotPairCount); dfree(d_stepFlags); dfree(d_ratioCoef); dfree(d_ratioBad);
    if (d_sortTemp) (void)hipFree(d_sortTemp);
    d_sortTemp = nullptr;
    sortTempBytes = 0;
    dfree(spans.spans);
    dfree(mask.words);
    dfree(table.slots); dfree(table.epoch); dfree(table.overflow);
    capacity = 0;
}

// --------------------------------------------------------------- dispatcher
KernelLaunchFn KernelDispatcher::find(const KernelVariant& v, PhaseType ph) const {
    auto it = m_map.find(key(ph, v));
    if (it != m_map.end()) return it->second;
    if (v.grid == GridKind::Hashed) {
        it = m_map.find(key(ph, KernelVariant{GridKind::Compact, v.path}));
        if (it != m_map.end()) return it->second;
    }
    it = m_map.find(key(ph, KernelVariant{GridKind::Dense, v.path}));
    return it != m_map.end() ? it->second : nullptr;
}

void InitializeKernelDispatcher(KernelDispatcher& kd) {
    // Grid-kind differences live inside the launch functions (GridBuffers::kind);
    // the registry is keyed so a specialised variant can be slotted in per kind.
    kd.registerKernel(PhaseType::Integrate, {GridKind::Dense, PathKind::Reference}, RefIntegrate);
    kd.registerKernel(PhaseType::GridBuild, {GridKind::Dense, PathKind::Reference}, RefGridBuild);
    kd.registerKernel(PhaseType::SolveIterations, {GridKind::Dense, PathKind::Reference}, RefSolve);
    kd.registerKernel(PhaseType::Velocity, {GridKind::Dense, PathKind::Reference}, RefVelocity);
    kd.registerKernel(PhaseType::Post, {GridKind::Dense, PathKind::Reference}, RefPost);
    kd.registerKernel(PhaseType::Integrate, {GridKind::Dense, PathKind::Reordered}, SortedPredict);
    kd.registerKernel(PhaseType::GridBuild, {GridKind::Dense, PathKind::Reordered}, SortedGridBuild);
    kd.registerKernel(PhaseType::SolveIterations, {GridKind::Dense, PathKind::Reordered}, SortedSolve);
    kd.registerKernel(PhaseType::Post, {GridKind::Dense, PathKind::Reordered}, SortedPost);
}

void BuildDefaultPipelines(PhasePipeline& pl, PathKind path) {
    pl.clear();
    pl.add(std::make_unique<DispatchedPhase>(PhaseType::Integrate, "integrate"));
    pl.add(std::make_unique<DispatchedPhase>(PhaseType::GridBuild, "grid_build"));
    pl.add(std::make_unique<DispatchedPhase>(PhaseType::SolveIterations, "solve"));
    // the reordered path fuses the velocity phase into the last delta pass
    if (path == PathKind::Reference) pl.add(std::make_unique<DispatchedPhase>(PhaseType::Velocity, "velocity"));
    pl.add(std::make_unique<DispatchedPhase>(PhaseType::Post, "post"));
}

// ---------------------------------------------------------------- simulator
float JitterUniform(uint64_t seed, uint64_t particle, uint32_t axis) {
    return pbfx::jitter_uniform(seed, particle, axis);
}

bool Simulator::initialize(const SimParams& p, const SimOptions& opt) {
    shutdown();
    if (p.grid.dim.x <= 0 || p.grid.dim.y <= 0 || p.grid.dim.z <= 0 || !(p.grid.cellSize > 0.f)) {
        set_error(PBFX_E_INVALID_ARGUMENT, "Simulator::initialize: empty grid");
        return false;
    }
    const uint64_t cells = (uint64_t)p.grid.dim.x * (uint64_t)p.grid.dim.y * (uint64_t)p.grid.dim.z;
    return initCommon(p, opt, cells);
}

bool Simulator::initializeEnsemble(const SimParams& base, const std::vector<SceneDesc>& scenes,
                                   const SimOptions& opt) {
    shutdown();
    if (scenes.empty() || opt.path != PathKind::Reordered || !(base.grid.cellSize > 0.f)) {
        set_error(PBFX_E_INVALID_ARGUMENT, "initializeEnsemble: needs scenes, the reordered path and a cell size");
        return false;
    }
    const float cs = base.grid.cellSize;
    uint64_t cells = 0, parts = 0;
    m_scenes.clear();
    std::vector<uint32_t> cellBase;
    for (const SceneDesc& d : scenes) {
        pbfx::SceneParams sp{};
        sp.grid.mins = d.boxMin;
        sp.grid.maxs = d.boxMax;
        sp.grid.cellSize = cs;
        // dims as the reference derives them (console.cpp:258-265): max(1, ceil(extent / cell))
        auto dim1 = [cs](float lo, float hi) { return std::max(1, (int)std::ceil((hi - lo) / cs)); };
        sp.grid.dim = make_int3(dim1(d.boxMin.x, d.boxMax.x), dim1(d.boxMin.y, d.boxMax.y),
                                dim1(d.boxMin.z, d.boxMax.z));
        sp.gravity = d.gravity;
        sp.xsph_c = d.xsph_c;
        sp.restitution = d.restitution;
        sp.cellBase = (uint32_t)cells;
        sp.particleBase = (uint32_t)parts;
        sp.numParticles = d.numParticles;
        cellBase.push_back((uint32_t)cells);
        cells += (uint64_t)sp.grid.dim.x * sp.grid.dim.y * sp.grid.dim.z;
        parts += d.numParticles;
        if (cells > 0xFFFFFFFFull || parts > 0xFFFFFFFFull) {
            set_error(PBFX_E_INVALID_ARGUMENT, "initializeEnsemble: more than 2^32 cells or particles");
            return false;
        }
        m_scenes.push_back(sp);
    }
    cellBase.push_back((uint32_t)std::min<uint64_t>(cells, 0xFFFFFFFFull));
    SimParams p = base;
    p.numParticles = p.maxParticles = (uint32_t)parts;
    p.grid = m_scenes[0].grid;
    p.gravity = m_scenes[0].gravity;
    p.xsph_c = m_scenes[0].xsph_c;
    p.boundaryRestitution = m_scenes[0].restitution;
    if (!initCommon(p, opt, cells)) return false;
    const size_t B = m_scenes.size();
    std::vector<uint32_t> sceneOf(parts);
    for (size_t b = 0; b < B; ++b)
        std::fill(sceneOf.begin() + m_scenes[b].particleBase,
                  sceneOf.begin() + m_scenes[b].particleBase + m_scenes[b].numParticles, (uint32_t)b);
    bool ok = dalloc(d_sceneParams, B) && dalloc(d_sceneOf, parts) && dalloc(d_sceneOfSlot, parts) &&
              dalloc(d_cellBase, B + 1);
    ok = ok && copyOnStream(d_sceneParams, m_scenes.data(), sizeof(pbfx::SceneParams) * B, hipMemcpyHostToDevice);
    ok = ok && copyOnStream(d_sceneOf, sceneOf.data(), sizeof(uint32_t) * parts, hipMemcpyHostToDevice);
    ok = ok && copyOnStream(d_cellBase, cellBase.data(), sizeof(uint32_t) * (B + 1), hipMemcpyHostToDevice);
    ok = ok && syncStream();
    if (!ok) {
        set_error(PBFX_E_CAPACITY, "initializeEnsemble: scene table upload failed");
        shutdown();
        return false;
    }
    m_grid.scenes = pbfx::Scenes{d_sceneParams, d_sceneOf, d_sceneOfSlot, d_cellBase, (uint32_t)B};
    return true;
}

bool Simulator::initCommon(const SimParams& p, const SimOptions& opt, uint64_t cells) {
    if (cells > 0xFFFFFFFFull) {
        set_error(PBFX_E_INVALID_ARGUMENT, "Simulator::initialize: more than 2^32 cells");
        return false;
    }
    if (opt.grid == GridKind::Hashed) {
        if (opt.path != PathKind::Reordered || opt.tableLog2 < 4 || opt.tableLog2 > 30 ||
            (uint64_t(1) << opt.tableLog2) < std::max<uint64_t>(p.numParticles, 1)) {
            set_error(PBFX_E_INVALID_ARGUMENT, "Simulator::initialize: hash table smaller than the particle count");
            return false;
        }
    }
    m_params = p;
    m_opt = opt;
    if (hipStreamCreateWithFlags(&m_stream, hipStreamNonBlocking) != hipSuccess) {
        set_error(PBFX_E_INVALID_ARGUMENT, "stream creation failed");
        return false;
    }
    const uint32_t cap = std::max<uint32_t>(p.maxParticles, p.numParticles);
    if (!m_bufs.allocate(cap)) return false;
    if (!m_grid.allocate(cap, (uint32_t)cells, opt.grid, opt.tableLog2, reach_for(p),
                         opt.path == PathKind::Reordered))
        return false;
    if (opt.path == PathKind::Reference && opt.grid != GridKind::Dense) {
        if (!EnsureCellCompactScratch(cap, 256)) return false;
    }
    // the drop-in neighbour launchers on this simulator's stream walk cell-sorted copies
    if (opt.path == PathKind::Reference && !EnsureDropInStaging(cap, m_stream)) return false;
    InitializeKernelDispatcher(m_dispatcher);
    BuildDefaultPipelines(m_pipeline, opt.path);
    m_ctx.bufs = &m_bufs;
    m_ctx.grid = &m_grid;
    m_ctx.useHashedGrid = opt.grid != GridKind::Dense;
    m_ctx.path = opt.path;
    m_ctx.dispatcher = &m_dispatcher;
    m_parity = 0;
    m_frameIndex = 0;
    m_graphDirty = true;
    m_initialized = true;
    checkScorrRatio();
    return true;
}

void Simulator::checkScorrRatio() {
    if (m_grid.d_ratioCoef)
        pbfx::launch_ratio_check(MakeDeviceParams(m_params), m_grid.d_ratioCoef, m_grid.d_ratioBad, m_stream);
}

void Simulator::releaseScenes() {
    dfree(d_sceneParams);
    dfree(d_sceneOf);
    dfree(d_sceneOfSlot);
    dfree(d_cellBase);
    m_scenes.clear();
    m_grid.scenes = pbfx::Scenes{};
}

void Simulator::destroyGraphs() {
    for (auto& e : m_exec) {
        if (e) (void)hipGraphExecDestroy(e);
        e = nullptr;
    }
}

void Simulator::shutdown() {
    if (m_stream) (void)hipStreamSynchronize(m_stream);
    for (auto& e : m_eventPool) (void)hipEventDestroy(e);
    m_eventPool.clear();
    destroyGraphs();
    releaseAdjoint();
    releaseScenes();
    m_bufs.release();
    m_grid.release();
    if (m_stream) {
        (void)hipStreamSynchronize(m_stream);
        ReleaseDropInStaging(m_stream);
        (void)hipStreamDestroy(m_stream);
    }
    m_stream = nullptr;
    m_initialized = false;
}

void Simulator::runPipeline(hipStream_t s) { m_pipeline.runFull(m_ctx, m_params, s); }

bool Simulator::captureGraphs() {
    destroyGraphs();
    for (int k = 0; k < 2; ++k) {
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(m_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
        runPipeline(m_stream);
        if (hipStreamEndCapture(m_stream, &g) != hipSuccess || !g) return false;
        const int slot = m_parity ^ k;
        const hipError_t e = hipGraphInstantiate(&m_exec[slot], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            set_error((int)e, "hipGraphInstantiate");
            return false;
        }
        m_bufs.swapPositionPingPong();
    }
    // two swaps: orientation restored
    m_graphDirty = false;
    return true;
}

// Structural changes, as the reference's step takes them (simulator.cpp:583-607 with
// ParamChangeTracker::structuralChanged, param_change_tracker.h:37-44): a particle
// count beyond the buffers re-allocates them, a new grid (box, cell size or dims) or
// smoothing length re-allocates the cell grid, and the step is re-captured.  Particle
// state is kept: the first min(old, new) particles keep their positions and
// velocities, new ones start at rest at the origin (or wherever the caller seeds them,
// e.g. pbfSimSeedLattice at an offset).
bool Simulator::restructure(const SimParams& p) {
    if (!m_scenes.empty()) {
        set_error(PBFX_E_INVALID_ARGUMENT,
                  "Simulator::setParams: an ensemble's particle count, grid, h and iterations are fixed");
        return false;
    }
    if (!m_tapes.empty()) {
        set_error(PBFX_E_INVALID_ARGUMENT, "Simulator::setParams: structural change while adjoint tapes are held");
        return false;
    }
    if (p.grid.dim.x <= 0 || p.grid.dim.y <= 0 || p.grid.dim.z <= 0 || !(p.grid.cellSize > 0.f) ||
        p.solverIters < 0) {
        set_error(PBFX_E_INVALID_ARGUMENT, "Simulator::setParams: empty grid or negative iterations");
        return false;
    }
    const uint64_t cells = (uint64_t)p.grid.dim.x * (uint64_t)p.grid.dim.y * (uint64_t)p.grid.dim.z;
    if (cells > 0xFFFFFFFFull) {
        set_error(PBFX_E_INVALID_ARGUMENT, "Simulator::setParams: more than 2^32 cells");
        return false;
    }
    const uint32_t need = std::max<uint32_t>(p.maxParticles, p.numParticles);
    const bool grow = need > m_bufs.capacity;
    if (grow && m_bufs.externalPos) {
        set_error(PBFX_E_CAPACITY, "Simulator::setParams: more particles than the bound external buffers hold");
        return false;
    }
    if (m_opt.grid == GridKind::Hashed && (uint64_t(1) << m_opt.tableLog2) < std::max<uint64_t>(p.numParticles, 1)) {
        set_error(PBFX_E_INVALID_ARGUMENT, "Simulator::setParams: hash table smaller than the particle count");
        return false;
    }
    if (hipStreamSynchronize(m_stream) != hipSuccess) return false;
    destroyGraphs();
    const uint32_t cap = grow ? need : m_bufs.capacity;
    if (grow) {
        // keep the current state of the existing particles across the re-allocation
        const uint32_t keep = std::min(m_params.numParticles, m_bufs.capacity);
        float4 *pos = nullptr, *vel = nullptr;
        if (!dalloc(pos, keep) || !dalloc(vel, keep)) {
            dfree(pos);
            return false;
        }
        bool ok = copyOnStream(pos, m_bufs.d_pos, sizeof(float4) * keep, hipMemcpyDeviceToDevice) &&
                  copyOnStream(vel, m_bufs.d_vel, sizeof(float4) * keep, hipMemcpyDeviceToDevice) && syncStream();
        ok = ok && m_bufs.allocate(cap);
        m_parity = 0;
        ok = ok && copyOnStream(m_bufs.d_pos, pos, sizeof(float4) * keep, hipMemcpyDeviceToDevice) &&
             copyOnStream(m_bufs.d_pos_pred, pos, sizeof(float4) * keep, hipMemcpyDeviceToDevice) &&
             copyOnStream(m_bufs.d_vel, vel, sizeof(float4) * keep, hipMemcpyDeviceToDevice) && syncStream();
        dfree(pos);
        dfree(vel);
        if (!ok) {
            set_error(PBFX_E_CAPACITY, "Simulator::setParams: re-allocation failed");
            m_initialized = false;
            return false;
        }
    }
    const pbfx::SpanTable keepSpans = m_grid.spans;
    const bool keepCandidate = m_grid.candidateMask, keepLaneOrder = m_grid.laneOrder;
    const int items = m_grid.sortItems, digits = m_grid.sortDigitBits;
    if (!m_grid.allocate(cap, (uint32_t)cells, m_opt.grid, m_opt.tableLog2, reach_for(p),
                         m_opt.path == PathKind::Reordered)) {
        m_initialized = false;
        return false;
    }
    // launch knobs survive the re-allocation
    m_grid.candidateMask = keepCandidate;
    m_grid.laneOrder = keepLaneOrder;
    m_grid.sortItems = items;
    m_grid.sortDigitBits = digits;
    m_grid.spans.streamRatio = keepSpans.streamRatio;
    m_grid.spans.chunkBits = keepSpans.chunkBits;
    m_grid.spans.block = keepSpans.block;
    if (m_opt.path == PathKind::Reference) {
        if (m_opt.grid != GridKind::Dense && !EnsureCellCompactScratch(cap, 256)) return false;
        if (!EnsureDropInStaging(cap, m_stream)) return false;
    }
    m_params = p;
    m_graphDirty = true;
    checkScorrRatio();
    return true;
}

bool Simulator::setParams(const SimParams& p) {
    if (!m_initialized) return false;
    if (std::memcmp(&p, &m_params, sizeof(SimParams)) == 0) return true;
    if (p.numParticles != m_params.numParticles || p.maxParticles != m_params.maxParticles ||
        std::memcmp(&p.grid, &m_params.grid, sizeof(p.grid)) != 0 || p.kernel.h != m_params.kernel.h ||
        p.solverIters != m_params.solverIters)
        return restructure(p);
    // an ensemble holds box, gravity, XSPH and restitution per scene; only the shared
    // parameters (dt, solver coefficients, clamps, neighbour cap) can change here
    if (!m_scenes.empty() &&
        (std::memcmp(&p.grid, &m_params.grid, sizeof(p.grid)) != 0 || p.gravity.x != m_params.gravity.x ||
         p.gravity.y != m_params.gravity.y || p.gravity.z != m_params.gravity.z || p.xsph_c != m_params.xsph_c ||
         p.boundaryRestitution != m_params.boundaryRestitution)) {
        set_error(PBFX_E_INVALID_ARGUMENT,
                  "Simulator::setParams: box, gravity, xsph_c and restitution are per scene in an ensemble");
        return false;
    }
    m_params = p;
    m_graphDirty = true;
    checkScorrRatio();
    return true;
}

bool Simulator::step(const SimParams& p) { return setParams(p) && stepMany(1); }

bool Simulator::stepMany(int steps) {
    if (!m_initialized) return false;
    for (int k = 0; k < steps; ++k) {
        if (m_opt.useGraph) {
            if (m_graphDirty && !captureGraphs()) {
                set_error(PBFX_E_INVALID_ARGUMENT, "graph capture failed");
                return false;
            }
            const hipError_t e = hipGraphLaunch(m_exec[m_parity], m_stream);
            if (e != hipSuccess) {
                set_error((int)e, "hipGraphLaunch");
                return false;
            }
        } else {
            runPipeline(m_stream);
        }
        m_bufs.swapPositionPingPong();
        m_parity ^= 1;
        ++m_frameIndex;
    }
    return true;
}

bool Simulator::profilePhases(std::vector<std::pair<std::string, float>>& out) {
    out.clear();
    if (!m_initialized) return false;
    LaunchTimer timer;
    timer.stream = m_stream;
    timer.pool = &m_eventPool;
    timer.mark("start");
    t_timer = &timer;
    runPipeline(m_stream);
    t_timer = nullptr;
    bool ok = !timer.events.empty() && hipEventSynchronize(timer.events.back()) == hipSuccess;
    for (size_t i = 1; ok && i < timer.events.size(); ++i) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, timer.events[i - 1], timer.events[i]);
        out.emplace_back(timer.names[i], ms);
    }
    if (!ok) return false;
    m_bufs.swapPositionPingPong();
    m_parity ^= 1;
    ++m_frameIndex;
    return true;
}

bool Simulator::seedBoxLattice(uint32_t first, uint32_t nx, uint32_t ny, uint32_t nz, float3 o, float spacing,
                               float amp, uint64_t seed) {
    return fillLattice(m_params.grid, first, nx, ny, nz, o, spacing, amp, seed);
}

bool Simulator::seedSceneLattice(uint32_t scene, uint32_t nx, uint32_t ny, uint32_t nz, float3 o, float spacing,
                                 float amp, uint64_t seed) {
    if (scene >= m_scenes.size() || (uint64_t)nx * ny * nz > m_scenes[scene].numParticles) {
        set_error(PBFX_E_INVALID_ARGUMENT, "seedSceneLattice: bad scene or block larger than the scene");
        return false;
    }
    return fillLattice(m_scenes[scene].grid, m_scenes[scene].particleBase, nx, ny, nz, o, spacing, amp, seed);
}

bool Simulator::getSceneState(uint32_t scene, float4* pos, float4* vel) const {
    if (!m_initialized || scene >= m_scenes.size()) return false;
    if (hipStreamSynchronize(m_stream) != hipSuccess) return false;
    const pbfx::SceneParams& sp = m_scenes[scene];
    const size_t bytes = sizeof(float4) * sp.numParticles;
    bool ok = true;
    if (pos) ok = ok && copyOnStream(pos, m_bufs.d_pos + sp.particleBase, bytes, hipMemcpyDeviceToHost);
    if (vel) ok = ok && copyOnStream(vel, m_bufs.d_vel + sp.particleBase, bytes, hipMemcpyDeviceToHost);
    return syncStream() && ok;
}

bool Simulator::fillLattice(const GridBounds& g, uint32_t first, uint32_t nx, uint32_t ny, uint32_t nz, float3 o,
                            float spacing, float amp, uint64_t seed) {
    const uint64_t n64 = (uint64_t)nx * ny * nz;
    if (!m_initialized || first + n64 > m_bufs.capacity) {
        set_error(PBFX_E_CAPACITY, "seedBoxLattice: block exceeds capacity");
        return false;
    }
    const uint32_t n = (uint32_t)n64;
    std::vector<float4> h(n);
    uint32_t idx = 0;
    for (uint32_t iz = 0; iz < nz; ++iz)
        for (uint32_t iy = 0; iy < ny; ++iy)
            for (uint32_t ix = 0; ix < nx; ++ix) {
                float x = o.x + (float)ix * spacing;
                float y = o.y + (float)iy * spacing;
                float z = o.z + (float)iz * spacing;
                x = std::fmin(std::fmax(x, g.mins.x), g.maxs.x);
                y = std::fmin(std::fmax(y, g.mins.y), g.maxs.y);
                z = std::fmin(std::fmax(z, g.mins.z), g.maxs.z);
                h[idx++] = make_float4(x, y, z, 1.0f);
            }
    if (amp > 0.f) {
        for (uint32_t i = 0; i < n; ++i) {
            const float jx = (2.0f * JitterUniform(seed, i, 0) - 1.0f) * amp;
            const float jy = (2.0f * JitterUniform(seed, i, 1) - 1.0f) * amp;
            const float jz = (2.0f * JitterUniform(seed, i, 2) - 1.0f) * amp;
            h[i].x = std::fmin(std::fmax(h[i].x + jx, g.mins.x), g.maxs.x);
            h[i].y = std::fmin(std::fmax(h[i].y + jy, g.mins.y), g.maxs.y);
            h[i].z = std::fmin(std::fmax(h[i].z + jz, g.mins.z), g.maxs.z);
        }
    }
    return uploadBlock(first, h, "seedBoxLattice: copy failed");
}

bool Simulator::copyOnStream(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) const {
    return bytes == 0 || hipMemcpyAsync(dst, src, bytes, kind, m_stream) == hipSuccess;
}

bool Simulator::syncStream() const { return hipStreamSynchronize(m_stream) == hipSuccess; }

bool Simulator::uploadBlock(uint32_t first, const std::vector<float4>& h, const char* where) {
    const size_t bytes = sizeof(float4) * h.size();
    if (hipStreamSynchronize(m_stream) != hipSuccess) {
        set_error(PBFX_E_INVALID_ARGUMENT, where);
        return false;
    }
    bool ok = copyOnStream(m_bufs.d_pos + first, h.data(), bytes, hipMemcpyHostToDevice) &&
              copyOnStream(m_bufs.d_pos_pred + first, h.data(), bytes, hipMemcpyHostToDevice) &&
              hipMemsetAsync(m_bufs.d_vel + first, 0, bytes, m_stream) == hipSuccess;
    ok = syncStream() && ok;
    if (!ok) set_error(PBFX_E_INVALID_ARGUMENT, where);
    return ok;
}

bool Simulator::seedCubeMix(uint32_t first, uint32_t groups, const float3* centers, uint32_t edge, float spacing,
                            float amp, uint64_t seed) {
    const uint64_t n64 = (uint64_t)groups * edge * edge * edge;
    if (!m_initialized || (groups && !centers) || first + n64 > m_bufs.capacity) {
        set_error(PBFX_E_CAPACITY, "seedCubeMix: blocks exceed capacity");
        return false;
    }
    std::vector<float4> h((size_t)n64);
    pbfx::cube_mix_fill(m_params.grid, groups, centers, edge, spacing, amp, seed, h.data());
    return uploadBlock(first, h, "seedCubeMix: copy failed");
}

bool Simulator::bindExternalPositions(float4* a, float4* b, uint32_t capacity) {
    if (!m_initialized || !a || !b || a == b || capacity < m_bufs.capacity) {
        set_error(PBFX_E_INVALID_ARGUMENT, "bindExternalPositions: two distinct buffers of >= capacity float4");
        return false;
    }
    const size_t bytes = sizeof(float4) * m_bufs.capacity;
    bool ok = hipMemcpyAsync(a, m_bufs.d_pos_curr, bytes, hipMemcpyDeviceToDevice, m_stream) == hipSuccess &&
              hipMemcpyAsync(b, m_bufs.d_pos_next, bytes, hipMemcpyDeviceToDevice, m_stream) == hipSuccess &&
              hipStreamSynchronize(m_stream) == hipSuccess;
    if (!ok) {
        set_error(PBFX_E_INVALID_ARGUMENT, "bindExternalPositions: copy failed");
        return false;
    }
    m_bufs.adoptExternalPositions(a, b);
    m_graphDirty = true;  // the captured launches hold the old pointers
    return true;
}

bool Simulator::setTuning(const Tuning& t) {
    const bool itemsOk = t.sortItems == 0 || t.sortItems == 4 || t.sortItems == 8 || t.sortItems == 16;
    const bool digitOk = t.sortDigitBits == 0 || (t.sortDigitBits >= 4 && t.sortDigitBits <= 8);
    const bool blockOk = t.walkBlock == 0 || t.walkBlock == 128 || t.walkBlock == 256;
    if (!m_initialized || t.streamRatio < 1 || !itemsOk || !digitOk || !blockOk) {
        set_error(PBFX_E_INVALID_ARGUMENT,
                  "setTuning: streamRatio >= 1, sortItems in {0, 4, 8, 16}, sortDigitBits 0 or 4..8, "
                  "walkBlock in {0, 128, 256}");
        return false;
    }
    m_grid.sortItems = t.sortItems;
    m_grid.sortDigitBits = t.sortDigitBits == 0 ? 8 : t.sortDigitBits;
    m_grid.candidateMask = t.candidateMask;
    m_grid.laneOrder = t.laneOrder;
    m_grid.spans.streamRatio = t.streamRatio;
    m_grid.spans.chunkBits = t.chunkBits ? 1u : 0u;
    m_grid.spans.block = t.walkBlock == 0 ? 256u : (uint32_t)t.walkBlock;
    m_graphDirty = true;
    return true;
}

bool Simulator::bindExternalVelocity(void* dst, size_t bytes, uint32_t strideBytes) {
    if (!m_initialized || !dst || bytes == 0 || strideBytes < sizeof(float4)) {
        set_error(PBFX_E_INVALID_ARGUMENT, "bindExternalVelocity: buffer, size and stride >= 16 needed");
        return false;
    }
    m_extVel = dst;
    m_extVelBytes = bytes;
    m_extVelStride = strideBytes;
    return true;
}

bool Simulator::publishVelocity(uint32_t count) {
    if (!m_initialized || !m_extVel) {
        set_error(PBFX_E_INVALID_ARGUMENT, "publishVelocity: no external velocity buffer bound");
        return false;
    }
    count = std::min<uint32_t>(count, m_params.numParticles);
    // rows that fit: the last one needs 16 bytes, the others a full stride
    uint64_t rows = m_extVelBytes < sizeof(float4) ? 0 : (m_extVelBytes - sizeof(float4)) / m_extVelStride + 1;
    rows = std::min<uint64_t>(rows, count);
    if (rows == 0) return true;
    const hipError_t e = hipMemcpy2DAsync(m_extVel, m_extVelStride, m_bufs.d_vel, sizeof(float4), sizeof(float4),
                                          (size_t)rows, hipMemcpyDeviceToDevice, m_stream);
    if (e != hipSuccess) {
        set_error((int)e, "publishVelocity");
        return false;
    }
    return true;
}

bool Simulator::setState(const float4* pos, const float4* vel, uint32_t N, bool fromDevice) {
    if (!m_initialized || N > m_bufs.capacity) {
        set_error(PBFX_E_CAPACITY, "setState: N exceeds capacity");
        return false;
    }
    const hipMemcpyKind k = fromDevice ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const size_t bytes = sizeof(float4) * N;
    bool ok = hipStreamSynchronize(m_stream) == hipSuccess;  // no step may still use the buffers
    if (pos) {
        ok = ok && copyOnStream(m_bufs.d_pos, pos, bytes, k);
        ok = ok && copyOnStream(m_bufs.d_pos_pred, pos, bytes, k);
    }
    if (vel) ok = ok && copyOnStream(m_bufs.d_vel, vel, bytes, k);
    ok = syncStream() && ok;
    if (!ok) set_error(PBFX_E_INVALID_ARGUMENT, "setState: copy failed");
    return ok;
}

bool Simulator::getState(float4* pos, float4* vel, bool toDevice) const {
    if (!m_initialized) return false;
    if (hipStreamSynchronize(m_stream) != hipSuccess) {
        set_error(PBFX_E_INVALID_ARGUMENT, "getState: stream error");
        return false;
    }
    const hipMemcpyKind k = toDevice ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    const size_t bytes = sizeof(float4) * m_params.numParticles;
    bool ok = true;
    if (pos) ok = ok && copyOnStream(pos, m_bufs.d_pos, bytes, k);
    if (vel) ok = ok && copyOnStream(vel, m_bufs.d_vel, bytes, k);
    return syncStream() && ok;
}

bool Simulator::densityStats(double* out5) {
    if (!m_initialized) return false;
    if (!m_scenes.empty()) {
        set_error(PBFX_E_INVALID_ARGUMENT, "densityStats: single-scene simulators only");
        return false;
    }
    const uint32_t N = m_params.numParticles;
    const hipStream_t s = m_stream;
    GridBuffers& g = m_grid;
    LaunchHashKeys(g.d_cellKeys, g.d_indices, m_bufs.d_pos, m_params.grid, N, s);
    pbfx::radix_sort_pairs(g.d_sortTemp, g.sortTempBytes, g.d_cellKeys, g.d_cellKeys_sorted, nullptr,
                           g.d_indices_sorted, N, 0, pbfx::key_bits_for(g.numCells), s);
    pbfx::compact_segments(g.d_cellUniqueKeys, g.d_cellOffsets, g.d_compactCount, g.d_segOf, g.d_cellKeys_sorted, N,
                           g.d_blockCounts, g.d_blockOffsets, s);
    pbfx::launch_spans_gather(pbfx::CellLookup::Compact, m_bufs.d_pos, g.d_indices_sorted, m_bufs.d_pos_sorted[0], N,
                              g.d_cellUniqueKeys, g.d_cellOffsets, g.d_compactCount, nullptr, nullptr, g.table,
                              g.spans, m_params.grid, g.scenes, s);
    return pbfx::density_stats_sorted(m_bufs.d_pos_sorted[0], m_bufs.d_vel, g.d_indices_sorted, g.d_segOf, g.spans,
                                      MakeDeviceParams(m_params), N, out5, s);
}

void Simulator::releaseAdjoint() {
    for (auto& t : m_tapes) {
        dfree(t.X); dfree(t.V); dfree(t.idxS); dfree(t.segOf); dfree(t.spans); dfree(t.Vs);
        dfree(t.predClamp); dfree(t.finalClamp); dfree(t.xsphGm); dfree(t.sceneOfSlot);
        for (auto& q : t.P) dfree(q);
        for (auto& q : t.lamA) dfree(q);
        for (auto& q : t.lamB) dfree(q);
        for (auto& q : t.dispU) dfree(q);
    }
    m_tapes.clear();
    dfree(m_adj.gPf); dfree(m_adj.gVo); dfree(m_adj.gVn); dfree(m_adj.gP); dfree(m_adj.a);
    dfree(m_adj.lamQ); dfree(m_adj.lamE); dfree(m_adj.partial); dfree(m_adj.gcSlot); dfree(m_adj.gmax);
    m_adj = pbfx::AdjointScratch{};
}

bool Simulator::adjointInit(int slots) {
    releaseAdjoint();
    if (!m_initialized || slots < 1 || m_opt.path != PathKind::Reordered ||
        m_params.solverIters < 1 || m_params.solverIters > pbfx::kMaxTapeIters) {
        set_error(PBFX_E_INVALID_ARGUMENT, "adjointInit: needs a reordered simulator, 1..8 iterations");
        return false;
    }
    const uint32_t cap = m_bufs.capacity;
    const size_t spanWords = (size_t)m_grid.spans.rows * m_grid.spans.stride;
    m_tapes.resize(slots);
    bool ok = true;
    for (auto& t : m_tapes) {
        ok = ok && dalloc(t.X, cap) && dalloc(t.V, cap) && dalloc(t.idxS, cap) && dalloc(t.segOf, cap) &&
             dalloc(t.spans, spanWords) && dalloc(t.Vs, (size_t)cap + pbfx::kSlotPad) && dalloc(t.predClamp, cap) &&
             dalloc(t.finalClamp, cap) && dalloc(t.xsphGm, cap);
        if (!m_scenes.empty()) ok = ok && dalloc(t.sceneOfSlot, cap);
        ok = ok && hipMemset(t.Vs, 0, sizeof(float4) * ((size_t)cap + pbfx::kSlotPad)) == hipSuccess;
        for (int it = 0; ok && it <= m_params.solverIters; ++it) {
            ok = dalloc(t.P[it], (size_t)cap + pbfx::kSlotPad) &&
                 hipMemset(t.P[it], 0, sizeof(float4) * ((size_t)cap + pbfx::kSlotPad)) == hipSuccess;
        }
        for (int it = 0; ok && it < m_params.solverIters; ++it)
            ok = dalloc(t.lamA[it], cap) && dalloc(t.lamB[it], cap) && dalloc(t.dispU[it], cap);
    }
    m_adj.blocks = pbfx::ceil_div(cap, 256);
    ok = ok && dalloc(m_adj.gPf, cap) && dalloc(m_adj.gVo, cap) && dalloc(m_adj.gVn, cap) && dalloc(m_adj.gP, cap) &&
         dalloc(m_adj.a, cap) && dalloc(m_adj.lamQ, cap) && dalloc(m_adj.lamE, cap) &&
         dalloc(m_adj.partial, m_adj.blocks) && dalloc(m_adj.gmax, std::max<size_t>(1, m_scenes.size()) * pbfx::kGmaxStride);
    if (!m_scenes.empty()) ok = ok && dalloc(m_adj.gcSlot, cap);
    if (!ok) {
        releaseAdjoint();
        return false;
    }
    ok = hipMemset(m_adj.gmax, 0, sizeof(uint32_t) * std::max<size_t>(1, m_scenes.size()) * pbfx::kGmaxStride) ==
         hipSuccess;
    if (!(null_stream_sync() && ok)) {
        releaseAdjoint();
        return false;
    }
    return true;
}

bool Simulator::stepRecord(int slot) {
    if (!m_initialized || slot < 0 || slot >= (int)m_tapes.size()) {
        set_error(PBFX_E_INVALID_ARGUMENT, "stepRecord: no such tape slot (adjointInit first)");
        return false;
    }
    pbfx::StepTape& t = m_tapes[slot];
    t.params = m_params;
    t.recorded = true;
    const uint32_t N = m_params.numParticles;
    const size_t b4 = sizeof(float4) * N;
    bool ok = hipMemcpyAsync(t.X, m_bufs.d_pos, b4, hipMemcpyDeviceToDevice, m_stream) == hipSuccess &&
              hipMemcpyAsync(t.V, m_bufs.d_vel, b4, hipMemcpyDeviceToDevice, m_stream) == hipSuccess;
    m_bufs.rec = t;
    runPipeline(m_stream);
    m_bufs.rec = pbfx::StepTape{};
    const size_t b1 = sizeof(uint32_t) * N;
    ok = ok && hipMemcpyAsync(t.idxS, m_grid.d_indices_sorted, b1, hipMemcpyDeviceToDevice, m_stream) == hipSuccess &&
         hipMemcpyAsync(t.segOf, m_grid.d_segOf, b1, hipMemcpyDeviceToDevice, m_stream) == hipSuccess &&
         hipMemcpyAsync(t.spans, m_grid.spans.spans, sizeof(uint2) * m_grid.spans.rows * m_grid.spans.stride,
                        hipMemcpyDeviceToDevice, m_stream) == hipSuccess;
    if (t.sceneOfSlot)
        ok = ok && hipMemcpyAsync(t.sceneOfSlot, m_grid.scenes.sceneOfSlot, b1, hipMemcpyDeviceToDevice,
                                  m_stream) == hipSuccess;
    if (!ok) {
        set_error(PBFX_E_INVALID_ARGUMENT, "stepRecord: tape copy failed");
        return false;
    }
    m_bufs.swapPositionPingPong();
    m_parity ^= 1;
    ++m_frameIndex;
    return true;
}

bool Simulator::backwardStep(int slot, const float4* gXin, const float4* gVin, float4* gXout, float4* gVout,
                             double* gcAccum, int32_t* log2Scale) {
    if (!m_initialized || slot < 0 || slot >= (int)m_tapes.size() || !m_tapes[slot].recorded) {
        set_error(PBFX_E_INVALID_ARGUMENT, "backwardStep: no step recorded in that tape slot");
        return false;
    }
    if (gXin == gXout || gVin == gVout) {
        set_error(PBFX_E_ALIAS, "backwardStep: outputs alias inputs");
        return false;
    }
    // the tape's own parameters: a parameter change after recording does not alter
    // which step is differentiated (the structural fields cannot change at all)
    const pbfx::StepTape& t = m_tapes[slot];
    pbfx::adjoint_step(t, m_grid.spans, t.params, m_grid.scenes, gXin, gVin, gXout, gVout, gcAccum, m_adj, m_stream,
                       log2Scale);
    return true;
}

bool Simulator::hashOverflowed() const {
    if (!m_grid.table.overflow) return false;
    uint32_t v = 0;
    if (!copyOnStream(&v, m_grid.table.overflow, sizeof(v), hipMemcpyDeviceToHost) || !syncStream()) return false;
    return v != 0;
}

}  // namespace sim
