// sim_host.h — headless host orchestration of the PBF step (C++17).
//
// Mirrors the reference's sim/ phase API: DeviceBuffers / GridBuffers
// (reference: sim/device_buffers.cuh:17-47, sim/grid_buffers.cuh:6-85),
// SimulationContext (sim/simulation_context.h:6-18), IPhase / PhasePipeline
// (sim/phase_pipeline.h:10-40), post ops (sim/post_ops.h:8-38), the kernel
// dispatcher registry (sim/kernel_dispatcher.h:11-36) and Simulator
// (sim/simulator.h:47-90) without the D3D12 interop.
//
// Two pipelines are registered in the dispatcher:
//   * Reference layout — the phases call the drop-in Launch* functions exactly
//     as sim/phases.cpp and sim/post_ops.cpp do (XSPH reads a snapshot);
//   * Reordered (default) — the same step with the solver working on
//     cell-sorted copies: predict+key, sort, compact segments, cell lookup
//     (dense ranges / hash table / binary search), neighbour-row spans, then
//     lambda / delta (last one fused with boundary + velocity) and XSPH fused
//     with the final boundary and the scatter back to particle order.
// Both advance state in original particle order and swap the position
// ping-pong after each step (reference: sim/simulator.cpp:144-159).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "pbf_params.h"
#include "adjoint.hpp"
#include "sorted_path.hpp"

namespace sim {

enum class GridKind : int { Dense = 0, Compact = 1, Hashed = 2 };
enum class PathKind : int { Reordered = 0, Reference = 1 };

struct DeviceBuffers {
    float4* d_pos = nullptr;       // current positions (alias of d_pos_curr)
    float4* d_vel = nullptr;
    float4* d_pos_pred = nullptr;  // predicted / next positions (alias of d_pos_next)
    float* d_lambda = nullptr;
    float4* d_delta = nullptr;
    float4* d_vel_snapshot = nullptr;  // Jacobi XSPH input
    float4* d_pos_curr = nullptr;
    float4* d_pos_next = nullptr;
    float4* d_pos_sorted[2] = {nullptr, nullptr};  // cell-sorted working copies
    float4* d_vel_sorted = nullptr;
    // Recording a step for the adjoint: when rec.P[0] is set, iteration it works in
    // rec.P[it] (so every iteration's positions and lambdas survive), velocities in
    // rec.Vs, and the clamps write their bits.
    pbfx::StepTape rec{};
    uint32_t capacity = 0;
    bool externalPos = false;  // d_pos_curr / d_pos_next belong to the caller

    bool allocate(uint32_t cap);
    void adoptExternalPositions(float4* a, float4* b);
    void release();
    void swapPositionPingPong();
    ~DeviceBuffers() { release(); }
};

struct GridBuffers {
    uint32_t* d_cellKeys = nullptr;
    uint32_t* d_cellKeys_sorted = nullptr;
    uint32_t* d_indices = nullptr;
    uint32_t* d_indices_sorted = nullptr;
    uint32_t* d_cellStart = nullptr;
    uint32_t* d_cellEnd = nullptr;
    uint32_t numCells = 0;
    uint32_t* d_cellUniqueKeys = nullptr;
    uint32_t* d_cellOffsets = nullptr;
    uint32_t* d_compactCount = nullptr;
    uint32_t* d_segOf = nullptr;
    uint32_t* d_blockCounts = nullptr;
    uint32_t* d_blockOffsets = nullptr;
    void* d_sortTemp = nullptr;
    size_t sortTempBytes = 0;
    int keyBits = 32;
    pbfx::SpanTable spans{};
    pbfx::CandMask mask{};  // per-step candidate mask (reordered path)
    bool candidateMask = true;  // later passes walk the mask (tuning knob)
    uint32_t* d_laneOrder = nullptr;  // lane -> slot of the lambda and delta passes (reordered path)
    uint2* d_slotPairs = nullptr;     // each cell's slots two at a time, for XSPH (reordered path)
    uint32_t* d_slotPairCount = nullptr;
    // the step's flags word (reordered path; bits pbfx::kStepVelocityUnbounded and
    // pbfx::kStepNanPosition): cleared by the predict kernel, set by the solver passes
    uint32_t* d_stepFlags = nullptr;
    // the s_corr ratio check's record and scratch word (reordered path; launch_ratio_check)
    pbfx::RatioCoef* d_ratioCoef = nullptr;
    uint32_t* d_ratioBad = nullptr;
    bool laneOrder = true;            // use d_laneOrder (tuning knob)
    int sortItems = 0;          // radix sort keys per thread (0: by N; tuning knob)
    int sortDigitBits = 8;      // radix sort widest digit of a pass (4..8 bits; tuning knob)
    pbfx::HashTable table{};
    pbfx::Scenes scenes{};  // ensemble view (empty for one scene)
    uint32_t tableLog2 = 0;
    uint32_t capacity = 0;
    GridKind kind = GridKind::Dense;

    bool allocate(uint32_t cap, uint32_t numCells, GridKind kind, uint32_t tableLog2, int reach, bool reordered);
    void release();
    ~GridBuffers() { release(); }
};

class KernelDispatcher;

struct SimulationContext {
    DeviceBuffers* bufs = nullptr;
    GridBuffers* grid = nullptr;
    bool useHashedGrid = false;  // compact or hashed lookup (reference: simulation_context.h:9)
    PathKind path = PathKind::Reordered;
    KernelDispatcher* dispatcher = nullptr;
    bool pingPongPos = true;
};

enum class PhaseType { Integrate, GridBuild, SortPairs, CellRanges, SolveIterations, Velocity, Post };

// Dispatcher registry (reference: kernel_dispatcher.h:11-36): one launch function
// per (phase, grid kind, path).  Lookup falls back from the exact grid kind to
// the compact variant, then the dense one.
using KernelLaunchFn = void (*)(DeviceBuffers&, GridBuffers&, const SimParams&, hipStream_t);
struct KernelVariant {
    GridKind grid = GridKind::Dense;
    PathKind path = PathKind::Reordered;
};
class KernelDispatcher {
public:
    void registerKernel(PhaseType ph, KernelVariant v, KernelLaunchFn fn) { m_map[key(ph, v)] = fn; }
    KernelLaunchFn find(const KernelVariant& v, PhaseType ph) const;
    size_t size() const { return m_map.size(); }

private:
    static uint32_t key(PhaseType ph, KernelVariant v) {
        return ((uint32_t)ph << 8) | ((uint32_t)v.grid << 4) | (uint32_t)v.path;
    }
    std::unordered_map<uint32_t, KernelLaunchFn> m_map;
};
void InitializeKernelDispatcher(KernelDispatcher& kd);

class IPhase {
public:
    virtual ~IPhase() = default;
    virtual PhaseType type() const = 0;
    virtual const char* name() const = 0;
    virtual void run(SimulationContext& ctx, const SimParams& p, hipStream_t s) = 0;
};

class PhasePipeline {
public:
    void add(std::unique_ptr<IPhase> ph) { m_phases.push_back(std::move(ph)); }
    void clear() { m_phases.clear(); }
    void runFull(SimulationContext& ctx, const SimParams& p, hipStream_t s) {
        for (auto& ph : m_phases) ph->run(ctx, p, s);
    }
    const std::vector<std::unique_ptr<IPhase>>& phases() const { return m_phases; }

private:
    std::vector<std::unique_ptr<IPhase>> m_phases;
};

// Builds the default phase list for a path (reference: phases.cpp:108-117 and
// post_ops.cpp:73-78: integrate, grid build, solve, velocity, XSPH, boundary).
void BuildDefaultPipelines(PhasePipeline& pipeline, PathKind path);

// One scene of a batched ensemble (see Simulator::initializeEnsemble).
struct SceneDesc {
    float3 boxMin;
    float3 boxMax;
    float3 gravity;
    float xsph_c;
    float restitution;
    uint32_t numParticles;
};

struct SimOptions {
    GridKind grid = GridKind::Dense;
    PathKind path = PathKind::Reordered;
    uint32_t tableLog2 = 22;
    bool useGraph = true;
};

class Simulator {
public:
    Simulator() = default;
    ~Simulator() { shutdown(); }
    Simulator(const Simulator&) = delete;
    Simulator& operator=(const Simulator&) = delete;

    bool initialize(const SimParams& p, const SimOptions& opt);
    // Batched ensemble: independent scenes packed into one particle array and one
    // key space, advanced by the same launches (one per phase for all scenes).
    // `base` gives the shared constants (h, rest density, mass, dt, iterations,
    // tuning, cell size); each scene has its own box, gravity, XSPH coefficient,
    // restitution and particle count.  Reordered path only.
    bool initializeEnsemble(const SimParams& base, const std::vector<SceneDesc>& scenes, const SimOptions& opt);
    void shutdown();
    // One step with the given parameters (reference: simulator.cpp:560-643).
    bool step(const SimParams& p);
    // New parameters for the following steps (same particle count and grid).
    bool setParams(const SimParams& p);
    bool stepMany(int steps);
    // Per-phase device times of one eager step, in milliseconds.
    bool profilePhases(std::vector<std::pair<std::string, float>>& out);

    // Lattice seeding (reference: simulator.cpp:808-878): block of nx*ny*nz at
    // particle offset `first`, x fastest, clamped to the box, then a per-axis
    // jitter of +-jitterAmp from SplitMix64(seed, particle, axis).  Zero velocity.
    bool seedBoxLattice(uint32_t first, uint32_t nx, uint32_t ny, uint32_t nz, float3 origin, float spacing,
                        float jitterAmp, uint64_t seed);
    bool seedSceneLattice(uint32_t scene, uint32_t nx, uint32_t ny, uint32_t nz, float3 origin, float spacing,
                          float jitterAmp, uint64_t seed);
    bool getSceneState(uint32_t scene, float4* pos, float4* vel) const;
    // CubeMix seeding (reference: simulator.cpp:920-975) at particle offset `first`.
    bool seedCubeMix(uint32_t first, uint32_t groups, const float3* centers, uint32_t edge, float spacing,
                     float jitterAmp, uint64_t seed);
    // External buffers (reference: simulator.cpp:676-802, 1069-1079) as device pointers.
    bool bindExternalPositions(float4* a, float4* b, uint32_t capacity);
    bool bindExternalVelocity(void* dst, size_t bytes, uint32_t strideBytes);
    bool publishVelocity(uint32_t count);
    // Launch-shape knobs of the reordered solver (the fields of PbfSimTuning).
    struct Tuning {
        bool candidateMask;
        uint32_t streamRatio;
        bool chunkBits;
        int sortItems;
        int sortDigitBits;
        int walkBlock;
        bool laneOrder;
    };
    bool setTuning(const Tuning& t);
    Tuning tuning() const {
        return Tuning{m_grid.candidateMask, m_grid.spans.streamRatio, m_grid.spans.chunkBits != 0,
                      m_grid.sortItems,     m_grid.sortDigitBits,     (int)m_grid.spans.block,
                      m_grid.laneOrder};
    }
    uint32_t sceneCount() const { return (uint32_t)m_scenes.size(); }
    const pbfx::SceneParams* sceneParams(uint32_t scene) const {
        return scene < m_scenes.size() ? &m_scenes[scene] : nullptr;
    }
    bool setState(const float4* pos, const float4* vel, uint32_t N, bool fromDevice);
    bool getState(float4* pos, float4* vel, bool toDevice) const;

    const SimParams& params() const { return m_params; }
    hipStream_t stream() const { return m_stream; }
    const float4* devicePositions() const { return m_bufs.d_pos; }
    const float4* deviceVel() const { return m_bufs.d_vel; }
    const DeviceBuffers& buffers() const { return m_bufs; }
    const GridBuffers& grid() const { return m_grid; }
    uint64_t frameIndex() const { return m_frameIndex; }
    bool hashOverflowed() const;
    // Density-constraint statistics of the current state (see pbfSimDensityStats).
    bool densityStats(double* out5);

    // ---- differentiable rollout (adjoint of the reordered step) ----
    // Allocates `slots` step tapes and the reverse-pass scratch.
    bool adjointInit(int slots);
    // One eager forward step recorded into tape `slot`.
    bool stepRecord(int slot);
    // Reverse of the step in tape `slot` (device pointers, particle order), with the
    // parameters the step was recorded with; adds dL/dc into gcAccum (device double).
    // log2Scale (device int, or nullptr): the renormalised form of adjoint_step.
    bool backwardStep(int slot, const float4* gXin, const float4* gVin, float4* gXout, float4* gVout,
                      double* gcAccum, int32_t* log2Scale = nullptr);

private:
    bool restructure(const SimParams& p);
    // re-checks the two-instruction s_corr ratio for m_params (reordered path)
    void checkScorrRatio();
    bool captureGraphs();
    void destroyGraphs();
    void runPipeline(hipStream_t s);
    bool initCommon(const SimParams& p, const SimOptions& opt, uint64_t cells);
    bool fillLattice(const GridBounds& g, uint32_t first, uint32_t nx, uint32_t ny, uint32_t nz, float3 o,
                     float spacing, float amp, uint64_t seed);
    bool uploadBlock(uint32_t first, const std::vector<float4>& h, const char* where);
    // Copies enqueued on the simulator's stream.  That stream does not wait for the
    // null stream, and a blocking hipMemcpy between device buffers (or a pageable
    // upload, or hipMemset) may return before its data lands, so state moves in and
    // out through these and the caller then synchronises the stream.
    bool copyOnStream(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) const;
    bool syncStream() const;
    void* m_extVel = nullptr;
    size_t m_extVelBytes = 0;
    uint32_t m_extVelStride = 0;
    void releaseScenes();

    std::vector<pbfx::StepTape> m_tapes;
    pbfx::AdjointScratch m_adj{};
    void releaseAdjoint();

    std::vector<pbfx::SceneParams> m_scenes;
    pbfx::SceneParams* d_sceneParams = nullptr;
    uint32_t* d_sceneOf = nullptr;
    uint32_t* d_sceneOfSlot = nullptr;
    uint32_t* d_cellBase = nullptr;

    SimParams m_params{};
    SimOptions m_opt{};
    DeviceBuffers m_bufs;
    GridBuffers m_grid;
    KernelDispatcher m_dispatcher;
    PhasePipeline m_pipeline;
    SimulationContext m_ctx;
    hipStream_t m_stream = nullptr;
    hipGraphExec_t m_exec[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> m_eventPool;  // events of profilePhases, reused step after step
    int m_parity = 0;
    uint64_t m_frameIndex = 0;
    bool m_graphDirty = true;
    bool m_initialized = false;
};

// Host scene seeding helpers shared with the C ABI.
float JitterUniform(uint64_t seed, uint64_t particle, uint32_t axis);

}  // namespace sim
