// pbf_sim.h — C ABI of the headless simulator in libpbfx.so.
//
// A handle over sim::Simulator (reference: sim/simulator.h:47-90 —
// initialize / step / seedBoxLattice / devicePositions / deviceVel), minus the
// D3D12 interop.  State stays on the device between steps in the reference's
// layout (float4 AoS, original particle order); the position ping-pong swap
// happens after every step.
#pragma once

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime.h>

#include "pbf_params.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct PbfSim PbfSim;

#define PBF_GRID_DENSE 0    /* cellStart/cellEnd over every cell (reference: dense grid) */
#define PBF_GRID_COMPACT 1  /* unique keys + offsets, binary search (reference: "hashed" grid) */
#define PBF_GRID_HASHED 2   /* open-addressing cell table, 2^tableLog2 slots */

#define PBF_PATH_REORDERED 0  /* cell-sorted solver (default) */
#define PBF_PATH_REFERENCE 1  /* phases over the drop-in Launch* functions */

// reference: Simulator::initialize (simulator.cpp:217).  Returns NULL on error
// (see pbfGetLastError).  `params->grid.dim` must be set.
PbfSim* pbfSimCreate(const sim::SimParams* params, int gridKind, int tableLog2, int path, int useGraph);
void pbfSimDestroy(PbfSim* sim);

// reference: Simulator::seedBoxLattice (simulator.cpp:808-878), placed at particle offset `first`,
// with a per-axis jitter of +-jitterAmp drawn from SplitMix64(seed, particle, axis).
int pbfSimSeedLattice(PbfSim* sim, uint32_t first, uint32_t nx, uint32_t ny, uint32_t nz, float ox, float oy,
                      float oz, float spacing, float jitterAmp, uint64_t seed);
// Copy state in/out (host or device pointers, N = params->numParticles on output).
// Both wait for the simulator's stream, copy on it and wait again, so the copy is
// complete on return and ordered with the steps before and after it; device memory
// written on another stream must be complete before the call.
int pbfSimSetState(PbfSim* sim, const float4* pos, const float4* vel, uint32_t N, int fromDevice);
int pbfSimGetState(PbfSim* sim, float4* pos, float4* vel, int toDevice);

// reference: Simulator::step (simulator.cpp:560-643).  Enqueues `steps` steps on the
// simulator's stream (graph replay when enabled) without synchronising.
int pbfSimStep(PbfSim* sim, int steps);
// New parameters for the following steps (reference: the parameter handling of
// Simulator::step, simulator.cpp:575-607, ParamChangeTracker param_change_tracker.h:
// 23-61).  Coefficients, gravity, dt, clamps and the neighbour cap re-capture the
// step.  A new particle count, grid (box, cell size, dims), smoothing length or
// iteration count re-allocates what depends on it (the particle buffers only when
// they must grow) and keeps the state of the particles that remain; particles added
// start at rest at the origin until seeded.  Refused (-1, PBFX_E_*): structural
// changes of an ensemble or while adjoint tapes are held, growth beyond bound external
// position buffers, a hash table smaller than the particle count.
int pbfSimSetParams(PbfSim* sim, const sim::SimParams* params);
// The parameters the simulator steps with (for an ensemble: the shared ones, with the
// ensemble's particle count and key space), as the base for pbfSimSetParams.
int pbfSimGetParams(PbfSim* sim, sim::SimParams* out);
int pbfSimSynchronize(PbfSim* sim);
hipStream_t pbfSimStream(PbfSim* sim);
const float4* pbfSimDevicePositions(PbfSim* sim);
const float4* pbfSimDeviceVelocities(PbfSim* sim);
// Sorted keys / indices of the last grid build (device pointers).
const uint32_t* pbfSimDeviceSortedKeys(PbfSim* sim);
const uint32_t* pbfSimDeviceSortedIndices(PbfSim* sim);

// One eager step with an event around every phase.  Writes up to maxPhases
// times (ms) and NUL-separated names into names (namesBytes long); returns the phase count.
int pbfSimProfilePhases(PbfSim* sim, float* ms, int maxPhases, char* names, int namesBytes);
// Density-constraint statistics of the current state (grid rebuilt on the current
// positions, lambda-pass neighbour order and cap): out5 = {mean(rho/rho0-1),
// max(rho/rho0-1), mean neighbours, mean |v|, mean(max(rho/rho0-1, 0))}.
// Synchronises the simulator's stream.
int pbfSimDensityStats(PbfSim* sim, double* out5);
// 1 when a hash-table insert found no free slot since creation.
int pbfSimHashOverflow(PbfSim* sim);

// ---- differentiable rollout (reordered simulators and ensembles) ------------
// Hand-written adjoint of the step.  A recorded step keeps the sort permutation,
// neighbour spans, clamp bits and every solver iteration's positions and lambdas;
// the reverse pass differentiates the step with those decisions frozen.  No
// reference counterpart (new form).
// Allocates `tapeSlots` recorded steps plus the reverse-pass scratch.
int pbfSimAdjointInit(PbfSim* sim, int tapeSlots);
// One eager forward step recorded into tape `slot` (state advances as pbfSimStep).
int pbfSimStepRecord(PbfSim* sim, int slot);
// Reverse of the step recorded in `slot`.  gXin/gVin = dL/dX_{t+1}, dL/dV_{t+1};
// gXout/gVout = dL/dX_t, dL/dV_t (device float4, particle order, w ignored; outputs
// must not alias inputs).  Adds dL/d(xsph_c) of the step to gcAccum (device doubles:
// one for a simulator, one per scene for an ensemble handle).
int pbfSimBackwardStep(PbfSim* sim, int slot, const float4* gXin, const float4* gVin, float4* gXout, float4* gVout,
                       double* gcAccum);
// The same reverse step, renormalised, for long rollouts whose gradients outgrow
// float32.  log2Scale holds device int32 exponents E, one per scene for an ensemble
// handle (else one): on entry the true dL/dX_{t+1}, dL/dV_{t+1} of scene q are its
// entries of gXin, gVin times 2^E[q].  The step adds its true dL/dc (times 2^E, in
// double) to gcAccum, then writes gXout, gVout as dL/dX_t, dL/dV_t times a power of two
// per scene, chosen so the scene's largest magnitude lies in [0.5, 1), and adds that
// power's exponent to E[q].  Exact: every kernel is linear in the incoming gradients
// and the rescale is a power-of-two product.  No reference counterpart (new form).
int pbfSimBackwardStepScaled(PbfSim* sim, int slot, const float4* gXin, const float4* gVin, float4* gXout,
                             float4* gVout, double* gcAccum, int32_t* log2Scale);

// ---- batched ensembles ----------------------------------------------------
// Independent scenes packed into one particle array and one cell-key space and
// advanced by the same launches (one per phase for all scenes).  `base` holds the
// shared constants (h, rest density, mass, dt, solver iterations, PBF tuning,
// neighbour cap, cell size in base->grid.cellSize); each scene has its own box,
// gravity, XSPH coefficient, restitution and particle count.  Scene s occupies
// particles [first_s, first_s + numParticles_s) in scene order.  Every scene
// evolves bit for bit as it would alone.  No reference counterpart (new form).
typedef struct {
    float boxMin[3];
    float boxMax[3];
    float gravity[3];
    float xsph_c;
    float restitution;
    uint32_t numParticles;
} PbfSceneDesc;
PbfSim* pbfEnsembleCreate(const sim::SimParams* base, const PbfSceneDesc* scenes, uint32_t count, int gridKind,
                          int tableLog2, int useGraph);
int pbfEnsembleSeedLattice(PbfSim* sim, uint32_t scene, uint32_t nx, uint32_t ny, uint32_t nz, float ox, float oy,
                           float oz, float spacing, float jitterAmp, uint64_t seed);
int pbfEnsembleGetScene(PbfSim* sim, uint32_t scene, float4* pos, float4* vel);
// first particle, particle count and grid dims (3 ints) of a scene
int pbfEnsembleSceneInfo(PbfSim* sim, uint32_t scene, uint32_t* first, uint32_t* count, int* dims3);

// ---- launch tuning --------------------------------------------------------
// Knobs of the reordered solver that change speed, never results (every setting
// gives the same bits).  No reference counterpart; picked per workload by the
// autotuner (pbfx/tune.py).
typedef struct {
    int candidateMask;     // 1: a pass before the first iteration records which candidates can
                           // come within h this step; the lambda passes walk only those
    uint32_t streamRatio;  // a wave streams a neighbour row through LDS when the union of its
                           // lanes' spans is <= streamRatio x the longest span (>= 1)
    int chunkBits;         // streamed rows: 1 = test a 64-candidate chunk into a bit mask, then
                           // queue the set bits; 0 = test and queue candidate by candidate
    int sortItems;         // radix sort keys per thread: 4, 8 or 16 (0 = chosen by N)
    int sortDigitBits;     // radix sort: widest digit of a pass, 4..8 bits (0 = 8); the key bits
                           // are spread evenly over ceil(bits / sortDigitBits) passes
    int walkBlock;         // lanes per block of the neighbour kernels: 128 or 256 (0 = 256)
                           // (pbfSimGetTuning reports the effective values: 8 and 256 for 0)
    int laneOrder;         // 1: the lambda and delta passes deal each 128-slot group to its two
                           // waves by the particles' place inside their cells, so a wave's
                           // lanes meet their neighbours together; 0: lane i takes slot i
} PbfSimTuning;
int pbfSimSetTuning(PbfSim* sim, const PbfSimTuning* tuning);
int pbfSimGetTuning(PbfSim* sim, PbfSimTuning* tuning);

// ---- scene generators -----------------------------------------------------
// reference: Simulator::seedBoxLatticeAuto (simulator.cpp:882-941).  Host only.  The
// lattice dims (nx, ny, nz) for about `total` particles at `spacing` in the box, and
// originInOut pulled 0.25 spacing inside it.  The reference then reseeds its
// simulator with nx*ny*nz particles; here the caller creates the simulator with
// that count and seeds it with pbfSimSeedLattice.
int pbfLatticeAutoDims(uint32_t total, const float* boxMin3, const float* boxMax3, float spacing,
                       float* originInOut3, uint32_t* dimsOut3);
// reference: the CubeMix settings of RuntimeConsole::Simulation (console.h:184-208).
typedef struct {
    uint32_t groupCount;
    uint32_t layers;
    float groupSpacing;   // x/z distance between group centres
    float layerSpacing;   // y distance between layers
    float baseHeight;     // y of the first layer
    int layerJitter;      // per-layer y offset of +-layerJitterAmp
    float layerJitterAmp;
    uint32_t layerJitterSeed;
} PbfCubeMixLayout;
// reference: GenerateCubeMixCenters (console.cpp:277-316).  Host only; writes
// 3 * groupCount floats.  The layer jitter draws from SplitMix64(seed ^ layer)
// instead of std::mt19937.
int pbfCubeMixCenters(const PbfCubeMixLayout* layout, float* centersOut);
// reference: Simulator::seedCubeMix (simulator.cpp:920-975), at particle offset
// `first`: groupCount blocks of edge^3 particles centred on centers (3 floats each),
// x fastest, clamped to the box, w = group index, velocity 0, per-axis jitter of
// +-jitterAmp from SplitMix64(seed, particle, axis) (the reference: unit-ball rejection
// from std::mt19937).
int pbfSimSeedCubeMix(PbfSim* sim, uint32_t first, uint32_t groupCount, const float* centers,
                      uint32_t edgeParticles, float spacing, float jitterAmp, uint64_t seed);

// ---- external buffers -----------------------------------------------------
// reference: Simulator::bindExternalPosPingPong (simulator.cpp:676-757), with device
// pointers in place of imported D3D12 resources.  The simulator adopts posA (current)
// and posB (next) as its position ping-pong; each holds >= capacity float4 (capacity
// >= numParticles).  The current state is copied into them; the simulator's own
// position buffers are released; the buffers stay owned by the caller and must
// outlive the simulator or the next bind.  Which of the two holds the current state
// after a step is given by pbfSimDevicePositions.
int pbfSimBindExternalPositions(PbfSim* sim, float4* posA, float4* posB, uint32_t capacity);
// reference: bindExternalVelocityBuffer / publishExternalVelocity (simulator.cpp:
// 760-802, 1069-1079).  publish copies `count` velocities (float4, original particle
// order) to dst, one every strideBytes (>= 16), clipped to `bytes`, on the
// simulator's stream.
int pbfSimBindExternalVelocity(PbfSim* sim, void* dst, size_t bytes, uint32_t strideBytes);
int pbfSimPublishVelocity(PbfSim* sim, uint32_t count);

// ABI self-checks (host only): copy by-value structs back out, so a foreign
// caller can verify it passes them with the same layout and calling convention.
void pbfDebugEchoDeviceParams(sim::DeviceParams dp, sim::DeviceParams* out);
void pbfDebugEchoGridBounds(sim::GridBounds g, sim::GridBounds* out);
void pbfDebugEchoFloat3(float3 v, float* out3);

// Arithmetic self-checks (device; no reference counterpart).  The neighbour kernels
// evaluate the pair square roots and divisions without the compiler's range-scaling
// steps when the parameters bound every operand away from the ends of the exponent
// range, which returns the same bits as IEEE sqrtf and division.
// pbfDebugPairUnscaled: 1 when the kernels take that form for dp, 0 when they take
// the general one, -1 on error.
int pbfDebugPairUnscaled(sim::DeviceParams dp);
// pbfDebugCheckUnscaledMath: counts where the unscaled sequences differ from IEEE,
// out5 = {square roots over every finite float >= 2^-96, roots of smaller inputs
// that are not in [+0, 2^-47], quotients over divSamples pseudo-random pairs (seeded)
// in the admitted operand range, zero numerators whose quotient is not +0, integer
// max results that differ from fmaxf}.  Synchronous; 0 on success, -1 on error.
int pbfDebugCheckUnscaledMath(uint64_t divSamples, uint64_t seed, unsigned long long* out5);
// pbfDebugScorrRatio: the pair-arithmetic check a simulator runs whenever its parameters
// change, for dp: out12 = {zh, zl, ok, w_q, poly6, h2, -scorr_k, scorr_min, rsq_ok, h,
// grad_r_eps, -3 spiky}, where ok is 1 when fma(W, zh, W * zl) gives the division's s_corr
// term for every float W from +0 to poly6 h2^3, and rsq_ok is 1 when the spiky gradient's
// coefficient with its reciprocal started from the square root's rsq equals the one
// started from v_rcp for every float r2 from +0 to h2 (with both at 1 the delta kernels take both),
// else 0.  Synchronous; 0 on success, -1 on error.
int pbfDebugScorrRatio(sim::DeviceParams dp, float* out12);

#ifdef __cplusplus
}
#endif
