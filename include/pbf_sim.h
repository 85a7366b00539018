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
int pbfSimSetState(PbfSim* sim, const float4* pos, const float4* vel, uint32_t N, int fromDevice);
int pbfSimGetState(PbfSim* sim, float4* pos, float4* vel, int toDevice);

// reference: Simulator::step (simulator.cpp:560-643).  Enqueues `steps` steps on the
// simulator's stream (graph replay when enabled) without synchronising.
int pbfSimStep(PbfSim* sim, int steps);
// New parameters for the following steps (same grid and particle count).
int pbfSimSetParams(PbfSim* sim, const sim::SimParams* params);
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

// ---- differentiable rollout (single-scene reordered simulators) -------------
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
// must not alias inputs).  Adds dL/d(xsph_c) of the step to *gcAccum (device double).
int pbfSimBackwardStep(PbfSim* sim, int slot, const float4* gXin, const float4* gVin, float4* gXout, float4* gVout,
                       double* gcAccum);

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

// ABI self-checks (host only): copy by-value structs back out, so a foreign
// caller can verify it passes them with the same layout and calling convention.
void pbfDebugEchoDeviceParams(sim::DeviceParams dp, sim::DeviceParams* out);
void pbfDebugEchoGridBounds(sim::GridBounds g, sim::GridBounds* out);
void pbfDebugEchoFloat3(float3 v, float* out3);

#ifdef __cplusplus
}
#endif
