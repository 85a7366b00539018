// pbf_hip.h — C ABI of the gfx950 PBF solver core (libpbfx.so).
//
// Drop-in launchers for the sim/ phase pipeline: each declaration below has the
// signature of the reference's extern "C" launcher it replaces, with
// cudaStream_t -> hipStream_t.  Pointers are device pointers; particle arrays
// are float4 AoS (x, y, z, w) in original particle order; GridBounds and
// DeviceParams are passed by value (include/pbf_params.h).
//
// Conventions kept from the reference:
//   * launchers return void, never allocate (except EnsureCellCompactScratch,
//     which sizes the compaction scratch before graph capture, and
//     EnsureDropInStaging below), never
//     synchronise, and are safe to record into a hipGraph;
//   * N == 0, null state pointers (LaunchIntegratePred) and xsph_c <= 0
//     (LaunchXSPH*) return early without launching.
// Deliberate changes (the reference's behaviour is racy there):
//   * LaunchLambda* write `lambda` and read `pos_pred` from their arguments
//     (the reference ignores both and uses device globals, pbf_lambda.cu:256);
//   * LaunchXSPH* require vel_out != vel_in (Jacobi update; the reference
//     smooths in place, post_ops.cpp:59,65).
// Errors (launch failures, invalid arguments) are recorded and returned by
// pbfGetLastError(); LaunchCellRanges no longer aborts the process.
#pragma once

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime.h>

#include "pbf_params.h"

#ifdef __cplusplus
extern "C" {
#endif

// ---- integrate / boundary / velocity ---------------------------------------
// reference: sim/integrate.cu:76-89 (KIntegratePred :29-52)
void LaunchIntegratePred(float4* posCurr, const float4* velSrc, float4* posPred, float3 gravity, float dt,
                         uint32_t N, hipStream_t s);
// reference: sim/boundary.cu:37-40 (KBoundary :11-25)
void LaunchBoundary(float4* pos_pred, float4* vel, sim::GridBounds grid, float restitution, uint32_t N,
                    hipStream_t s);
// reference: sim/boundary.cu:41-44 (KVelocity :28-34)
void LaunchVelocity(float4* vel, const float4* pos, const float4* pos_pred, float inv_dt, uint32_t N,
                    hipStream_t s);

// ---- grid build --------------------------------------------------------------
// reference: sim/hash_keys.cu:30-33 (KHash :14-27)
void LaunchHashKeys(uint32_t* keys, uint32_t* indices, const float4* pos, sim::GridBounds grid, uint32_t N,
                    hipStream_t s);
// reference: sim/sort_pairs.cu:5-16.  Stable sort of (key, value) pairs by the
// full 32-bit key; inputs are not modified.
void LaunchSortPairsQuery(size_t* tempBytes, const uint32_t* d_keys_in, uint32_t* d_keys_out,
                          const uint32_t* d_vals_in, uint32_t* d_vals_out, uint32_t N, hipStream_t s);
// reference: sim/sort_pairs.cu:18-27
void LaunchSortPairs(void* d_temp_storage, size_t tempBytes, uint32_t* d_keys_in, uint32_t* d_keys_out,
                     uint32_t* d_vals_in, uint32_t* d_vals_out, uint32_t N, hipStream_t s);
// reference: sim/cell_ranges.cu:54-80.  start = 0xFFFFFFFF / end = 0 for empty cells.
void LaunchCellRanges(uint32_t* cellStart, uint32_t* cellEnd, const uint32_t* keysSorted, uint32_t N,
                      uint32_t numCells, hipStream_t s);
// reference: sim/cell_compact.cu:58-93.  Sizes the compaction scratch for N
// particles; call before capturing LaunchCellRangesCompact into a graph.
bool EnsureCellCompactScratch(uint32_t N, uint32_t threadsPerBlock);
// reference: sim/cell_compact.cu:187-231.  uniqueKeys[0..M), offsets[0..M] with
// offsets[M] = N, *compactCount = M.
void LaunchCellRangesCompact(uint32_t* d_uniqueKeys, uint32_t* d_offsets, uint32_t* d_compactCount,
                             const uint32_t* d_keysSorted, uint32_t N, hipStream_t s);
// No counterpart in the reference: LaunchCellRangesCompact for callers that cannot
// allocate (code under graph capture without a prior EnsureCellCompactScratch, tight
// loops, harnesses that forbid hipMalloc).  Same outputs; the compaction scratch is a
// static device array of the library (512 KiB per GPU, held until the process ends),
// so it allocates nothing.  N <= 2^26 sorted keys (larger N records
// PBFX_E_CAPACITY and launches nothing).  Calls that overlap on one GPU share that
// array and race: order them on one stream.
void LaunchCellRangesCompactNoAlloc(uint32_t* d_uniqueKeys, uint32_t* d_offsets, uint32_t* d_compactCount,
                                    const uint32_t* d_keysSorted, uint32_t N, hipStream_t s);

// ---- neighbour solves --------------------------------------------------------
// reference: sim/pbf_lambda.cu:246-258 (KLambdaGlobalsDense :10-108)
void LaunchLambda(float* lambda, const float4* pos_pred, const uint32_t* indicesSorted, const uint32_t* keysSorted,
                  const uint32_t* cellStart, const uint32_t* cellEnd, sim::DeviceParams dp, uint32_t N,
                  hipStream_t s);
// reference: sim/pbf_lambda.cu:260-274 (KLambdaGlobalsCompact :110-210)
void LaunchLambdaCompact(float* lambda, const float4* pos_pred, const uint32_t* indicesSorted,
                         const uint32_t* keysSorted, const uint32_t* uniqueKeys, const uint32_t* offsets,
                         const uint32_t* compactCount, sim::DeviceParams dp, uint32_t N, hipStream_t s);
// reference: sim/pbf_delta_apply.cu:239-267 (KDeltaCompute :10-113, KApplyDelta :223-235)
void LaunchDeltaApply(float4* pos_pred, float4* delta, const float* lambda, const uint32_t* indicesSorted,
                      const uint32_t* keysSorted, const uint32_t* cellStart, const uint32_t* cellEnd,
                      sim::DeviceParams dp, uint32_t N, hipStream_t s);
// reference: sim/pbf_delta_apply.cu:269-280 (KDeltaComputeCompact :116-220)
void LaunchDeltaApplyCompact(float4* pos_pred, float4* delta, const float* lambda, const uint32_t* indicesSorted,
                             const uint32_t* keysSorted, const uint32_t* uniqueKeys, const uint32_t* offsets,
                             const uint32_t* compactCount, sim::DeviceParams dp, uint32_t N, hipStream_t s);
// reference: sim/xsph.cu:189-205 (KXSPH :9-96)
void LaunchXSPH(float4* vel_out, const float4* vel_in, const float4* pos_pred, const uint32_t* indicesSorted,
                const uint32_t* keysSorted, const uint32_t* cellStart, const uint32_t* cellEnd, sim::DeviceParams dp,
                uint32_t N, hipStream_t s);
// reference: sim/xsph.cu:207-228 (KXSPHCompact :98-185)
void LaunchXSPHCompact(float4* vel_out, const float4* vel_in, const float4* pos_pred, const uint32_t* indicesSorted,
                       const uint32_t* keysSorted, const uint32_t* uniqueKeys, const uint32_t* offsets,
                       const uint32_t* compactCount, sim::DeviceParams dp, uint32_t N, hipStream_t s);

// ---- staging for the neighbour launchers (no counterpart in the reference) -----
// EnsureDropInStaging sizes two cell-sorted staging buffers of N particles for
// launches on `stream` (and the cell table and slot pairs the staged XSPH builds,
// about 56 bytes a particle in all).  With them, LaunchLambda / LaunchDeltaApply /
// LaunchXSPH and their compact forms on that stream first gather their inputs into slot
// order (one extra short kernel) and walk the gathered copy, instead of reading every
// candidate through indicesSorted; XSPH also pairs each cell's slots two to a lane
// (six more short kernels): the same pairs in the same order, the same bits, faster.
// Launches on other streams, or with more particles than reserved, read through
// indicesSorted as before.  Like EnsureCellCompactScratch it allocates, so call it
// before graph capture; growing the reservation frees the old buffers, so a graph
// captured before then must be captured again.  Launches on the one stream are
// ordered, and buffers are never shared between streams.  It also gives the stream its
// own compaction scratch for LaunchCellRangesCompact (the reference's one scratch per
// process, cell_compact.cu:50-56, is shared by every stream, so two streams compacting
// at once would overwrite each other's tile counts).  ReleaseDropInStaging frees a
// stream's buffers (call it once the stream is idle, before destroying it).
bool EnsureDropInStaging(uint32_t N, hipStream_t stream);
void ReleaseDropInStaging(hipStream_t stream);

// ---- statistics (diagnostics: read back to the host and synchronise `s`) -------
// reference: sim/stats.h:17-32 (StatsKernel, stats.cu:25-98).  Means over the
// sampled particles (i % sampleStride == 0) of in-range neighbours (r2 < h2),
// |v|, sum W and m * sum W.  Cross-particle sums are deterministic (double).
bool LaunchComputeStats(const float4* pos_pred, const float4* vel, const uint32_t* indices_sorted,
                        const uint32_t* cellStart, const uint32_t* cellEnd, sim::GridBounds grid,
                        sim::KernelCoeffs kc, float particleMass, uint32_t N, uint32_t numCells,
                        uint32_t sampleStride, double* outAvgNeighbors, double* outAvgSpeed, double* outAvgRhoRel,
                        double* outAvgRho, hipStream_t s);
// reference: sim/stats.h:34-45 (StatsBruteKernel, stats_bruteforce.cu:8-62), O(N^2).
bool LaunchComputeStatsBruteforce(const float4* pos_pred, const float4* vel, sim::KernelCoeffs kc,
                                  float particleMass, uint32_t N, uint32_t sampleStride, uint32_t maxISamples,
                                  double* outAvgNeighbors, double* outAvgSpeed, double* outAvgRhoRel,
                                  double* outAvgRho, hipStream_t s);

// ---- extensions (no reference counterpart) ----------------------------------
// Last recorded error (hipError_t value, or PBFX_E_* below), then cleared.
#define PBFX_E_INVALID_ARGUMENT 1001
#define PBFX_E_ALIAS 1002
#define PBFX_E_CAPACITY 1003
int pbfGetLastError(void);
const char* pbfGetLastErrorString(void);

// Stable LSD radix sort of (key, value) pairs over key bits [beginBit, endBit).
// Same temp-size contract as LaunchSortPairs (query with d_temp_storage == NULL).
// If d_vals_in is NULL the values are taken as the identity 0..N-1.
void pbfSortPairsBits(void* d_temp_storage, size_t* tempBytes, const uint32_t* d_keys_in, uint32_t* d_keys_out,
                      const uint32_t* d_vals_in, uint32_t* d_vals_out, uint32_t N, int beginBit, int endBit,
                      hipStream_t s);

#ifdef __cplusplus
}
#endif
