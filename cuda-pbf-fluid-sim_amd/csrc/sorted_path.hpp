// sorted_path.hpp — launchers of the cell-reordered step used by the headless
// pipeline.  State lives in original particle order between steps (the
// reference's layout); inside a step the solver works on copies gathered into
// cell-sorted order, where every neighbour row is a contiguous slot span.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbf_params.h"

namespace pbfx {

enum class CellLookup : int { Dense = 0, Compact = 1, Hashed = 2 };

// One slot of the open-addressing cell table: cell key -> [beg, end) of its
// sorted slots.  (key, epoch) form the 64-bit word claimed by CAS; a slot whose
// epoch differs from the current step's is empty, so the table is never cleared.
struct alignas(16) HashSlot {
    uint32_t key;
    uint32_t epoch;
    uint32_t beg;
    uint32_t end;
};

struct HashTable {
    HashSlot* slots = nullptr;
    uint32_t mask = 0;          // table size - 1 (power of two)
    uint32_t scatter = 0;       // 1 when the key space exceeds the table: homes are hashed
    uint32_t* epoch = nullptr;  // device counter, advanced once per step
    uint32_t* overflow = nullptr;  // device flag, set when an insert finds no slot
};

// Neighbour-row spans: spans[r * stride + m] = [lo, hi) slots of row r of
// occupied cell m, rows ordered dz-major then dy, (2*reach+1)^2 rows.
struct SpanTable {
    uint2* spans = nullptr;
    uint32_t stride = 0;
    int reach = 1;
    int rows = 9;
    // a wave streams a row through LDS when the union of its lanes' spans is at most
    // streamRatio times the longest one; otherwise each lane walks its own span
    uint32_t streamRatio = 3;
    // streamed rows: 1 = each lane tests a whole 64-candidate chunk into a bit mask,
    // then queues its set bits; 0 = each candidate is tested and queued in turn
    uint32_t chunkBits = 1;
    // lanes per block of the neighbour kernels (128 or 256)
    uint32_t block = 256;
};

// Sorted position arrays hold this many slots past the last particle, so that the
// neighbour walks may load a few candidates past the end of a span without a bounds
// select (the loaded values are never used).
constexpr uint32_t kSlotPad = 4;

// Candidate mask words per particle: the mask covers the first 32 * kCandMaskWords
// candidates of a particle's walk; later candidates are tested directly.
constexpr uint32_t kCandMaskWords = 64;

// Per-step candidate mask (built and walked in walk.hpp): words[w * stride + s] holds
// bits 32w..32w+31 of slot s's walk; rs2 is the squared skin radius.
// The pair-arithmetic check's record (launch_ratio_check): zh = RN(1 / w_q) and
// zl = RN(1 / w_q - zh), ok = 1 when fma(W, zh, W * zl) gives the division's s_corr term
// for every float W from +0 to the largest a pair can have; rsq_ok = 1 when
// spiky_coef_rsq gives spiky_coef's bits for every float r2 from +0 to h2; and the
// parameters the check was made for (the delta kernels take the two shortcuts, together,
// only when both checks passed and theirs match).
struct RatioCoef {
    float zh, zl, ok, w_q, poly6, h2, neg_k, smin;
    float rsq_ok, h, eps, neg3_spiky;
};

struct CandMask {
    uint32_t* words = nullptr;
    uint32_t stride = 0;
    float rs2 = 0.f;
};

// Squared skin radius of the candidate mask for p's solver iterations, or 0 when no
// mask applies (no displacement clamp, or a skin wider than 2h).
float mask_skin_radius2(const sim::SimParams& p);

// Per-scene parameters of a batched ensemble.  All scenes share h, rest density,
// mass, dt, solver settings and the PBF tuning; these fields may differ.
struct SceneParams {
    sim::GridBounds grid;   // box and cell grid of the scene
    float3 gravity;
    float xsph_c;
    float restitution;
    uint32_t cellBase;      // first key of the scene in the ensemble's key space
    uint32_t particleBase;  // first particle (and first sorted slot) of the scene
    uint32_t numParticles;
};
static_assert(sizeof(SceneParams) == 72, "SceneParams layout");

// Device view of an ensemble.  With params == nullptr the kernels describe one
// scene through their own parameters.
struct Scenes {
    const SceneParams* params = nullptr;
    const uint32_t* sceneOf = nullptr;   // scene of each particle, original order
    uint32_t* sceneOfSlot = nullptr;     // scene of each sorted slot (written by spans_gather)
    const uint32_t* cellBase = nullptr;  // count + 1 ascending key bases
    uint32_t count = 0;
};

// Bits of the step's flags word (reordered path), which the predict kernel clears:
//  - kStepVelocityUnbounded: a velocity the step wrote is NaN or has a component of
//    magnitude 2^126 or more (XSPH then takes its general form);
//  - kStepNanPosition: a solver pass before the last wrote a NaN position.  The
//    candidate mask's skin argument does not cover a NaN position (the reference's
//    range test admits it at any distance), so the later lambda passes then walk
//    every candidate instead of the mask.
constexpr uint32_t kStepVelocityUnbounded = 1u;
constexpr uint32_t kStepNanPosition = 2u;

// P = X + (V + g dt) dt, AABB clamp, write Xn = (P, 1) and its cell key.
// Thread 0 also advances the hash-table epoch when `epoch` is not null, and clears the
// step's flags when `stepFlags` is not null.
// clampBits (optional, particle order): axes moved by the clamp, for the adjoint.
void launch_predict_hash(const float4* X, const float4* V, float4* Xn, uint32_t* keys, uint32_t* epoch,
                         const sim::SimParams& p, const Scenes& sc, hipStream_t s, uint32_t* clampBits = nullptr,
                         uint32_t* stepFlags = nullptr);

// Insert the M occupied cells (count on device) into the table.
void launch_hash_insert(const HashTable& t, const uint32_t* uniq, const uint32_t* off, const uint32_t* count,
                        uint32_t maxCells, hipStream_t s);

// Gather Ps[s] = Xn[idxS[s]] and build every occupied cell's row spans.
void launch_spans_gather(CellLookup lookup, const float4* Xn, const uint32_t* idxS, float4* Ps, uint32_t N,
                         const uint32_t* uniq, const uint32_t* off, const uint32_t* count,
                         const uint32_t* cellStart, const uint32_t* cellEnd, const HashTable& table,
                         const SpanTable& st, const sim::GridBounds& grid, const Scenes& sc, hipStream_t s);

// lambda of every sorted slot, stored in Ps[s].w (full walk, or the candidate mask's walk).
// Recording (non-null recA/recB): recA[s] = (sum of spiky gradients, b0) and recB[s] =
// (bits of e0, slot of the neighbour that reached the cap or ~0), the reverse pass's
// coefficients (lambda_adjoint_coefs).
// Build the candidate mask from the positions Ps (one lane per slot pair of
// build_slot_pairs, every candidate of the cell's rows): the skin bit of each candidate
// in walk order, mk.words[w * stride + s].
void launch_mask_build(const float4* Ps, const uint32_t* segOf, const SpanTable& st, const CandMask& mk, uint32_t N,
                       const uint2* pairs, const uint32_t* pairCount, hipStream_t s);

// stepFlags: with kStepNanPosition set, the full walk replaces the mask's.
void launch_lambda_sorted(float4* Ps, const uint32_t* segOf, const SpanTable& st, const CandMask& mk,
                          bool useMask, const sim::DeviceParams& dp, uint32_t N, hipStream_t s,
                          float4* recA = nullptr, uint2* recB = nullptr, const uint32_t* perm = nullptr,
                          const uint32_t* stepFlags = nullptr);

// Check the two-instruction s_corr ratio for the parameters dp on every W it can meet and
// write the record `coef` (bad: one word of device scratch).  Runs on stream s; the
// simulator calls it whenever its parameters change.
void launch_ratio_check(const sim::DeviceParams& dp, RatioCoef* coef, uint32_t* bad, hipStream_t s);

// Pb[s] = Pa[s] + delta(s).  On the last iteration also clamps to the box and
// writes Vs[s] = (Pb[s] - X[idxS[s]]) * invDt.  Every candidate is walked.  Recording:
// clampBits[s] (last iteration)
// and recU[s] = the scaled correction before the displacement clamp.  stepFlags: the
// last iteration sets kStepVelocityUnbounded, the others kStepNanPosition (see there).
void launch_delta_sorted(const float4* Pa, float4* Pb, const uint32_t* segOf, const SpanTable& st,
                         const sim::DeviceParams& dp, uint32_t N, bool last,
                         const float4* X, const uint32_t* idxS, float4* Vs, float invDt, const Scenes& sc,
                         hipStream_t s, uint32_t* clampBits = nullptr, float4* recU = nullptr,
                         const uint32_t* perm = nullptr, uint32_t* stepFlags = nullptr,
                         const RatioCoef* ratioCoef = nullptr);

// perm[t] = the slot lane t of the lambda and delta passes takes: each 128-slot group
// dealt to its two waves by the particles' place inside their cells (positions Ps).
void launch_lane_order(const float4* Ps, uint32_t N, const sim::GridBounds& grid, const Scenes& sc, uint32_t* perm,
                       hipStream_t s);

// No solver iterations: clamp Ps and compute Vs (the post-solve boundary and velocity phases),
// setting kStepVelocityUnbounded as launch_delta_sorted's last iteration does.
void launch_finalize_sorted(float4* Ps, const float4* X, const uint32_t* idxS, float4* Vs,
                            const sim::GridBounds& grid, float invDt, uint32_t N, const Scenes& sc, hipStream_t s,
                            uint32_t* stepFlags = nullptr);

// XSPH (Jacobi over Vs, skipped when xsph_c <= 0), boundary with restitution,
// scatter Xn[idxS[s]] = (P, 1) and V[idxS[s]] = (v, 0), one slot pair a lane
// (pairs / pairCount from build_slot_pairs).  Recording (non-null recGm):
// recGm[s] = gate * m / rho0, walked also when xsph_c is 0.  stepFlags: the step's flags
// (kStepVelocityUnbounded clear: the weighted-select form of the walk); null: the general form.
void launch_xsph_scatter(const float4* Ps, const float4* Vs, const uint32_t* segOf, const SpanTable& st,
                         const sim::DeviceParams& dp, uint32_t N, const uint2* pairs, const uint32_t* pairCount,
                         const uint32_t* idxS, float4* Xn, float4* V, const Scenes& sc, hipStream_t s,
                         float* recGm = nullptr, const uint32_t* stepFlags = nullptr);

}  // namespace pbfx
