// adjoint.hpp — reverse-mode kernels of the reordered PBF step.
//
// A recorded step (StepTape) freezes what the forward pass decided: the sort
// permutation, the occupied-cell index and neighbour-row spans, the clamp bits of
// the predict and post-solve box clamps, the sorted positions of every solver
// iteration (lambda in w) and the sorted velocities before XSPH.  The reverse pass
// differentiates the step on that frozen structure: the neighbour sets, the
// neighbour cap, the n < 2 gate, the lambda and displacement clamps and the box
// clamps keep the branch the forward pass took.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbf_params.h"
#include "sorted_path.hpp"

namespace pbfx {

constexpr int kMaxTapeIters = 8;

struct StepTape {
    float4* X = nullptr;            // positions at the start of the step (particle order)
    float4* V = nullptr;            // velocities at the start of the step (particle order)
    uint32_t* idxS = nullptr;       // sorted slot -> particle
    uint32_t* segOf = nullptr;      // sorted slot -> occupied cell
    uint2* spans = nullptr;         // neighbour-row spans of the step (rows x stride)
    float4* P[kMaxTapeIters + 1] = {};  // P[it]: sorted positions of iteration it (w = lambda); P[I]: final
    float4* Vs = nullptr;           // sorted velocities after the velocity phase
    uint32_t* predClamp = nullptr;  // particle order: bit a set when the predict clamp moved axis a
    uint32_t* finalClamp = nullptr; // sorted: bit a set when the post-solve clamp moved axis a
    // written by the forward kernels while recording (sorted order, per iteration):
    float4* lamA[kMaxTapeIters] = {};   // (sum of spiky gradients, b0) of the lambda pass
    uint2* lamB[kMaxTapeIters] = {};    // (bits of e0, slot that reached the neighbour cap or ~0)
    float4* dispU[kMaxTapeIters] = {};  // scaled correction before the displacement clamp
    float* xsphGm = nullptr;            // XSPH gate * m / rho0
    uint32_t* sceneOfSlot = nullptr;    // ensembles: scene of each sorted slot
    // the parameters the step was recorded with: its reverse uses them, whatever the
    // simulator's parameters are by the time the reverse pass runs
    sim::SimParams params{};
    bool recorded = false;
};

// Scratch of the reverse pass (sorted order unless noted).
// Words between the per-scene maxima of AdjointScratch::gmax: one 128-byte line per
// scene, so the atomics of different scenes do not queue on one cache line.
constexpr uint32_t kGmaxStride = 32;

struct AdjointScratch {
    float4* gPf = nullptr;    // dL/d(final position)
    float4* gVo = nullptr;    // dL/d(velocity after XSPH); w: the recorded XSPH factor
    float4* gVn = nullptr;    // dL/d(velocity before XSPH)
    float4* gP = nullptr;     // dL/d(positions of the current iteration)
    float4* a = nullptr;      // delta adjoint vectors
    float4* lamQ = nullptr;   // lambda adjoint: (sum of spiky gradients, beta = gLam b0)
    uint2* lamE = nullptr;    // lambda adjoint: (bits of eps = gLam e0, cap cut-off slot)
    double* partial = nullptr;  // per-block partial sums of dL/dc (one scene)
    double* gcSlot = nullptr;   // ensembles: dL/dc contribution of each sorted slot
    uint32_t* gmax = nullptr;   // renormalised reverse: bits of max |dL/dX_t|, |dL/dV_t| (scene q at q * kGmaxStride)
    uint32_t blocks = 0;
};

// Reverse of one recorded step.  gXin/gVin: dL/dX_{t+1}, dL/dV_{t+1} (particle
// order).  gXout/gVout: dL/dX_t, dL/dV_t (particle order; may not alias the
// inputs).  Adds dL/dc of the step into gcAccum (device doubles: one, or one per
// scene of an ensemble, whose scene s owns sorted slots [particleBase_s,
// particleBase_s + n_s) because its keys follow the previous scene's).
//
// Renormalised form (log2Scale != nullptr: device ints, one per scene of an ensemble,
// else one): the true gradients of scene q are its buffer entries times 2^log2Scale[q].
// The step's dL/dc is added unscaled (times 2^E_in, in double); then each scene's
// gXout and gVout entries are rescaled by a power of two so their largest magnitude
// lies in [0.5, 1) (exact: a power-of-two product of a normal float) and the scene's
// exponent grows by the power taken out.  Every kernel of the step is linear in
// the incoming gradients, so the rescaled buffers hold the same digits; the
// dynamics' growth of the gradient (about 8x per step on the dam break) then never
// overflows float32 over a long rollout.
void adjoint_step(const StepTape& t, const SpanTable& st, const sim::SimParams& p, const Scenes& sc,
                  const float4* gXin, const float4* gVin, float4* gXout, float4* gVout, double* gcAccum,
                  AdjointScratch& w, hipStream_t s, int32_t* log2Scale = nullptr);

// Clamp bits (bit a: axis a was moved by the box clamp) are written by the
// forward kernels when recording: launch_predict_hash(..., clampBits) in particle
// order and launch_delta_sorted(..., clampBits) on the last iteration, sorted.

}  // namespace pbfx
