// scenes.hpp — deterministic scene generators shared by the simulator and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "pbf_params.h"
#include "pbf_sim.h"

namespace pbfx {

// Uniform variate in [0, 1) with a 24-bit mantissa: SplitMix64 over
// (seed, 3 * particle + axis + 1).  The seeders' jitter source.
float jitter_uniform(uint64_t seed, uint64_t particle, uint32_t axis);

// Lattice dimensions for about `total` particles at `spacing` inside the box, and the
// origin pulled a quarter spacing inside it (reference: seedBoxLatticeAuto).
bool lattice_auto_dims(uint32_t total, float3 mins, float3 maxs, float spacing, float3& origin, uint32_t dims[3]);

struct CubeMixLayout {
    uint32_t groupCount = 8;
    uint32_t layers = 2;
    float groupSpacing = 200.f;
    float layerSpacing = 200.f;
    float baseHeight = 200.f;
    bool layerJitter = false;
    float layerJitterAmp = 0.f;
    uint32_t layerJitterSeed = 0x13579BDFu;
};

// Group centres, groups dealt to layers as evenly as the reference does, each layer a
// ceil(sqrt(n)) wide grid in x/z (reference: GenerateCubeMixCenters).
bool cube_mix_centers(const CubeMixLayout& c, float3* out);

// groups * edge^3 positions (w = group index), jittered by +-amp per axis and clamped.
void cube_mix_fill(const sim::GridBounds& b, uint32_t groups, const float3* centers, uint32_t edge, float spacing,
                   float amp, uint64_t seed, float4* out);

}  // namespace pbfx
