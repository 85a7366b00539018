// scenes.cpp — deterministic scene generators (host side): the auto-sized box
// lattice, the CubeMix group centres and the CubeMix cube blocks.
#include "scenes.hpp"

#include <algorithm>
#include <cmath>

namespace pbfx {

float jitter_uniform(uint64_t seed, uint64_t particle, uint32_t axis) {
    // SplitMix64 finaliser (Steele, Lea, Flood 2014) over (seed, 3*particle+axis+1).
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (particle * 3ull + (uint64_t)axis + 1ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (float)(uint32_t)(z >> 40) * (1.0f / 16777216.0f);
}

bool lattice_auto_dims(uint32_t total, float3 mins, float3 maxs, float spacing, float3& origin, uint32_t dims[3]) {
    // reference: simulator.cpp:882-941 (seedBoxLatticeAuto)
    if (!(spacing > 0.f)) return false;
    const float sx = std::fmax(0.f, maxs.x - mins.x);
    const float sy = std::fmax(0.f, maxs.y - mins.y);
    const float sz = std::fmax(0.f, maxs.z - mins.z);
    auto fit = [spacing](float len) { return std::max<uint32_t>(1u, (uint32_t)std::floor(len / spacing)); };
    const uint32_t maxX = fit(sx), maxY = fit(sy), maxZ = fit(sz);
    const uint64_t T = std::max<uint32_t>(1u, total);
    uint32_t nx = std::min<uint32_t>(maxX, (uint32_t)std::ceil(std::pow((double)T, 1.0 / 3.0)));
    const uint64_t remXY = (T + nx - 1) / nx;
    uint32_t ny = std::min<uint32_t>(maxY, (uint32_t)std::ceil(std::sqrt((double)remXY)));
    uint32_t nz = std::min<uint32_t>(maxZ, (uint32_t)std::ceil((double)remXY / std::max<uint32_t>(1u, ny)));
    while ((uint64_t)nx * ny * nz < T) {
        if (nx < maxX) ++nx;
        else if (ny < maxY) ++ny;
        else if (nz < maxZ) ++nz;
        else break;
    }
    const float m = spacing * 0.25f;
    origin.x = std::fmin(std::fmax(origin.x, mins.x + m), maxs.x - m);
    origin.y = std::fmin(std::fmax(origin.y, mins.y + m), maxs.y - m);
    origin.z = std::fmin(std::fmax(origin.z, mins.z + m), maxs.z - m);
    dims[0] = nx;
    dims[1] = ny;
    dims[2] = nz;
    return true;
}

bool cube_mix_centers(const CubeMixLayout& c, float3* out) {
    // reference: console.cpp:277-316 (GenerateCubeMixCenters).  The per-layer counts
    // keep the reference's unsigned arithmetic, including its wrap when there are
    // more layers than groups (such layers are never reached).
    if (c.layers == 0) return c.groupCount == 0;
    std::vector<uint32_t> perLayer(c.layers, 0u);
    uint32_t remaining = c.groupCount;
    for (uint32_t L = 0; L < c.layers; ++L) {
        uint32_t take = remaining / (c.layers - L);
        if (take == 0) take = 1;
        perLayer[L] = take;
        remaining -= take;
    }
    uint32_t g = 0;
    for (uint32_t layer = 0; layer < c.layers && g < c.groupCount; ++layer) {
        const uint32_t inLayer = perLayer[layer];
        const uint32_t nx = (uint32_t)std::ceil(std::sqrt((double)inLayer));
        float y = c.baseHeight + (float)layer * c.layerSpacing;
        if (c.layerJitter) {
            // one variate per layer from the seeder's generator (the reference draws it
            // from std::mt19937 seeded with seed ^ layer)
            y += (2.0f * jitter_uniform(c.layerJitterSeed ^ layer, 0, 1) - 1.0f) * c.layerJitterAmp;
        }
        for (uint32_t k = 0; k < inLayer && g < c.groupCount; ++k) {
            const uint32_t ix = k % nx, iz = k / nx;
            out[g++] = make_float3((float)ix * c.groupSpacing, y, (float)iz * c.groupSpacing);
        }
    }
    return true;
}

void cube_mix_fill(const sim::GridBounds& b, uint32_t groups, const float3* centers, uint32_t edge, float spacing,
                   float amp, uint64_t seed, float4* out) {
    // reference: simulator.cpp:920-975 (seedCubeMix): per group an edge^3 block centred
    // on the group centre, x fastest, clamped to the box, w = group index; then jitter
    // and a second clamp.
    const float start = -0.5f * spacing * (float)(edge - 1);
    uint64_t i = 0;
    for (uint32_t g = 0; g < groups; ++g) {
        const float3 c = centers[g];
        for (uint32_t iz = 0; iz < edge; ++iz)
            for (uint32_t iy = 0; iy < edge; ++iy)
                for (uint32_t ix = 0; ix < edge; ++ix) {
                    float x = c.x + start + (float)ix * spacing;
                    float y = c.y + start + (float)iy * spacing;
                    float z = c.z + start + (float)iz * spacing;
                    x = std::fmin(std::fmax(x, b.mins.x), b.maxs.x);
                    y = std::fmin(std::fmax(y, b.mins.y), b.maxs.y);
                    z = std::fmin(std::fmax(z, b.mins.z), b.maxs.z);
                    out[i++] = make_float4(x, y, z, (float)g);
                }
    }
    if (amp > 0.f) {
        for (uint64_t k = 0; k < i; ++k) {
            const float jx = (2.0f * jitter_uniform(seed, k, 0) - 1.0f) * amp;
            const float jy = (2.0f * jitter_uniform(seed, k, 1) - 1.0f) * amp;
            const float jz = (2.0f * jitter_uniform(seed, k, 2) - 1.0f) * amp;
            out[k].x = std::fmin(std::fmax(out[k].x + jx, b.mins.x), b.maxs.x);
            out[k].y = std::fmin(std::fmax(out[k].y + jy, b.mins.y), b.maxs.y);
            out[k].z = std::fmin(std::fmax(out[k].z + jz, b.mins.z), b.maxs.z);
        }
    }
}

}  // namespace pbfx

// ------------------------------------------------------------------ C ABI
extern "C" int pbfLatticeAutoDims(uint32_t total, const float* boxMin, const float* boxMax, float spacing,
                                  float* originInOut, uint32_t* dimsOut) {
    if (!boxMin || !boxMax || !originInOut || !dimsOut) return -1;
    float3 o = make_float3(originInOut[0], originInOut[1], originInOut[2]);
    if (!pbfx::lattice_auto_dims(total, make_float3(boxMin[0], boxMin[1], boxMin[2]),
                                 make_float3(boxMax[0], boxMax[1], boxMax[2]), spacing, o, dimsOut))
        return -1;
    originInOut[0] = o.x;
    originInOut[1] = o.y;
    originInOut[2] = o.z;
    return 0;
}

extern "C" int pbfCubeMixCenters(const PbfCubeMixLayout* layout, float* centersOut) {
    if (!layout || (!centersOut && layout->groupCount)) return -1;
    pbfx::CubeMixLayout c;
    c.groupCount = layout->groupCount;
    c.layers = layout->layers;
    c.groupSpacing = layout->groupSpacing;
    c.layerSpacing = layout->layerSpacing;
    c.baseHeight = layout->baseHeight;
    c.layerJitter = layout->layerJitter != 0;
    c.layerJitterAmp = layout->layerJitterAmp;
    c.layerJitterSeed = layout->layerJitterSeed;
    std::vector<float3> tmp(c.groupCount);
    if (!pbfx::cube_mix_centers(c, tmp.data())) return -1;
    for (uint32_t g = 0; g < c.groupCount; ++g) {
        centersOut[3 * g + 0] = tmp[g].x;
        centersOut[3 * g + 1] = tmp[g].y;
        centersOut[3 * g + 2] = tmp[g].z;
    }
    return 0;
}
