// pbf_common.hpp — device helpers and error state shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <climits>

#include "pbf_params.h"

namespace pbfx {

// ---------------------------------------------------------------- errors
// Launchers return void (the reference's ABI); failures are recorded here and
// read back with pbfGetLastError().
void set_error(int code, const char* where);
void check_launch(const char* where);

#define PBFX_CHECK(expr, where)                                        \
    do {                                                               \
        hipError_t e__ = (expr);                                       \
        if (e__ != hipSuccess) ::pbfx::set_error((int)e__, where);     \
    } while (0)

// ---------------------------------------------------------- device helpers
// Saturating float->int32, NaN -> 0 (hardware conversion rule, made explicit
// so the compiler cannot treat out-of-range values as undefined).
__device__ __forceinline__ int sat_f2i(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return INT_MAX;
    if (f < -2147483648.0f) return INT_MIN;
    return (int)f;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Cell key of a position: floor((p - min) / cs), clamped per axis, row-major
// (reference: hash_keys.cu:6-26).
__device__ __forceinline__ uint32_t cell_key(float px, float py, float pz, const sim::GridBounds& g) {
    const float rx = (px - g.mins.x) / g.cellSize;
    const float ry = (py - g.mins.y) / g.cellSize;
    const float rz = (pz - g.mins.z) / g.cellSize;
    const int cx = clampi(sat_f2i(floorf(rx)), 0, g.dim.x - 1);
    const int cy = clampi(sat_f2i(floorf(ry)), 0, g.dim.y - 1);
    const int cz = clampi(sat_f2i(floorf(rz)), 0, g.dim.z - 1);
    return (uint32_t)((cz * g.dim.y + cy) * g.dim.x + cx);
}

// Key -> (x, y, z) cell (reference: pbf_lambda.cu:31-36).
__device__ __forceinline__ void decode_key(uint32_t key, int3 dim, int& x, int& y, int& z) {
    x = (int)(key % (uint32_t)dim.x);
    const uint32_t kdx = key / (uint32_t)dim.x;
    y = (int)(kdx % (uint32_t)dim.y);
    z = (int)(kdx / (uint32_t)dim.y);
}

__device__ __forceinline__ int reach_of(float h, float cs) {
    return (cs > 0.f) ? max(1, sat_f2i(ceilf(h / cs))) : 1;
}

// AABB clamp of one particle (reference: boundary.cu:11-25).  Returns the
// clamped position (w = 1) and updates v.xyz; w of v becomes 0 at the caller.
__device__ __forceinline__ void clamp_box(float& px, float& py, float& pz, float& vx, float& vy, float& vz,
                                          const sim::GridBounds& g, float e) {
    if (px < g.mins.x) { px = g.mins.x; vx *= -e; }
    if (py < g.mins.y) { py = g.mins.y; vy *= -e; }
    if (pz < g.mins.z) { pz = g.mins.z; vz *= -e; }
    if (px > g.maxs.x) { px = g.maxs.x; vx *= -e; }
    if (py > g.maxs.y) { py = g.maxs.y; vy *= -e; }
    if (pz > g.maxs.z) { pz = g.maxs.z; vz *= -e; }
}

// Axes the box clamp will move: bit a set when p_a < min_a or p_a > max_a.
__device__ __forceinline__ uint32_t clamp_bits(float px, float py, float pz, const sim::GridBounds& g) {
    return ((px < g.mins.x || px > g.maxs.x) ? 1u : 0u) | ((py < g.mins.y || py > g.maxs.y) ? 2u : 0u) |
           ((pz < g.mins.z || pz > g.maxs.z) ? 4u : 0u);
}

__device__ __forceinline__ float clamp_restitution(float r) { return fminf(fmaxf(r, 0.0f), 1.0f); }

// Blocks of a 1-D launch are dealt round-robin to the 8 XCDs.  This remaps the
// hardware block index so that each XCD receives one contiguous run of logical
// tiles; neighbouring tiles (neighbouring cells) then share an L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t bid, uint32_t nblocks) {
    constexpr uint32_t kXcd = 8;
    const uint32_t per = nblocks / kXcd;
    const uint32_t rem = nblocks % kXcd;
    const uint32_t xcd = bid % kXcd;
    const uint32_t slot = bid / kXcd;
    // XCDs [0, rem) own per+1 tiles, the rest own per.
    const uint32_t start = xcd * per + (xcd < rem ? xcd : rem);
    return start + slot;
}

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace pbfx
