// pbf_params.h — parameter block of the PBF solver core (gfx950 / HIP).
//
// These are the POD types the phase API passes to every launcher.  Field order,
// defaults and sizes are the ABI of the sim/ phase pipeline
// (reference: sim/parameters.h:9-123), so a caller compiled against that
// pipeline can hand its structs to this library unchanged.  The vector members
// use HIP's own float3/int3 (12 bytes, 4-byte alignment), which lay out exactly
// like the reference's.
#pragma once

#include <stdint.h>
#include <hip/hip_runtime.h>

namespace sim {

// Smoothing-kernel constants (reference: sim/parameters.h:9-16).
struct KernelCoeffs {
    float h;
    float inv_h;
    float h2;
    float poly6;   // 315 / (64 pi h^9)
    float spiky;   // 15 / (pi h^6)
    float visc;    // 15 / (2 pi h^3)
};

// Simulation box and its uniform cell grid (reference: sim/parameters.h:18-24).
struct GridBounds {
    float3 mins;
    float3 maxs;
    float  cellSize;
    int3   dim;
};

// PBF tuning knobs (reference: sim/parameters.h:26-57).  Fields that no kernel
// reads (xpbd_enable, lambda_warm_start_*, semi_implicit_integration_enable)
// are kept so the layout stays identical.
struct PbfTuning {
    int   scorr_enable = 1;
    float scorr_k = 0.003f;
    float scorr_n = 4.0f;
    float scorr_dq_h = 0.3f;
    float wq_min = 1e-12f;
    float scorr_min = -0.25f;

    float grad_r_eps = 1e-6f;
    float lambda_denom_eps = 1e-4f;

    float compliance = 0.0f;
    int   xpbd_enable = 0;

    int   enable_lambda_clamp = 1;
    float lambda_max_abs = 50.0f;
    int   enable_disp_clamp = 1;
    float disp_clamp_max_h = 0.05f;

    int   enable_relax = 1;
    float relax_omega = 0.75f;

    int   xsph_gate_enable = 0;
    int   xsph_n_min = 0;
    int   xsph_n_max = 8;

    int   lambda_warm_start_enable = 0;
    float lambda_warm_start_decay = 0.5f;

    int   semi_implicit_integration_enable = 0;
};

// Host-side parameter aggregate (reference: sim/parameters.h:59-77).
struct SimParams {
    uint32_t numParticles;
    uint32_t maxParticles;
    float3   gravity;
    float    dt;
    float    cfl;
    float    restDensity;
    int      solverIters;
    int      maxNeighbors;
    int      sortEveryN;
    float    boundaryRestitution;
    float    particleMass;
    PbfTuning pbf{};
    float    xsph_c = 0.05f;
    KernelCoeffs kernel{};
    GridBounds  grid{};
    uint32_t ghostParticleCount = 0;
};

// By-value kernel parameter block (reference: sim/parameters.h:79-92).
struct DeviceParams {
    KernelCoeffs kernel;
    GridBounds   grid;
    float3       gravity;
    float        restDensity;
    float        dt;
    float        inv_dt;
    float        boundaryRestitution;
    int          maxNeighbors;
    float        particleMass;
    PbfTuning    pbf;
    float        xsph_c;
};

static_assert(sizeof(KernelCoeffs) == 24, "KernelCoeffs ABI");
static_assert(sizeof(GridBounds) == 40, "GridBounds ABI");
static_assert(sizeof(PbfTuning) == 88, "PbfTuning ABI");
static_assert(sizeof(DeviceParams) == 192, "DeviceParams ABI");
static_assert(offsetof(DeviceParams, pbf) == 100, "DeviceParams.pbf offset");
static_assert(offsetof(DeviceParams, xsph_c) == 188, "DeviceParams.xsph_c offset");
static_assert(sizeof(SimParams) == 212, "SimParams ABI");
static_assert(offsetof(SimParams, pbf) == 52, "SimParams.pbf offset");
static_assert(offsetof(SimParams, grid) == 168, "SimParams.grid offset");

// Kernel constants from h, single precision, same operation order as the
// reference (sim/parameters.h:94-107) so poly6/spiky are bit-identical.
inline KernelCoeffs MakeKernelCoeffs(float h) {
    const float pi = 3.14159265358979323846f;
    KernelCoeffs kc{};
    kc.h = h;
    kc.inv_h = 1.0f / h;
    kc.h2 = h * h;
    const float h3 = h * kc.h2;
    const float h6 = h3 * h3;
    const float h9 = h6 * h3;
    kc.poly6 = 315.0f / (64.0f * pi * h9);
    kc.spiky = 15.0f / (pi * h6);
    kc.visc = 15.0f / (2.0f * pi * h3);
    return kc;
}

// SimParams -> DeviceParams (reference: sim/parameters.h:109-123).
inline DeviceParams MakeDeviceParams(const SimParams& sp) {
    DeviceParams dp{};
    dp.kernel = sp.kernel;
    dp.grid = sp.grid;
    dp.gravity = sp.gravity;
    dp.restDensity = sp.restDensity;
    dp.dt = sp.dt;
    dp.inv_dt = (sp.dt > 0.f) ? (1.0f / sp.dt) : 0.0f;
    dp.boundaryRestitution = sp.boundaryRestitution;
    dp.maxNeighbors = sp.maxNeighbors;
    dp.particleMass = sp.particleMass;
    dp.pbf = sp.pbf;
    dp.xsph_c = sp.xsph_c;
    return dp;
}

}  // namespace sim
