// k_integrate.hip — prediction, AABB boundary and velocity update.
//
// Streaming kernels: one float4 AoS element per lane, 16-byte loads/stores,
// 256-lane blocks.  Each is HBM-bound (48, 64 and 48 bytes per particle).
#include <cstdio>
#include <cstring>

#include "pbf_common.hpp"
#include "pbf_hip.h"

namespace pbfx {

namespace {
thread_local int t_last_error = 0;
thread_local char t_last_where[128] = {0};
}  // namespace

void set_error(int code, const char* where) {
    t_last_error = code;
    std::snprintf(t_last_where, sizeof(t_last_where), "%s", where ? where : "");
}

void check_launch(const char* where) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) set_error((int)e, where);
}

namespace {

constexpr int kBS = 256;

// P = X + (V + g dt) dt, w copied from X (reference: integrate.cu:29-52).
__global__ __launch_bounds__(kBS) void k_integrate_pred(const float4* __restrict__ X, const float4* __restrict__ V,
                                                        float4* __restrict__ P, float3 g, float dt, uint32_t N) {
    const uint32_t i = blockIdx.x * kBS + threadIdx.x;
    if (i >= N) return;
    const float4 p = X[i];
    const float4 v = V[i];
    const float vx = v.x + g.x * dt;
    const float vy = v.y + g.y * dt;
    const float vz = v.z + g.z * dt;
    P[i] = make_float4(p.x + vx * dt, p.y + vy * dt, p.z + vz * dt, p.w);
}

// AABB clamp with restitution, w reset (reference: boundary.cu:11-25).
__global__ __launch_bounds__(kBS) void k_boundary(float4* __restrict__ P, float4* __restrict__ V, sim::GridBounds g,
                                                  float restitution, uint32_t N) {
    const uint32_t i = blockIdx.x * kBS + threadIdx.x;
    if (i >= N) return;
    float4 p = P[i];
    float4 v = V[i];
    const float e = clamp_restitution(restitution);
    clamp_box(p.x, p.y, p.z, v.x, v.y, v.z, g, e);
    P[i] = make_float4(p.x, p.y, p.z, 1.0f);
    V[i] = make_float4(v.x, v.y, v.z, 0.0f);
}

// V = (P - X) * inv_dt, w = 0 (reference: boundary.cu:28-34).
__global__ __launch_bounds__(kBS) void k_velocity(float4* __restrict__ V, const float4* __restrict__ X,
                                                  const float4* __restrict__ P, float inv_dt, uint32_t N) {
    const uint32_t i = blockIdx.x * kBS + threadIdx.x;
    if (i >= N) return;
    const float4 x = X[i];
    const float4 p = P[i];
    V[i] = make_float4((p.x - x.x) * inv_dt, (p.y - x.y) * inv_dt, (p.z - x.z) * inv_dt, 0.0f);
}

}  // namespace
}  // namespace pbfx

using namespace pbfx;

extern "C" int pbfGetLastError(void) {
    const int e = t_last_error;
    t_last_error = 0;
    return e;
}

extern "C" const char* pbfGetLastErrorString(void) { return t_last_where; }

extern "C" void LaunchIntegratePred(float4* posCurr, const float4* velSrc, float4* posPred, float3 gravity, float dt,
                                    uint32_t N, hipStream_t s) {
    if (!posCurr || !velSrc || !posPred || N == 0) return;
    k_integrate_pred<<<ceil_div(N, kBS), kBS, 0, s>>>(posCurr, velSrc, posPred, gravity, dt, N);
    check_launch("LaunchIntegratePred");
}

extern "C" void LaunchBoundary(float4* pos_pred, float4* vel, sim::GridBounds grid, float restitution, uint32_t N,
                               hipStream_t s) {
    if (N == 0) return;
    if (!pos_pred || !vel) { set_error(PBFX_E_INVALID_ARGUMENT, "LaunchBoundary: null pointer"); return; }
    k_boundary<<<ceil_div(N, kBS), kBS, 0, s>>>(pos_pred, vel, grid, restitution, N);
    check_launch("LaunchBoundary");
}

extern "C" void LaunchVelocity(float4* vel, const float4* pos, const float4* pos_pred, float inv_dt, uint32_t N,
                               hipStream_t s) {
    if (N == 0) return;
    if (!vel || !pos || !pos_pred) { set_error(PBFX_E_INVALID_ARGUMENT, "LaunchVelocity: null pointer"); return; }
    k_velocity<<<ceil_div(N, kBS), kBS, 0, s>>>(vel, pos, pos_pred, inv_dt, N);
    check_launch("LaunchVelocity");
}
