// consumer.cpp — a C++ caller compiled against the public headers only, as the
// reference's own host code would be after the switch: the parameter structs from
// pbf_params.h, the drop-in launchers from pbf_hip.h and the simulator from pbf_sim.h.
// Exit status 0 and a final "OK" line mean every check passed.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "pbf_hip.h"
#include "pbf_params.h"
#include "pbf_sim.h"

#define CHECK(c, msg)                                                    \
    do {                                                                 \
        if (!(c)) {                                                      \
            std::printf("FAIL: %s (pbf error %d)\n", msg, pbfGetLastError()); \
            return 1;                                                    \
        }                                                                \
    } while (0)

// One whole step through the drop-in launchers on caller-owned buffers, in the
// reference's phase order: PhaseIntegrate (phases.cpp:33-37), PhaseGridBuildFull
// (:40-63), PhaseSolveIterations (:70-95), PhaseVelocity (:100-105), then XsphOp
// (post_ops.cpp:54-71, over a velocity snapshot) and BoundaryOp (:25-29).  X and V
// advance in place; returns false on a launcher error.
static bool dropin_step(const sim::SimParams& p, bool compact, std::vector<float4>& X, std::vector<float4>& V,
                        hipStream_t st) {
    const uint32_t N = p.numParticles;
    const uint32_t cells = (uint32_t)(p.grid.dim.x * p.grid.dim.y * p.grid.dim.z);
    const sim::DeviceParams dp = sim::MakeDeviceParams(p);
    float4 *pos = nullptr, *pred = nullptr, *vel = nullptr, *velSnap = nullptr, *delta = nullptr;
    float* lambda = nullptr;
    uint32_t *keys = nullptr, *idx = nullptr, *keysS = nullptr, *idxS = nullptr, *a = nullptr, *b = nullptr,
             *count = nullptr;
    bool ok = hipMalloc(&pos, 16 * N) == hipSuccess && hipMalloc(&pred, 16 * N) == hipSuccess &&
              hipMalloc(&vel, 16 * N) == hipSuccess && hipMalloc(&velSnap, 16 * N) == hipSuccess &&
              hipMalloc(&delta, 16 * N) == hipSuccess && hipMalloc(&lambda, 4 * N) == hipSuccess &&
              hipMalloc(&keys, 4 * N) == hipSuccess && hipMalloc(&idx, 4 * N) == hipSuccess &&
              hipMalloc(&keysS, 4 * N) == hipSuccess && hipMalloc(&idxS, 4 * N) == hipSuccess &&
              hipMalloc(&count, 4) == hipSuccess;
    // dense: cellStart/cellEnd over every cell; compact: unique keys and N + 1 offsets
    ok = ok && hipMalloc(&a, 4 * (size_t)(compact ? N : cells)) == hipSuccess &&
         hipMalloc(&b, 4 * (size_t)(compact ? N + 1 : cells)) == hipSuccess;
    size_t tempBytes = 0;
    void* temp = nullptr;
    if (ok) LaunchSortPairsQuery(&tempBytes, keys, keysS, idx, idxS, N, st);
    ok = ok && hipMalloc(&temp, tempBytes ? tempBytes : 1) == hipSuccess;
    if (compact) ok = ok && EnsureCellCompactScratch(N, 256);
    ok = ok && hipMemcpy(pos, X.data(), 16 * N, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(vel, V.data(), 16 * N, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        LaunchIntegratePred(pos, vel, pred, p.gravity, p.dt, N, st);
        LaunchBoundary(pred, vel, p.grid, 0.0f, N, st);
        LaunchHashKeys(keys, idx, pred, p.grid, N, st);
        LaunchSortPairs(temp, tempBytes, keys, keysS, idx, idxS, N, st);
        if (compact)
            LaunchCellRangesCompact(a, b, count, keysS, N, st);
        else
            LaunchCellRanges(a, b, keysS, N, cells, st);
        for (int it = 0; it < p.solverIters; ++it) {
            if (compact) {
                LaunchLambdaCompact(lambda, pred, idxS, keysS, a, b, count, dp, N, st);
                LaunchDeltaApplyCompact(pred, delta, lambda, idxS, keysS, a, b, count, dp, N, st);
            } else {
                LaunchLambda(lambda, pred, idxS, keysS, a, b, dp, N, st);
                LaunchDeltaApply(pred, delta, lambda, idxS, keysS, a, b, dp, N, st);
            }
        }
        LaunchBoundary(pred, vel, p.grid, 0.0f, N, st);
        LaunchVelocity(vel, pos, pred, 1.0f / p.dt, N, st);
        if (p.xsph_c > 0.f) {
            ok = hipMemcpyAsync(velSnap, vel, 16 * N, hipMemcpyDeviceToDevice, st) == hipSuccess;
            if (compact)
                LaunchXSPHCompact(vel, velSnap, pred, idxS, keysS, a, b, count, dp, N, st);
            else
                LaunchXSPH(vel, velSnap, pred, idxS, keysS, a, b, dp, N, st);
        }
        LaunchBoundary(pred, vel, p.grid, p.boundaryRestitution, N, st);
        // the step's result is the predicted buffer (the ping-pong swap)
        ok = ok && hipStreamSynchronize(st) == hipSuccess && pbfGetLastError() == 0 &&
             hipMemcpy(X.data(), pred, 16 * N, hipMemcpyDeviceToHost) == hipSuccess &&
             hipMemcpy(V.data(), vel, 16 * N, hipMemcpyDeviceToHost) == hipSuccess;
    }
    hipFree(pos); hipFree(pred); hipFree(vel); hipFree(velSnap); hipFree(delta); hipFree(lambda);
    hipFree(keys); hipFree(idx); hipFree(keysS); hipFree(idxS); hipFree(a); hipFree(b); hipFree(count);
    hipFree(temp);
    return ok;
}

static bool same_bits(const std::vector<float4>& a, const std::vector<float4>& b) {
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(float4)) == 0;
}

int main() {
    // the 16^3 dam-break block in a 1 m box, parameters derived as the reference app does
    const float r = 0.005f, h = 2.5f * r;
    sim::SimParams p{};
    p.numParticles = p.maxParticles = 16 * 16 * 16;
    p.dt = 1.0f / 120.0f;
    p.gravity = make_float3(0.f, -9.81f, 0.f);
    p.restDensity = 1000.f;
    p.solverIters = 2;
    p.maxNeighbors = 64;
    p.sortEveryN = 1;
    p.xsph_c = 0.01f;
    p.kernel = sim::MakeKernelCoeffs(h);
    p.particleMass = 1000.f * h * h * h;
    p.grid.mins = make_float3(0.f, 0.f, 0.f);
    p.grid.maxs = make_float3(1.f, 1.f, 1.f);
    p.grid.cellSize = h;
    p.grid.dim = make_int3((int)std::ceil(1.f / h), (int)std::ceil(1.f / h), (int)std::ceil(1.f / h));

    PbfSim* s = pbfSimCreate(&p, PBF_GRID_HASHED, 16, PBF_PATH_REORDERED, 1);
    CHECK(s != nullptr, "pbfSimCreate");
    CHECK(pbfSimSeedLattice(s, 0, 16, 16, 16, r, r, r, 2.f * r, 0.1f * r, 7) == 0, "pbfSimSeedLattice");
    CHECK(pbfSimStep(s, 10) == 0, "pbfSimStep");
    CHECK(pbfSimSynchronize(s) == 0, "pbfSimSynchronize");
    const uint32_t N = p.numParticles;
    std::vector<float4> X(N), V(N);
    CHECK(pbfSimGetState(s, X.data(), V.data(), 0) == 0, "pbfSimGetState");
    for (uint32_t i = 0; i < N; ++i) {
        CHECK(std::isfinite(X[i].x) && std::isfinite(X[i].y) && std::isfinite(X[i].z), "finite positions");
        CHECK(X[i].x >= 0.f && X[i].x <= 1.f && X[i].y >= 0.f && X[i].y <= 1.f && X[i].z >= 0.f && X[i].z <= 1.f,
              "positions inside the box");
    }

    // the drop-in launchers on the simulator's current positions: keys, then the
    // two-call sort, then dense cell ranges
    float4* dPos = nullptr;
    uint32_t *keys = nullptr, *idx = nullptr, *keysS = nullptr, *idxS = nullptr, *start = nullptr, *end = nullptr;
    const uint32_t cells = (uint32_t)(p.grid.dim.x * p.grid.dim.y * p.grid.dim.z);
    CHECK(hipMalloc(&dPos, sizeof(float4) * N) == hipSuccess, "hipMalloc");
    CHECK(hipMalloc(&keys, 4 * N) == hipSuccess && hipMalloc(&idx, 4 * N) == hipSuccess &&
              hipMalloc(&keysS, 4 * N) == hipSuccess && hipMalloc(&idxS, 4 * N) == hipSuccess &&
              hipMalloc(&start, 4 * cells) == hipSuccess && hipMalloc(&end, 4 * cells) == hipSuccess,
          "hipMalloc");
    CHECK(hipMemcpy(dPos, X.data(), sizeof(float4) * N, hipMemcpyHostToDevice) == hipSuccess, "copy");
    hipStream_t st = pbfSimStream(s);
    LaunchHashKeys(keys, idx, dPos, p.grid, N, st);
    size_t tempBytes = 0;
    LaunchSortPairsQuery(&tempBytes, keys, keysS, idx, idxS, N, st);
    void* temp = nullptr;
    CHECK(hipMalloc(&temp, tempBytes) == hipSuccess, "hipMalloc temp");
    LaunchSortPairs(temp, tempBytes, keys, keysS, idx, idxS, N, st);
    LaunchCellRanges(start, end, keysS, N, cells, st);
    CHECK(hipStreamSynchronize(st) == hipSuccess && pbfGetLastError() == 0, "launchers");
    std::vector<uint32_t> hk(N), hs(cells), he(cells);
    hipMemcpy(hk.data(), keysS, 4 * N, hipMemcpyDeviceToHost);
    hipMemcpy(hs.data(), start, 4 * cells, hipMemcpyDeviceToHost);
    hipMemcpy(he.data(), end, 4 * cells, hipMemcpyDeviceToHost);
    uint64_t covered = 0;
    for (uint32_t i = 1; i < N; ++i) CHECK(hk[i - 1] <= hk[i], "sorted keys");
    for (uint32_t c = 0; c < cells; ++c)
        if (he[c] > 0) covered += he[c] - hs[c];
    CHECK(covered == N, "cell ranges cover every particle");

    hipFree(temp); hipFree(dPos); hipFree(keys); hipFree(idx); hipFree(keysS); hipFree(idxS);
    hipFree(start); hipFree(end);

    // three more steps from the simulator's state through the neighbour launchers
    // (dense and compact forms), against the simulator's reference path stepping from
    // the same state: bit for bit
    for (int compact = 0; compact < 2; ++compact) {
        std::vector<float4> Xl = X, Vl = V;
        hipStream_t ls = nullptr;
        CHECK(hipStreamCreateWithFlags(&ls, hipStreamNonBlocking) == hipSuccess, "stream");
        for (int k = 0; k < 3; ++k) CHECK(dropin_step(p, compact != 0, Xl, Vl, ls), "drop-in step");
        (void)hipStreamDestroy(ls);
        PbfSim* r = pbfSimCreate(&p, compact ? PBF_GRID_COMPACT : PBF_GRID_DENSE, 16, PBF_PATH_REFERENCE, 0);
        CHECK(r != nullptr, "pbfSimCreate (reference path)");
        CHECK(pbfSimSetState(r, X.data(), V.data(), N, 0) == 0, "pbfSimSetState");
        CHECK(pbfSimStep(r, 3) == 0, "pbfSimStep (reference path)");
        std::vector<float4> Xr(N), Vr(N);
        CHECK(pbfSimGetState(r, Xr.data(), Vr.data(), 0) == 0, "pbfSimGetState (reference path)");
        pbfSimDestroy(r);
        CHECK(same_bits(Xl, Xr) && same_bits(Vl, Vr),
              compact ? "compact neighbour launchers equal the simulator" : "dense neighbour launchers equal the simulator");
    }
    pbfSimDestroy(s);
    std::printf("OK %u particles, 10 steps, launchers checked, 3 drop-in steps (dense, compact) bit-identical\n", N);
    return 0;
}
