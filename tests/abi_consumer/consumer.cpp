// consumer.cpp — a C++ caller compiled against the public headers only, as the
// reference's own host code would be after the switch: the parameter structs from
// pbf_params.h, the drop-in launchers from pbf_hip.h and the simulator from pbf_sim.h.
// Exit status 0 and a final "OK" line mean every check passed.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
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
    pbfSimDestroy(s);
    std::printf("OK %u particles, 10 steps, launchers checked\n", N);
    return 0;
}
