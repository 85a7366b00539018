// k_stats.hip — neighbourhood statistics (mean neighbours, speed, kernel sum, density).
//
// Drop-in for the reference's diagnostics (reference: sim/stats.cu:25-167,
// sim/stats_bruteforce.cu:8-119).  Per particle the sums are formed in float in
// the reference's order; across particles they are reduced in double through
// per-block partials summed in block order, so the result is deterministic (the
// reference accumulates with float atomics, whose order varies run to run).
// Like the reference these launchers read their result back to the host and
// synchronise the stream; they are diagnostics, not part of the captured step.
// Their small device scratch is allocated once per device and reused.
#include <mutex>
#include <vector>

#include "pbf_common.hpp"
#include "pbf_hip.h"
#include "stats.hpp"

namespace pbfx {
namespace {

constexpr int kBS = 256;

// sums[0..3] = neighbours, speed, sum W, samples (per block, double)
struct BlockSums {
    double v[4];
};

__device__ __forceinline__ void block_reduce_store(double a, double b, double c, double d, BlockSums* out) {
    __shared__ double sh[4][kBS];
    sh[0][threadIdx.x] = a;
    sh[1][threadIdx.x] = b;
    sh[2][threadIdx.x] = c;
    sh[3][threadIdx.x] = d;
    __syncthreads();
    for (int w = kBS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) out[blockIdx.x].v[q] = sh[q][0];
    }
}

// Grid statistics over original particle order (reference: stats.cu:25-98):
// every sampled particle, neighbours j != i with r2 < h2 (strict, as the reference).
__global__ __launch_bounds__(kBS) void k_stats_grid(const float4* __restrict__ P, const float4* __restrict__ V,
                                                    const uint32_t* __restrict__ idxS,
                                                    const uint32_t* __restrict__ cellStart,
                                                    const uint32_t* __restrict__ cellEnd, sim::GridBounds g,
                                                    sim::KernelCoeffs kc, uint32_t N, uint32_t numCells,
                                                    uint32_t stride, BlockSums* out) {
    const uint32_t i = blockIdx.x * kBS + threadIdx.x;
    double nb = 0.0, sp = 0.0, sw = 0.0, cnt = 0.0;
    if (i < N && !(stride > 1 && (i % stride) != 0)) {
        const float4 pi = P[i];
        const float4 vi = V[i];
        sp = sqrtf(vi.x * vi.x + vi.y * vi.y + vi.z * vi.z);
        const int cx = clampi(sat_f2i(floorf((pi.x - g.mins.x) / g.cellSize)), 0, g.dim.x - 1);
        const int cy = clampi(sat_f2i(floorf((pi.y - g.mins.y) / g.cellSize)), 0, g.dim.y - 1);
        const int cz = clampi(sat_f2i(floorf((pi.z - g.mins.z) / g.cellSize)), 0, g.dim.z - 1);
        const int reach = (g.cellSize > 0.f) ? max(1, sat_f2i(ceilf(kc.h / g.cellSize))) : 1;
        float neighbors = 0.0f, sumW = 0.0f;
        for (int dz = -reach; dz <= reach; ++dz) {
            const int z = cz + dz;
            if (z < 0 || z >= g.dim.z) continue;
            for (int dy = -reach; dy <= reach; ++dy) {
                const int y = cy + dy;
                if (y < 0 || y >= g.dim.y) continue;
                for (int dx = -reach; dx <= reach; ++dx) {
                    const int x = cx + dx;
                    if (x < 0 || x >= g.dim.x) continue;
                    const uint32_t c = (uint32_t)(x + y * g.dim.x + z * g.dim.x * g.dim.y);
                    if (c >= numCells) continue;
                    const uint32_t b = cellStart[c], e = cellEnd[c];
                    if (b == 0xFFFFFFFFu || b >= e) continue;
                    for (uint32_t k = b; k < e; ++k) {
                        const uint32_t j = idxS[k];
                        if (j == i) continue;
                        const float4 pj = P[j];
                        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
                        const float r2 = rx * rx + ry * ry + rz * rz;
                        if (r2 < kc.h2) {
                            neighbors += 1.0f;
                            const float t = kc.h2 - r2;
                            sumW += kc.poly6 * t * t * t;
                        }
                    }
                }
            }
        }
        nb = neighbors;
        sw = sumW;
        cnt = 1.0;
    }
    block_reduce_store(nb, sp, sw, cnt, out);
}

// O(N^2) statistics (reference: stats_bruteforce.cu:8-62): thread t of T takes
// i = t, t+T, ... (stride-filtered) until it has maxISamples samples.
__global__ __launch_bounds__(kBS) void k_stats_brute(const float4* __restrict__ P, const float4* __restrict__ V,
                                                     sim::KernelCoeffs kc, uint32_t N, uint32_t stride,
                                                     uint32_t maxISamples, uint32_t T, BlockSums* out) {
    const uint32_t tid = blockIdx.x * kBS + threadIdx.x;
    double nb = 0.0, sp = 0.0, sw = 0.0, cnt = 0.0;
    uint32_t taken = 0;
    if (tid < T) {
        for (uint32_t i = tid; i < N && taken < maxISamples; i += T) {
            if (stride > 1 && (i % stride) != 0) continue;
            ++taken;
            const float4 pi = P[i];
            const float4 vi = V[i];
            sp += (double)sqrtf(vi.x * vi.x + vi.y * vi.y + vi.z * vi.z);
            float neighbors = 0.0f, sumW = 0.0f;
            for (uint32_t j = 0; j < N; ++j) {
                if (j == i) continue;
                const float4 pj = P[j];
                const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
                const float r2 = rx * rx + ry * ry + rz * rz;
                if (r2 < kc.h2) {
                    neighbors += 1.0f;
                    const float t = kc.h2 - r2;
                    sumW += kc.poly6 * t * t * t;
                }
            }
            nb += neighbors;
            sw += sumW;
            cnt += 1.0;
        }
    }
    block_reduce_store(nb, sp, sw, cnt, out);
}

__global__ void k_sum_blocks(const BlockSums* __restrict__ in, uint32_t n, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t b = 0; b < n; ++b)
        for (int q = 0; q < 4; ++q) s[q] += in[b].v[q];
    for (int q = 0; q < 4; ++q) out[q] = s[q];
}

// Density-constraint statistics over cell-sorted slots, with the lambda pass's
// neighbour order and cap: per slot rho = m * sum W (float), err = rho/rho0 - 1.
// Per block: sum err, max err, sum neighbours, sum |v|, sum max(err, 0).
struct DensitySums {
    double sumErr, maxErr, sumN, sumSpeed, sumPos;
};

__global__ __launch_bounds__(kBS) void k_density_sorted(const float4* __restrict__ Ps, const float4* __restrict__ V,
                                                        const uint32_t* __restrict__ idxS,
                                                        const uint32_t* __restrict__ segOf, SpanTable st,
                                                        sim::DeviceParams dp, uint32_t N,
                                                        DensitySums* __restrict__ out) {
    __shared__ double sh[5][kBS];
    const uint32_t s = blockIdx.x * kBS + threadIdx.x;
    double e = 0.0, emax = -1e300, n = 0.0, sp = 0.0, pos = 0.0;
    if (s < N) {
        const float4 pi = Ps[s];
        const uint32_t m = segOf[s];
        const float h2 = dp.kernel.h2;
        const float mp = dp.particleMass * dp.kernel.poly6;
        const int cap = dp.maxNeighbors;
        float density = 0.f;
        int cnt = 0;
        for (int r = 0; r < st.rows && !(cap > 0 && cnt >= cap); ++r) {
            const uint2 spn = st.spans[(uint64_t)r * st.stride + m];
            for (uint32_t k = spn.x; k < spn.y; ++k) {
                if (k == s) continue;
                const float4 pj = Ps[k];
                const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
                const float r2 = rx * rx + ry * ry + rz * rz;
                if (r2 > h2) continue;
                ++cnt;
                const float hr2 = h2 - r2;
                density += mp * hr2 * hr2 * hr2;
                if (cap > 0 && cnt >= cap) break;
            }
        }
        e = (double)density / (double)dp.restDensity - 1.0;
        emax = e;
        pos = e > 0.0 ? e : 0.0;
        n = cnt;
        const float4 v = V[idxS[s]];
        sp = sqrt((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z);
    }
    sh[0][threadIdx.x] = e;
    sh[1][threadIdx.x] = emax;
    sh[2][threadIdx.x] = n;
    sh[3][threadIdx.x] = sp;
    sh[4][threadIdx.x] = pos;
    __syncthreads();
    for (int w = kBS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
            sh[1][threadIdx.x] = fmax(sh[1][threadIdx.x], sh[1][threadIdx.x + w]);
            sh[2][threadIdx.x] += sh[2][threadIdx.x + w];
            sh[3][threadIdx.x] += sh[3][threadIdx.x + w];
            sh[4][threadIdx.x] += sh[4][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = DensitySums{sh[0][0], sh[1][0], sh[2][0], sh[3][0], sh[4][0]};
}

__global__ void k_density_finish(const DensitySums* __restrict__ in, uint32_t n, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    DensitySums t{0.0, -1e300, 0.0, 0.0, 0.0};
    for (uint32_t b = 0; b < n; ++b) {
        t.sumErr += in[b].sumErr;
        t.maxErr = fmax(t.maxErr, in[b].maxErr);
        t.sumN += in[b].sumN;
        t.sumSpeed += in[b].sumSpeed;
        t.sumPos += in[b].sumPos;
    }
    out[0] = t.sumErr;
    out[1] = t.maxErr;
    out[2] = t.sumN;
    out[3] = t.sumSpeed;
    out[4] = t.sumPos;
}

// Per-device scratch for the statistics reductions.
struct StatsScratch {
    void* buf = nullptr;
    size_t bytes = 0;
};
std::mutex g_mu;
std::vector<StatsScratch> g_scratch;

void* stats_scratch(size_t bytes) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    if ((int)g_scratch.size() <= dev) g_scratch.resize(dev + 1);
    StatsScratch& s = g_scratch[dev];
    if (s.bytes < bytes) {
        if (s.buf) (void)hipFree(s.buf);
        s.buf = nullptr;
        s.bytes = 0;
        if (hipMalloc(&s.buf, bytes) != hipSuccess) {
            set_error(PBFX_E_CAPACITY, "stats scratch allocation failed");
            return nullptr;
        }
        s.bytes = bytes;
    }
    return s.buf;
}

bool finish_stats(const BlockSums* partial, uint32_t blocks, float mass, double* oN, double* oS, double* oRR,
                  double* oR, hipStream_t s, double* dOut) {
    k_sum_blocks<<<1, 1, 0, s>>>(partial, blocks, dOut);
    double h[4] = {0, 0, 0, 0};
    if (hipMemcpyAsync(h, dOut, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        set_error(PBFX_E_INVALID_ARGUMENT, "stats readback failed");
        return false;
    }
    const double samples = h[3] > 0 ? h[3] : 1.0;
    if (oN) *oN = h[0] / samples;
    if (oS) *oS = h[1] / samples;
    if (oRR) *oRR = h[2] / samples;
    if (oR) *oR = h[2] / samples * (double)mass;
    return true;
}

void zero_outputs(double* a, double* b, double* c, double* d) {
    if (a) *a = 0.0;
    if (b) *b = 0.0;
    if (c) *c = 0.0;
    if (d) *d = 0.0;
}

}  // namespace

bool density_stats_sorted(const float4* Ps, const float4* V, const uint32_t* idxS, const uint32_t* segOf,
                          const SpanTable& st, const sim::DeviceParams& dp, uint32_t N, double* out5,
                          hipStream_t s) {
    for (int q = 0; q < 5; ++q) out5[q] = 0.0;
    if (N == 0) return true;
    const uint32_t blocks = ceil_div(N, kBS);
    const size_t bytes = sizeof(DensitySums) * blocks + 64;
    char* base = static_cast<char*>(stats_scratch(bytes));
    if (!base) return false;
    DensitySums* part = reinterpret_cast<DensitySums*>(base);
    double* dOut = reinterpret_cast<double*>(base + sizeof(DensitySums) * blocks);
    k_density_sorted<<<blocks, kBS, 0, s>>>(Ps, V, idxS, segOf, st, dp, N, part);
    k_density_finish<<<1, 1, 0, s>>>(part, blocks, dOut);
    double h[5];
    if (hipMemcpyAsync(h, dOut, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        set_error(PBFX_E_INVALID_ARGUMENT, "density stats readback failed");
        return false;
    }
    out5[0] = h[0] / N;
    out5[1] = h[1];
    out5[2] = h[2] / N;
    out5[3] = h[3] / N;
    out5[4] = h[4] / N;
    return true;
}

}  // namespace pbfx

using namespace pbfx;

extern "C" bool LaunchComputeStats(const float4* pos_pred, const float4* vel, const uint32_t* indices_sorted,
                                   const uint32_t* cellStart, const uint32_t* cellEnd, sim::GridBounds grid,
                                   sim::KernelCoeffs kc, float particleMass, uint32_t N, uint32_t numCells,
                                   uint32_t sampleStride, double* outAvgNeighbors, double* outAvgSpeed,
                                   double* outAvgRhoRel, double* outAvgRho, hipStream_t s) {
    if (!pos_pred || !vel || !indices_sorted || !cellStart || !cellEnd || N == 0) {
        zero_outputs(outAvgNeighbors, outAvgSpeed, outAvgRhoRel, outAvgRho);
        return true;
    }
    const uint32_t blocks = ceil_div(N, kBS);
    char* base = static_cast<char*>(stats_scratch(sizeof(BlockSums) * blocks + 64));
    if (!base) return false;
    BlockSums* part = reinterpret_cast<BlockSums*>(base);
    k_stats_grid<<<blocks, kBS, 0, s>>>(pos_pred, vel, indices_sorted, cellStart, cellEnd, grid, kc, N, numCells,
                                       sampleStride, part);
    check_launch("LaunchComputeStats");
    return finish_stats(part, blocks, particleMass, outAvgNeighbors, outAvgSpeed, outAvgRhoRel, outAvgRho, s,
                        reinterpret_cast<double*>(base + sizeof(BlockSums) * blocks));
}

extern "C" bool LaunchComputeStatsBruteforce(const float4* pos_pred, const float4* vel, sim::KernelCoeffs kc,
                                             float particleMass, uint32_t N, uint32_t sampleStride,
                                             uint32_t maxISamples, double* outAvgNeighbors, double* outAvgSpeed,
                                             double* outAvgRhoRel, double* outAvgRho, hipStream_t s) {
    if (!pos_pred || !vel || N == 0 || maxISamples == 0) {
        zero_outputs(outAvgNeighbors, outAvgSpeed, outAvgRhoRel, outAvgRho);
        return true;
    }
    // thread count of the reference launch: ceil(maxISamples / 128) blocks of 128
    const uint32_t T = ceil_div(maxISamples, 128) * 128u;
    const uint32_t blocks = ceil_div(T, kBS);
    char* base = static_cast<char*>(stats_scratch(sizeof(BlockSums) * blocks + 64));
    if (!base) return false;
    BlockSums* part = reinterpret_cast<BlockSums*>(base);
    k_stats_brute<<<blocks, kBS, 0, s>>>(pos_pred, vel, kc, N, sampleStride, maxISamples, T, part);
    check_launch("LaunchComputeStatsBruteforce");
    return finish_stats(part, blocks, particleMass, outAvgNeighbors, outAvgSpeed, outAvgRhoRel, outAvgRho, s,
                        reinterpret_cast<double*>(base + sizeof(BlockSums) * blocks));
}
