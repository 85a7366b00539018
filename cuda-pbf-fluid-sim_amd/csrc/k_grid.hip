// k_grid.hip — cell keys, dense cell ranges and compact cell segments.
//
// Dense ranges: a fill kernel writes the empty-cell sentinels (start =
// 0xFFFFFFFF, end = 0), then one lane per sorted slot marks the boundaries
// (reference: sim/cell_ranges.cu:21-80; the reference clears with memsets,
// which are replaced by a kernel so the phase can sit in a hipGraph).
// Compact segments: count starts per 1024-slot tile (each lane owns 4
// consecutive slots, one 16-byte load), scan the tile counts, scatter unique
// keys and offsets (reference: sim/cell_compact.cu:96-231).
#include <mutex>
#include <unordered_map>
#include <vector>

#include "grid.hpp"
#include "pbf_common.hpp"
#include "pbf_hip.h"

namespace pbfx {
namespace {

constexpr int kBS = 256;
constexpr int kItems = 4;
constexpr uint32_t kTile = kBS * kItems;

__global__ __launch_bounds__(kBS) void k_hash(uint32_t* __restrict__ keys, uint32_t* __restrict__ idx,
                                              const float4* __restrict__ P, sim::GridBounds g, uint32_t N) {
    const uint32_t i = blockIdx.x * kBS + threadIdx.x;
    if (i >= N) return;
    const float4 p = P[i];
    keys[i] = cell_key(p.x, p.y, p.z, g);
    idx[i] = i;
}

__global__ __launch_bounds__(kBS) void k_fill(uint32_t* __restrict__ p, uint32_t v, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * kBS;
    for (uint64_t i = (uint64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ __launch_bounds__(kBS) void k_fill2(uint32_t* __restrict__ a, uint32_t va, uint32_t* __restrict__ b,
                                               uint32_t vb, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * kBS;
    for (uint64_t i = (uint64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += stride) {
        a[i] = va;
        b[i] = vb;
    }
}

// Boundary marking (reference: cell_ranges.cu:21-51).
__global__ __launch_bounds__(kBS) void k_build_ranges(const uint32_t* __restrict__ keyS, uint32_t N,
                                                      uint32_t numCells, uint32_t* __restrict__ start,
                                                      uint32_t* __restrict__ end) {
    const uint32_t i = blockIdx.x * kBS + threadIdx.x;
    if (i >= N) return;
    const uint32_t k = keyS[i];
    if (k >= numCells) return;
    if (i == 0) {
        start[k] = 0;
    } else {
        const uint32_t kp = keyS[i - 1];
        if (k != kp) {
            if (kp < numCells) end[kp] = i;
            start[k] = i;
        }
    }
    if (i == N - 1) end[k] = N;
}

__device__ __forceinline__ void load4(const uint32_t* __restrict__ keyS, uint64_t first, uint32_t N, uint32_t (&k)[4],
                                      uint32_t& prev) {
    if (first + 4 <= N) {
        const uint4 v = *reinterpret_cast<const uint4*>(keyS + first);
        k[0] = v.x; k[1] = v.y; k[2] = v.z; k[3] = v.w;
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) k[j] = (first + j < N) ? keyS[first + j] : 0u;
    }
    prev = first > 0 && first <= N ? keyS[first - 1] : 0u;
}

// Segment starts in a 4-slot run: bit j set when slot first+j opens a segment.
__device__ __forceinline__ uint32_t start_bits(const uint32_t (&k)[4], uint32_t prev, uint64_t first, uint32_t N) {
    uint32_t bits = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint64_t i = first + j;
        const uint32_t kp = (j == 0) ? prev : k[j - 1];
        const bool f = (i < N) && (i == 0 || k[j] != kp);
        bits |= (f ? 1u : 0u) << j;
    }
    return bits;
}

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* sh, uint32_t& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kBS / 64; ++i) {
        const uint32_t s = sh[i];
        base += (i < w) ? s : 0u;
        tot += s;
    }
    total = tot;
    __syncthreads();
    return base + x - v;
}

__global__ __launch_bounds__(kBS) void k_compact_count(const uint32_t* __restrict__ keyS, uint32_t N,
                                                       uint32_t* __restrict__ blockCounts) {
    __shared__ uint32_t sh[kBS / 64];
    const uint64_t first = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kItems;
    uint32_t k[4], prev;
    load4(keyS, first, N, k, prev);
    const uint32_t c = __popc(start_bits(k, prev, first, N));
    uint32_t total;
    block_excl_scan(c, sh, total);
    if (threadIdx.x == 0) blockCounts[blockIdx.x] = total;
}

// Single-block exclusive scan of the tile counts (tiles <= a few thousand).
__global__ __launch_bounds__(kBS) void k_scan_tiles(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                    uint32_t n) {
    __shared__ uint32_t sh[kBS / 64];
    const uint32_t per = (n + kBS - 1) / kBS;
    const uint32_t b0 = threadIdx.x * per, b1 = min(n, b0 + per);
    uint32_t sum = 0;
    for (uint32_t t = b0; t < b1; ++t) sum += in[t];
    uint32_t total;
    uint32_t run = block_excl_scan(sum, sh, total);
    for (uint32_t t = b0; t < b1; ++t) {
        const uint32_t c = in[t];
        out[t] = run;
        run += c;
    }
}

__global__ __launch_bounds__(kBS) void k_compact_scatter(const uint32_t* __restrict__ keyS, uint32_t N,
                                                         const uint32_t* __restrict__ blockOffsets,
                                                         const uint32_t* __restrict__ blockCounts,
                                                         uint32_t* __restrict__ uniq, uint32_t* __restrict__ off,
                                                         uint32_t* __restrict__ count,
                                                         uint32_t* __restrict__ segOf) {
    __shared__ uint32_t sh[kBS / 64];
    const uint64_t first = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kItems;
    uint32_t k[4], prev;
    load4(keyS, first, N, k, prev);
    const uint32_t bits = start_bits(k, prev, first, N);
    uint32_t total;
    const uint32_t excl = block_excl_scan(__popc(bits), sh, total);
    uint32_t m = blockOffsets[blockIdx.x] + excl;  // segments opened before this lane's run
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint64_t i = first + j;
        if (i >= N) break;
        if ((bits >> j) & 1u) {
            uniq[m] = k[j];
            off[m] = (uint32_t)i;
            ++m;
        }
        if (segOf) segOf[i] = m - 1;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
        const uint32_t M = blockOffsets[blockIdx.x] + blockCounts[blockIdx.x];
        off[M] = N;
        *count = M;
    }
}

// Slot pairs: slot i opens a pair when its place inside its cell (i - off[segOf[i]])
// is even; the pair is (i, i + 1) when i + 1 is in the same cell, else (i, ~0).
// Four slots a lane, as in the compaction: count per tile, scan, scatter.
__device__ __forceinline__ uint32_t pair_start_bits(const uint32_t* __restrict__ segOf,
                                                    const uint32_t* __restrict__ off, uint64_t first, uint32_t N) {
    uint32_t bits = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint64_t i = first + j;
        if (i < N) bits |= (((uint32_t)i - off[segOf[i]]) & 1u) ? 0u : (1u << j);
    }
    return bits;
}

__global__ __launch_bounds__(kBS) void k_pair_count(const uint32_t* __restrict__ segOf,
                                                    const uint32_t* __restrict__ off, uint32_t N,
                                                    uint32_t* __restrict__ blockCounts) {
    __shared__ uint32_t sh[kBS / 64];
    const uint64_t first = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kItems;
    uint32_t total;
    block_excl_scan(__popc(pair_start_bits(segOf, off, first, N)), sh, total);
    if (threadIdx.x == 0) blockCounts[blockIdx.x] = total;
}

__global__ __launch_bounds__(kBS) void k_pair_scatter(const uint32_t* __restrict__ segOf,
                                                      const uint32_t* __restrict__ off, uint32_t N,
                                                      const uint32_t* __restrict__ blockOffsets,
                                                      const uint32_t* __restrict__ blockCounts,
                                                      uint2* __restrict__ pairs, uint32_t* __restrict__ count) {
    __shared__ uint32_t sh[kBS / 64];
    const uint64_t first = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kItems;
    const uint32_t bits = pair_start_bits(segOf, off, first, N);
    uint32_t total;
    uint32_t p = blockOffsets[blockIdx.x] + block_excl_scan(__popc(bits), sh, total);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if ((bits >> j) & 1u) {
            const uint32_t i = (uint32_t)(first + j);
            const uint32_t end = off[segOf[i] + 1u];
            pairs[p++] = make_uint2(i, i + 1u < end ? i + 1u : 0xFFFFFFFFu);
        }
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *count = blockOffsets[blockIdx.x] + blockCounts[blockIdx.x];
}

__global__ void k_compact_empty(uint32_t* off, uint32_t* count) {
    off[0] = 0;
    *count = 0;
}

// Per-device compaction scratch sized by EnsureCellCompactScratch.
struct CompactScratch {
    uint32_t* counts = nullptr;
    uint32_t* offsets = nullptr;
    uint32_t capacity = 0;  // tiles
};
std::mutex g_scratch_mu;
std::vector<CompactScratch> g_scratch;
std::unordered_map<hipStream_t, CompactScratch> g_stream_scratch;  // under g_scratch_mu

CompactScratch* scratch_for_current_device() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch.size() <= dev) g_scratch.resize(dev + 1);
    return &g_scratch[dev];
}

// Allocation-free compaction: the tile counts and tile offsets live in this file-scope
// device array instead of memory the library allocates.  Every GPU holds its own copy
// (at its own address) from module load until the process ends.  Its host-side name
// is only a stand-in: host code uses the device address hipGetSymbolAddress returns,
// looked up once per device on the first call there.  Nothing in it needs clearing
// between calls: k_compact_count writes every tile count, k_scan_tiles every offset.
constexpr uint32_t kNoAllocMaxKeys = 1u << 26;
constexpr uint32_t kNoAllocMaxTiles = kNoAllocMaxKeys / kTile;
static __device__ __align__(256) uint32_t scratch[2 * kNoAllocMaxTiles];

std::mutex g_noalloc_mu;
std::vector<uint32_t*> g_noalloc_ptr;  // per device: the address of `scratch` there

uint32_t* noalloc_scratch_for_current_device() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error(PBFX_E_INVALID_ARGUMENT, "LaunchCellRangesCompactNoAlloc: no current device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if ((int)g_noalloc_ptr.size() <= dev) g_noalloc_ptr.resize(dev + 1, nullptr);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error((int)e, "LaunchCellRangesCompactNoAlloc: hipGetSymbolAddress(scratch)");
            return nullptr;
        }
        g_noalloc_ptr[dev] = static_cast<uint32_t*>(p);
    }
    return g_noalloc_ptr[dev];
}

}  // namespace

uint32_t compact_scratch_elems(uint32_t N) { return ceil_div(N > 0 ? N : 1, kTile); }

bool reserve_stream_compact_scratch(uint32_t N, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    CompactScratch& cs = g_stream_scratch[s];
    const uint32_t need = compact_scratch_elems(N);
    if (need <= cs.capacity) return true;
    if (cs.counts) (void)hipFree(cs.counts);
    if (cs.offsets) (void)hipFree(cs.offsets);
    cs = CompactScratch{};
    if (hipMalloc(&cs.counts, sizeof(uint32_t) * need) != hipSuccess ||
        hipMalloc(&cs.offsets, sizeof(uint32_t) * need) != hipSuccess) {
        if (cs.counts) (void)hipFree(cs.counts);
        g_stream_scratch.erase(s);
        set_error(PBFX_E_CAPACITY, "stream compaction scratch: allocation failed");
        return false;
    }
    cs.capacity = need;
    return true;
}

void release_stream_compact_scratch(hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    const auto it = g_stream_scratch.find(s);
    if (it == g_stream_scratch.end()) return;
    if (it->second.counts) (void)hipFree(it->second.counts);
    if (it->second.offsets) (void)hipFree(it->second.offsets);
    g_stream_scratch.erase(it);
}

void fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s) {
    if (n == 0) return;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + kBS - 1) / kBS, 4096);
    k_fill<<<blocks, kBS, 0, s>>>(p, value, n);
    check_launch("fill_u32");
}

void compact_segments(uint32_t* uniq, uint32_t* offsets, uint32_t* count, uint32_t* segOf,
                      const uint32_t* keysSorted, uint32_t N, uint32_t* blockCounts, uint32_t* blockOffsets,
                      hipStream_t s) {
    if (N == 0) {
        k_compact_empty<<<1, 1, 0, s>>>(offsets, count);
        check_launch("compact_segments");
        return;
    }
    const uint32_t tiles = compact_scratch_elems(N);
    k_compact_count<<<tiles, kBS, 0, s>>>(keysSorted, N, blockCounts);
    k_scan_tiles<<<1, kBS, 0, s>>>(blockCounts, blockOffsets, tiles);
    k_compact_scatter<<<tiles, kBS, 0, s>>>(keysSorted, N, blockOffsets, blockCounts, uniq, offsets, count, segOf);
    check_launch("compact_segments");
}

void build_slot_pairs(uint2* pairs, uint32_t* count, const uint32_t* segOf, const uint32_t* offsets, uint32_t N,
                      uint32_t* blockCounts, uint32_t* blockOffsets, hipStream_t s) {
    if (N == 0) {
        fill_u32(count, 0u, 1, s);
        return;
    }
    const uint32_t tiles = compact_scratch_elems(N);
    k_pair_count<<<tiles, kBS, 0, s>>>(segOf, offsets, N, blockCounts);
    k_scan_tiles<<<1, kBS, 0, s>>>(blockCounts, blockOffsets, tiles);
    k_pair_scatter<<<tiles, kBS, 0, s>>>(segOf, offsets, N, blockOffsets, blockCounts, pairs, count);
    check_launch("build_slot_pairs");
}

}  // namespace pbfx

using namespace pbfx;

extern "C" void LaunchHashKeys(uint32_t* keys, uint32_t* indices, const float4* pos, sim::GridBounds grid,
                               uint32_t N, hipStream_t s) {
    if (N == 0) return;
    if (grid.dim.x <= 0 || grid.dim.y <= 0 || grid.dim.z <= 0) {
        set_error(PBFX_E_INVALID_ARGUMENT, "LaunchHashKeys: empty grid");
        return;
    }
    k_hash<<<ceil_div(N, kBS), kBS, 0, s>>>(keys, indices, pos, grid, N);
    check_launch("LaunchHashKeys");
}

extern "C" void LaunchCellRanges(uint32_t* cellStart, uint32_t* cellEnd, const uint32_t* keysSorted, uint32_t N,
                                 uint32_t numCells, hipStream_t s) {
    if (numCells == 0) return;
    const uint32_t fb = (uint32_t)std::min<uint64_t>(ceil_div(numCells, kBS), 4096);
    k_fill2<<<fb, kBS, 0, s>>>(cellStart, 0xFFFFFFFFu, cellEnd, 0u, numCells);
    if (N > 0) k_build_ranges<<<ceil_div(N, kBS), kBS, 0, s>>>(keysSorted, N, numCells, cellStart, cellEnd);
    check_launch("LaunchCellRanges");
}

extern "C" bool EnsureCellCompactScratch(uint32_t N, uint32_t /*threadsPerBlock*/) {
    CompactScratch* cs = scratch_for_current_device();
    if (!cs) return false;
    const uint32_t need = compact_scratch_elems(N);
    if (need <= cs->capacity) return true;
    if (cs->counts) (void)hipFree(cs->counts);
    if (cs->offsets) (void)hipFree(cs->offsets);
    cs->counts = cs->offsets = nullptr;
    cs->capacity = 0;
    if (hipMalloc(&cs->counts, sizeof(uint32_t) * need) != hipSuccess ||
        hipMalloc(&cs->offsets, sizeof(uint32_t) * need) != hipSuccess) {
        set_error(PBFX_E_CAPACITY, "EnsureCellCompactScratch: allocation failed");
        return false;
    }
    cs->capacity = need;
    return true;
}

extern "C" void LaunchCellRangesCompact(uint32_t* d_uniqueKeys, uint32_t* d_offsets, uint32_t* d_compactCount,
                                        const uint32_t* d_keysSorted, uint32_t N, hipStream_t s) {
    if (N == 0) {
        compact_segments(d_uniqueKeys, d_offsets, d_compactCount, nullptr, d_keysSorted, 0, nullptr, nullptr, s);
        return;
    }
    // A stream with its own scratch (reserved with its drop-in staging) uses it, so
    // compactions on different streams never share tile counts.
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        const auto it = g_stream_scratch.find(s);
        if (it != g_stream_scratch.end() && it->second.capacity >= compact_scratch_elems(N)) {
            compact_segments(d_uniqueKeys, d_offsets, d_compactCount, nullptr, d_keysSorted, N, it->second.counts,
                             it->second.offsets, s);
            return;
        }
    }
    // Like the reference (cell_compact.cu:204-210) the device-wide scratch is grown
    // here when needed; graph capture requires EnsureCellCompactScratch beforehand.
    if (!EnsureCellCompactScratch(N, kBS)) return;
    CompactScratch* cs = scratch_for_current_device();
    compact_segments(d_uniqueKeys, d_offsets, d_compactCount, nullptr, d_keysSorted, N, cs->counts, cs->offsets, s);
}

extern "C" void LaunchCellRangesCompactNoAlloc(uint32_t* d_uniqueKeys, uint32_t* d_offsets,
                                               uint32_t* d_compactCount, const uint32_t* d_keysSorted, uint32_t N,
                                               hipStream_t s) {
    if (N == 0) {
        compact_segments(d_uniqueKeys, d_offsets, d_compactCount, nullptr, d_keysSorted, 0, nullptr, nullptr, s);
        return;
    }
    if (N > kNoAllocMaxKeys) {
        set_error(PBFX_E_CAPACITY, "LaunchCellRangesCompactNoAlloc: more than 2^26 sorted keys (its static scratch)");
        return;
    }
    uint32_t* sc = noalloc_scratch_for_current_device();
    if (!sc) return;
    compact_segments(d_uniqueKeys, d_offsets, d_compactCount, nullptr, d_keysSorted, N, sc, sc + kNoAllocMaxTiles, s);
}
