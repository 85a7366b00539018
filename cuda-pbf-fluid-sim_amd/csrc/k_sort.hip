// k_sort.hip — stable LSD radix sort of (key, value) pairs for wave64.
//
// Per digit pass (digits of at most 8 bits; the bits to sort are spread evenly
// over ceil(bits / maxDigitBits) passes), three kernels:
//   hist     each tile counts its digits (wave64 match via 8 ballots, one LDS
//            atomic per distinct digit per wave), writes counts[digit][tile];
//   scanrows one block per digit turns its row into exclusive tile offsets and
//            the digit total;
//   scatter  each tile ranks its elements stably (per-wave LDS digit counters
//            advanced by the match leader), reorders the tile in LDS by digit,
//            then writes runs of equal digits to consecutive addresses.
// Stability: inside a tile the rank order is (wave, item, lane) = element
// order, and tiles are offset in tile order, so equal keys keep input order —
// the contract of the reference's cub::DeviceRadixSort::SortPairs
// (reference: sim/sort_pairs.cu:18-27).
// Only the key bits that can be non-zero are sorted when the caller knows them
// (pbfSortPairsBits); LaunchSortPairs sorts all 32 bits like the reference.
#include "pbf_common.hpp"
#include "pbf_hip.h"
#include "radix_sort.hpp"

namespace pbfx {
namespace {

constexpr int kRadixBits = 8;
constexpr int kR = 1 << kRadixBits;
constexpr int kBS = 256;
constexpr int kWaves = kBS / 64;

__device__ __forceinline__ uint32_t popc_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Lanes of this wave (among `valid`) holding the same digit of `bits` bits (<= 8).
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid, int bits) {
    uint64_t m = __ballot(valid);
    for (int b = 0; b < bits; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

// Exclusive scan of one value per thread over a 256-thread block.
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
    for (int i = 0; i < kWaves; ++i) {
        const uint32_t s = sh[i];
        base += (i < w) ? s : 0u;
        tot += s;
    }
    total = tot;
    __syncthreads();
    return base + x - v;
}

template <int ITEMS>
__global__ __launch_bounds__(kBS) void k_radix_hist(const uint32_t* __restrict__ keys, uint32_t N, int shift,
                                                    uint32_t mask, int bits, uint32_t* __restrict__ counts,
                                                    uint32_t numTiles) {
    constexpr uint32_t T = kBS * ITEMS;
    __shared__ uint32_t hist[kR];
    const uint32_t tile = blockIdx.x;
    hist[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)tile * T;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint64_t idx = base + (uint64_t)k * kBS + threadIdx.x;
        const bool valid = idx < N;
        const uint32_t d = valid ? ((keys[idx] >> shift) & mask) : 0u;
        const uint64_t peers = match_digit(d, valid, bits);
        if (valid && popc_below(peers) == 0) atomicAdd(&hist[d], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    counts[(uint64_t)threadIdx.x * numTiles + tile] = hist[threadIdx.x];
}

// One block per digit row: exclusive scan over tiles, row total to totals[d].
__global__ __launch_bounds__(kBS) void k_radix_scan_rows(uint32_t* __restrict__ counts, uint32_t numTiles,
                                                         uint32_t* __restrict__ totals) {
    __shared__ uint32_t sh[kWaves];
    uint32_t* row = counts + (uint64_t)blockIdx.x * numTiles;
    const uint32_t per = (numTiles + kBS - 1) / kBS;
    const uint32_t b0 = threadIdx.x * per;
    const uint32_t b1 = min(numTiles, b0 + per);
    uint32_t sum = 0;
    for (uint32_t t = b0; t < b1; ++t) sum += row[t];
    uint32_t total;
    uint32_t run = block_excl_scan(sum, sh, total);
    for (uint32_t t = b0; t < b1; ++t) {
        const uint32_t c = row[t];
        row[t] = run;
        run += c;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

template <int ITEMS>
__global__ __launch_bounds__(kBS) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                       const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
                                                       uint32_t* __restrict__ vout, uint32_t N, int shift,
                                                       uint32_t mask, int bits, const uint32_t* __restrict__ counts,
                                                       const uint32_t* __restrict__ totals, uint32_t numTiles) {
    constexpr uint32_t T = kBS * ITEMS;
    constexpr uint32_t TW = T / kWaves;  // elements per wave sub-tile
    __shared__ uint32_t sKeys[T];
    __shared__ uint32_t sVals[T];
    __shared__ uint32_t sWaveHist[kWaves][kR];
    __shared__ uint32_t sDigitStart[kR];
    __shared__ uint32_t sGlobalBase[kR];
    __shared__ uint32_t sScan[kWaves];

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63, w = tid >> 6;
    const uint32_t tile = blockIdx.x;
    const uint64_t tileBase = (uint64_t)tile * T;

    // global base of every digit for this tile
    {
        uint32_t tot;
        const uint32_t digitBase = block_excl_scan(totals[tid], sScan, tot);
        sGlobalBase[tid] = digitBase + counts[(uint64_t)tid * numTiles + tile];
    }
#pragma unroll
    for (int q = 0; q < kWaves; ++q) sWaveHist[q][tid] = 0;
    __syncthreads();

    uint32_t key[ITEMS], val[ITEMS], rank[ITEMS];
    const uint64_t waveBase = tileBase + (uint64_t)w * TW;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint64_t idx = waveBase + (uint64_t)k * 64 + lane;
        const bool valid = idx < N;
        key[k] = valid ? kin[idx] : 0u;
        val[k] = valid ? (vin ? vin[idx] : (uint32_t)idx) : 0u;
    }
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint64_t idx = waveBase + (uint64_t)k * 64 + lane;
        const bool valid = idx < N;
        const uint32_t d = (key[k] >> shift) & mask;
        const uint64_t peers = match_digit(d, valid, bits);
        const uint32_t pre = sWaveHist[w][d];
        rank[k] = pre + popc_below(peers);
        const uint32_t leader = 63u - (uint32_t)__clzll(peers);
        if (valid && lane == leader) sWaveHist[w][d] = pre + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // digit tid: prefix across waves, tile-local digit start
    {
        uint32_t c[kWaves];
        uint32_t sum = 0;
#pragma unroll
        for (int q = 0; q < kWaves; ++q) {
            c[q] = sWaveHist[q][tid];
            sWaveHist[q][tid] = sum;
            sum += c[q];
        }
        uint32_t tot;
        sDigitStart[tid] = block_excl_scan(sum, sScan, tot);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint64_t idx = waveBase + (uint64_t)k * 64 + lane;
        if (idx < N) {
            const uint32_t d = (key[k] >> shift) & mask;
            const uint32_t pos = sDigitStart[d] + sWaveHist[w][d] + rank[k];
            sKeys[pos] = key[k];
            sVals[pos] = val[k];
        }
    }
    __syncthreads();
    const uint32_t nValid = (uint32_t)min((uint64_t)T, (uint64_t)N - tileBase);
    for (uint32_t j = tid; j < nValid; j += kBS) {
        const uint32_t kk = sKeys[j];
        const uint32_t d = (kk >> shift) & mask;
        const uint32_t o = sGlobalBase[d] + (j - sDigitStart[d]);
        kout[o] = kk;
        vout[o] = sVals[j];
    }
}

__global__ void k_copy_pairs(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                             uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t N) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    kout[i] = kin[i];
    vout[i] = vin ? vin[i] : i;
}

int items_for(uint32_t N) {
    if (N <= (2u << 20)) return 4;
    if (N <= (8u << 20)) return 8;
    return 16;
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

struct Layout {
    int items;
    uint32_t tile, numTiles;
    size_t offK, offV, offCounts, offTotals, bytes;
};

Layout layout_for(uint32_t N, int items = 0) {
    Layout L{};
    L.items = (items == 4 || items == 8 || items == 16) ? items : items_for(N);
    L.tile = kBS * L.items;
    L.numTiles = ceil_div(N > 0 ? N : 1, L.tile);
    L.offK = 0;
    L.offV = align256((size_t)N * 4);
    L.offCounts = L.offV + align256((size_t)N * 4);
    L.offTotals = L.offCounts + align256((size_t)kR * L.numTiles * 4);
    L.bytes = L.offTotals + align256(kR * 4);
    return L;
}

// One pass over the digit (key >> shift) & mask of `bits` bits.  Digits above the
// mask never occur: their count rows stay zero from the histogram and are not
// scanned, and their totals (possibly left by an earlier, wider pass) only shift the
// bases of those absent digits.
template <int ITEMS>
void launch_pass(const Layout& L, const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout,
                 uint32_t N, int shift, uint32_t mask, int bits, uint32_t* counts, uint32_t* totals,
                 hipStream_t s) {
    k_radix_hist<ITEMS><<<L.numTiles, kBS, 0, s>>>(kin, N, shift, mask, bits, counts, L.numTiles);
    k_radix_scan_rows<<<(int)mask + 1, kBS, 0, s>>>(counts, L.numTiles, totals);
    k_radix_scatter<ITEMS><<<L.numTiles, kBS, 0, s>>>(kin, vin, kout, vout, N, shift, mask, bits, counts, totals,
                                                      L.numTiles);
}

}  // namespace

// Sized for the smallest tile (4 items a thread: the most tiles), so that every
// items-per-thread choice fits.
size_t radix_sort_temp_bytes(uint32_t N, int, int) { return layout_for(N, 4).bytes; }

bool radix_sort_pairs(void* temp, size_t tempBytes, const uint32_t* kin, uint32_t* kout, const uint32_t* vin,
                      uint32_t* vout, uint32_t N, int beginBit, int endBit, hipStream_t s, int items,
                      int digitBits) {
    if (N == 0) return true;
    if (digitBits == 0) digitBits = kRadixBits;
    if (digitBits < 1 || digitBits > kRadixBits) {
        set_error(PBFX_E_INVALID_ARGUMENT, "radix sort: digit width must be 1..8 bits");
        return false;
    }
    const Layout L = layout_for(N, items);
    if (!temp || tempBytes < L.bytes) {
        set_error(PBFX_E_CAPACITY, "radix sort: temp storage too small");
        return false;
    }
    if (kin == kout || (vin && vin == vout)) {
        set_error(PBFX_E_ALIAS, "radix sort: output aliases input");
        return false;
    }
    beginBit = beginBit < 0 ? 0 : beginBit;
    endBit = endBit > 32 ? 32 : endBit;
    char* base = static_cast<char*>(temp);
    uint32_t* altK = reinterpret_cast<uint32_t*>(base + L.offK);
    uint32_t* altV = reinterpret_cast<uint32_t*>(base + L.offV);
    uint32_t* counts = reinterpret_cast<uint32_t*>(base + L.offCounts);
    uint32_t* totals = reinterpret_cast<uint32_t*>(base + L.offTotals);
    const int bits = endBit - beginBit;
    if (bits <= 0) {
        k_copy_pairs<<<ceil_div(N, 256), 256, 0, s>>>(kin, vin, kout, vout, N);
        check_launch("radix sort copy");
        return true;
    }
    const int passes = (bits + digitBits - 1) / digitBits;
    // spread the bits evenly over the passes (e.g. 21 bits at 8 -> 7, 7, 7)
    const int width = (bits + passes - 1) / passes;
    const uint32_t* srcK = kin;
    const uint32_t* srcV = vin;
    for (int p = 0; p < passes; ++p) {
        const int shift = beginBit + p * width;
        const int wbits = (shift + width <= endBit) ? width : (endBit - shift);
        const uint32_t mask = (wbits >= 32) ? 0xFFFFFFFFu : ((1u << wbits) - 1u);
        const bool toOut = ((passes - 1 - p) % 2) == 0;
        uint32_t* dK = toOut ? kout : altK;
        uint32_t* dV = toOut ? vout : altV;
        switch (L.items) {
            case 4: launch_pass<4>(L, srcK, srcV, dK, dV, N, shift, mask, wbits, counts, totals, s); break;
            case 8: launch_pass<8>(L, srcK, srcV, dK, dV, N, shift, mask, wbits, counts, totals, s); break;
            default: launch_pass<16>(L, srcK, srcV, dK, dV, N, shift, mask, wbits, counts, totals, s); break;
        }
        srcK = dK;
        srcV = dV;
    }
    check_launch("radix sort");
    return true;
}

}  // namespace pbfx

using namespace pbfx;

extern "C" void LaunchSortPairsQuery(size_t* tempBytes, const uint32_t*, uint32_t*, const uint32_t*, uint32_t*,
                                     uint32_t N, hipStream_t) {
    if (tempBytes) *tempBytes = radix_sort_temp_bytes(N, 0, 32);
}

extern "C" void LaunchSortPairs(void* d_temp_storage, size_t tempBytes, uint32_t* d_keys_in, uint32_t* d_keys_out,
                                uint32_t* d_vals_in, uint32_t* d_vals_out, uint32_t N, hipStream_t s) {
    radix_sort_pairs(d_temp_storage, tempBytes, d_keys_in, d_keys_out, d_vals_in, d_vals_out, N, 0, 32, s);
}

extern "C" void pbfSortPairsBits(void* d_temp_storage, size_t* tempBytes, const uint32_t* d_keys_in,
                                 uint32_t* d_keys_out, const uint32_t* d_vals_in, uint32_t* d_vals_out, uint32_t N,
                                 int beginBit, int endBit, hipStream_t s) {
    if (!tempBytes) { set_error(PBFX_E_INVALID_ARGUMENT, "pbfSortPairsBits: tempBytes is null"); return; }
    if (!d_temp_storage) {
        *tempBytes = radix_sort_temp_bytes(N, beginBit, endBit);
        return;
    }
    radix_sort_pairs(d_temp_storage, *tempBytes, d_keys_in, d_keys_out, d_vals_in, d_vals_out, N, beginBit, endBit,
                     s);
}
