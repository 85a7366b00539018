// k_solve_sorted.hip — the cell-reordered PBF step.
//
// Layout inside a step: Ps (float4, cell-sorted) holds the predicted position
// in xyz and, between the lambda and delta passes, lambda in w, so a delta
// neighbour costs one 16-byte load; Vs (float4, cell-sorted) holds velocities
// for XSPH.  Every occupied cell carries its (2*reach+1)^2 neighbour-row spans,
// built once per step, so the five neighbour passes read spans instead of
// searching.  Slot s of a neighbour span is visited in the same order as the
// reference's per-cell loops (rows dz-major, then dy; cells ascending x; slots
// ascending), and slot identity replaces the j != i test, so every sum is
// formed in the reference's order.
//
// Launch shape: 256-lane blocks over sorted slots, block index remapped so each
// XCD owns one contiguous run of slots (its neighbours' rows then sit in the
// same L2).
#include <type_traits>

#include "pbf_common.hpp"
#include "pbf_hip.h"
#include "pbf_math.hpp"
#include "sorted_path.hpp"
#include "walk.hpp"

namespace pbfx {
namespace {

using namespace walk;

template <int BS>
__device__ __forceinline__ uint32_t slot_of_thread() {
    return xcd_tile(blockIdx.x, gridDim.x) * BS + threadIdx.x;
}

// A velocity whose components are all below 2^126 in magnitude (NaN fails the test):
// the difference of two such velocities is finite.  The kernels that write the step's
// velocities set kStepVelocityUnbounded for any other; XSPH reads it.
__device__ __forceinline__ bool velocity_bounded(float x, float y, float z) {
    return fabsf(x) < 0x1p126f && fabsf(y) < 0x1p126f && fabsf(z) < 0x1p126f;
}
__device__ __forceinline__ void note_velocity(uint32_t* stepFlags, float x, float y, float z) {
    if (stepFlags && !velocity_bounded(x, y, z)) atomicOr(stepFlags, kStepVelocityUnbounded);
}
// The checked shortcuts (the two-instruction s_corr ratio and the spiky coefficient's
// rsq start), when the record holds both checks' verdicts for this launch's constants
// (a record left from other parameters keeps the full sequences).
__device__ __forceinline__ void take_ratio_coef(PairConsts& c, const RatioCoef* __restrict__ rc) {
    if (rc && rc->ok != 0.f && rc->rsq_ok != 0.f && rc->w_q == c.w_q && rc->poly6 == c.poly6 &&
        rc->h2 == c.h2 && rc->neg_k == c.neg_scorr_k && rc->smin == c.scorr_min && rc->h == c.h &&
        rc->eps == c.grad_r_eps && rc->neg3_spiky == c.neg3_spiky) {
        c.rq_hi = rc->zh;
        c.rq_lo = rc->zl;
        c.ratio2 = true;
    }
}

// Every float W from +0 to the largest a pair can have (poly6 h2^3: r = 0), in a
// grid-stride loop: any W whose s_corr term from the two-instruction ratio differs from
// the division's sets bit 0 of *bad.  Every float r2 from +0 to h2 (and a NaN):
// any whose spiky coefficient from the rsq start differs from spiky_coef's (for NaN:
// is not NaN) sets bit 1.
__global__ __launch_bounds__(kBS) void k_ratio_check(sim::DeviceParams dp, uint32_t* __restrict__ bad) {
    const PairConsts c = make_pair_consts(dp);
    if (!pair_fixed(c)) return;
    const float zh = 1.0f / c.w_q;                               // RN(1 / w_q)
    const float zl = __builtin_fmaf(-c.w_q, zh, 1.0f) / c.w_q;  // 1 - w_q zh is exact
    const float wmax = c.poly6 * c.h2 * c.h2 * c.h2;
    if (!(wmax >= 0.f && wmax < 0x1p127f)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(bad, 3u);
        return;
    }
    const uint64_t top = __float_as_uint(wmax);
    const uint64_t first = (uint64_t)blockIdx.x * kBS + threadIdx.x, stride = (uint64_t)gridDim.x * kBS;
    uint32_t diff = 0;
    for (uint64_t u = first; u <= top; u += stride) {
        const float w = __uint_as_float((uint32_t)u);
        const float a = scorr_fixed(scorr_ratio2(w, zh, zl), c);
        const float b = scorr_fixed(div_unscaled(w, c.w_q, c.rcp_w_q), c);
        diff |= __float_as_uint(a) ^ __float_as_uint(b);
    }
    if (diff) atomicOr(bad, 1u);
    // h <= 2^16 in the common form (pair_unscaled_ok): h2 is a positive finite float
    const uint64_t topR = __float_as_uint(c.h2);
    uint32_t diffR = 0;
    for (uint64_t u = first; u <= topR; u += stride) {
        const float r2 = __uint_as_float((uint32_t)u);
        diffR |= __float_as_uint(spiky_coef_rsq(r2, c)) ^ __float_as_uint(spiky_coef(r2, c));
    }
    if (first == 0) {  // (r2 is never -0: a sum of squares from +0)
        const float nan = __uint_as_float(0x7FC00000u);
        diffR |= (__builtin_isnan(spiky_coef_rsq(nan, c)) && __builtin_isnan(spiky_coef(nan, c))) ? 0u : 1u;
    }
    if (diffR) atomicOr(bad, 2u);
}

__global__ void k_ratio_record(sim::DeviceParams dp, const uint32_t* __restrict__ bad, RatioCoef* __restrict__ rc) {
    const PairConsts c = make_pair_consts(dp);
    const float zh = 1.0f / c.w_q;
    RatioCoef r;
    r.zh = zh;
    r.zl = __builtin_fmaf(-c.w_q, zh, 1.0f) / c.w_q;
    r.ok = (pair_fixed(c) && (*bad & 1u) == 0u) ? 1.0f : 0.0f;
    r.w_q = c.w_q;
    r.poly6 = c.poly6;
    r.h2 = c.h2;
    r.neg_k = c.neg_scorr_k;
    r.smin = c.scorr_min;
    r.rsq_ok = (pair_fixed(c) && (*bad & 2u) == 0u) ? 1.0f : 0.0f;
    r.h = c.h;
    r.eps = c.grad_r_eps;
    r.neg3_spiky = c.neg3_spiky;
    *rc = r;
}

__global__ void k_clear_word(uint32_t* __restrict__ w) { *w = 0u; }

// Whether the step's flags say a solver pass wrote a NaN position (the candidate mask
// then no longer bounds the neighbours; see kStepNanPosition).
__device__ __forceinline__ bool nan_position_seen(const uint32_t* stepFlags) {
    return stepFlags != nullptr && (*stepFlags & kStepNanPosition) != 0u;
}

__global__ __launch_bounds__(kBS) void k_predict_hash(const float4* __restrict__ X, const float4* __restrict__ V,
                                                      float4* __restrict__ Xn, uint32_t* __restrict__ keys,
                                                      uint32_t* __restrict__ epoch, float3 g, float dt,
                                                      sim::GridBounds grid, uint32_t N, Scenes sc,
                                                      uint32_t* __restrict__ clampBits, uint32_t* __restrict__ stepFlags) {
    const uint32_t i = blockIdx.x * kBS + threadIdx.x;
    if (epoch && i == 0) epoch[0] = epoch[0] + 1u;
    if (stepFlags && i == 0) stepFlags[0] = 0u;  // the step's flags start clear
    if (i >= N) return;
    uint32_t keyBase = 0;
    if (sc.params) {
        const SceneParams& sp = sc.params[sc.sceneOf[i]];
        grid = sp.grid;
        g = sp.gravity;
        keyBase = sp.cellBase;
    }
    const float4 x = X[i];
    const float4 v = V[i];
    // reference: integrate.cu:43-51, then boundary.cu:11-25 with e = 0 (the
    // velocity side of that clamp is overwritten by the velocity phase)
    const float vx = v.x + g.x * dt;
    const float vy = v.y + g.y * dt;
    const float vz = v.z + g.z * dt;
    float px = x.x + vx * dt, py = x.y + vy * dt, pz = x.z + vz * dt;
    if (clampBits) clampBits[i] = clamp_bits(px, py, pz, grid);
    float ux = 0.f, uy = 0.f, uz = 0.f;
    clamp_box(px, py, pz, ux, uy, uz, grid, 0.0f);
    Xn[i] = make_float4(px, py, pz, 1.0f);
    keys[i] = keyBase + cell_key(px, py, pz, grid);
}

// Home slot of a cell key.  While the key space fits the table (scatter == 0) the home
// is the key itself, so consecutive keys (the cells of an x-window) share cache lines.
// Otherwise blocks of four consecutive keys (one 64-byte line of slots) are placed by a
// multiplicative (Fibonacci) hash of the block index.  Shifting whole key ranges by a
// constant instead lets the long runs of occupied cells of different ranges overlap, and
// linear probing then chains them: on the 8M splash's 4.1M cells in 2^24 slots that took
// 10.8 slots per lookup on average (145 at most), against 1.1 (10) with the blocks.
__device__ __forceinline__ uint32_t table_home(uint32_t key, uint32_t mask, uint32_t scatter) {
    if (!scatter) return key & mask;
    const uint32_t L = (uint32_t)__popc(mask);  // mask = 2^L - 1, 4 <= L <= 30
    const uint32_t blk = ((key >> 2) * 0x9E3779B1u) >> (34u - L);  // the product's top L - 2 bits
    return ((blk << 2) | (key & 3u)) & mask;
}

__global__ __launch_bounds__(kBS) void k_hash_insert(HashTable t, const uint32_t* __restrict__ uniq,
                                                     const uint32_t* __restrict__ off,
                                                     const uint32_t* __restrict__ count) {
    const uint32_t m = blockIdx.x * kBS + threadIdx.x;
    const uint32_t M = *count;
    if (m >= M) return;
    const uint32_t key = uniq[m];
    const uint32_t ep = *t.epoch;
    const unsigned long long mine = ((unsigned long long)ep << 32) | key;
    uint32_t slot = table_home(key, t.mask, t.scatter);
    for (uint32_t probe = 0; probe <= t.mask; ++probe, slot = (slot + 1) & t.mask) {
        unsigned long long* w = reinterpret_cast<unsigned long long*>(&t.slots[slot]);
        unsigned long long cur = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while ((uint32_t)(cur >> 32) != ep) {
            const unsigned long long prev = atomicCAS(w, cur, mine);
            if (prev == cur) {
                t.slots[slot].beg = off[m];
                t.slots[slot].end = off[m + 1];
                return;
            }
            cur = prev;
        }
        // slot taken this step by another cell: keep probing
    }
    *t.overflow = 1u;
}

struct HashRows {
    const HashSlot* __restrict__ slots;
    uint32_t mask;
    uint32_t scatter;
    uint32_t ep;
    __device__ __forceinline__ bool cell(uint32_t key, uint32_t& b, uint32_t& e) const {
        uint32_t slot = table_home(key, mask, scatter);
        for (uint32_t probe = 0; probe <= mask; ++probe, slot = (slot + 1) & mask) {
            const uint4 v = *reinterpret_cast<const uint4*>(&slots[slot]);
            if (v.y != ep) return false;
            if (v.x == key) { b = v.z; e = v.w; return b < e; }
        }
        return false;
    }
    __device__ __forceinline__ void span(uint32_t k0, uint32_t k1, uint32_t& lo, uint32_t& hi) const {
        lo = 0xFFFFFFFFu;
        hi = 0u;
        for (uint32_t c = k0; c <= k1; ++c) {
            uint32_t b, e;
            if (!cell(c, b, e)) continue;
            lo = min(lo, b);
            hi = max(hi, e);
        }
        if (lo == 0xFFFFFFFFu) lo = hi = 0;
    }
};

// Scene of a global cell key: the last scene whose key base is <= key.
__device__ __forceinline__ uint32_t scene_of_key(const Scenes& sc, uint32_t key) {
    uint32_t lo = 0, hi = sc.count;  // invariant: cellBase[lo] <= key < cellBase[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sc.cellBase[mid] <= key) lo = mid; else hi = mid;
    }
    return lo;
}

template <class Rows>
__device__ __forceinline__ void cell_spans(uint32_t key, uint32_t m, const Rows& rows, const SpanTable& st,
                                           sim::GridBounds g, const Scenes& sc) {
    uint32_t base0 = 0;
    if (sc.params) {
        const SceneParams& sp = sc.params[scene_of_key(sc, key)];
        g = sp.grid;
        base0 = sp.cellBase;
        key -= base0;
    }
    int cx, cy, cz;
    decode_key(key, g.dim, cx, cy, cz);
    const int R = st.reach;
    const int x0 = max(0, cx - R), x1 = min(g.dim.x - 1, cx + R);
    int r = 0;
    for (int dz = -R; dz <= R; ++dz) {
        const int z = cz + dz;
        for (int dy = -R; dy <= R; ++dy, ++r) {
            const int y = cy + dy;
            uint32_t lo = 0, hi = 0;
            if (z >= 0 && z < g.dim.z && y >= 0 && y < g.dim.y && x0 <= x1) {
                const uint32_t base = base0 + (uint32_t)((z * g.dim.y + y) * g.dim.x);
                rows.span(base + (uint32_t)x0, base + (uint32_t)x1, lo, hi);
            }
            st.spans[(uint64_t)r * st.stride + m] = make_uint2(lo, hi);
        }
    }
}

struct HashRowsDev;

// Row spans of a one-cell reach from the hash table, the lookups shared across a wave:
// each lane looks up only its own column's cell of every neighbour row and takes the
// cells on either side from the neighbouring lanes when those hold them (consecutive
// occupied cells of one x-line sit in consecutive lanes), looking them up itself
// otherwise.  The min / max over a window are exact, so the spans equal cell_spans'.
// Every lane of the wave takes part (has: the lane holds occupied cell m).
__device__ __forceinline__ void cell_spans_hashed_r1(uint32_t key, uint32_t m, bool has, const HashRows& hr,
                                                     const SpanTable& st, sim::GridBounds g, const Scenes& sc) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t base0 = 0;
    if (has && sc.params) {
        const SceneParams& sp = sc.params[scene_of_key(sc, key)];
        g = sp.grid;
        base0 = sp.cellBase;
        key -= base0;
    }
    int cx = 0, cy = 0, cz = 0;
    if (has) decode_key(key, g.dim, cx, cy, cz);
    int r = 0;
    for (int dz = -1; dz <= 1; ++dz) {
        for (int dy = -1; dy <= 1; ++dy, ++r) {
            const int z = cz + dz, y = cy + dy;
            const bool row = has && z >= 0 && z < g.dim.z && y >= 0 && y < g.dim.y;
            const uint32_t c = row ? base0 + (uint32_t)((z * g.dim.y + y) * g.dim.x + cx) : 0xFFFFFFFFu;
            uint32_t b = 0, e = 0;
            const bool occ = row && hr.cell(c, b, e);
            const uint32_t cL = (uint32_t)__shfl_up((int)c, 1), bL = (uint32_t)__shfl_up((int)b, 1),
                           eL = (uint32_t)__shfl_up((int)e, 1);
            const int occL = __shfl_up((int)occ, 1);
            const uint32_t cR = (uint32_t)__shfl_down((int)c, 1), bR = (uint32_t)__shfl_down((int)b, 1),
                           eR = (uint32_t)__shfl_down((int)e, 1);
            const int occR = __shfl_down((int)occ, 1);
            uint32_t lo = 0xFFFFFFFFu, hi = 0u;
            auto take = [&](bool o, uint32_t bb, uint32_t ee) {
                if (o) {
                    lo = min(lo, bb);
                    hi = max(hi, ee);
                }
            };
            if (row) {
                if (cx > 0) {
                    if (lane > 0u && cL == c - 1u) {
                        take(occL != 0, bL, eL);
                    } else {
                        uint32_t bb = 0, ee = 0;
                        const bool o = hr.cell(c - 1u, bb, ee);
                        take(o, bb, ee);
                    }
                }
                take(occ, b, e);
                if (cx + 1 < g.dim.x) {
                    if (lane < 63u && cR == c + 1u) {
                        take(occR != 0, bR, eR);
                    } else {
                        uint32_t bb = 0, ee = 0;
                        const bool o = hr.cell(c + 1u, bb, ee);
                        take(o, bb, ee);
                    }
                }
            }
            if (lo == 0xFFFFFFFFu) lo = hi = 0u;
            if (has) st.spans[(uint64_t)r * st.stride + m] = make_uint2(lo, hi);
        }
    }
}

template <class Rows>
__global__ __launch_bounds__(kBS) void k_spans_gather(const float4* __restrict__ Xn, const uint32_t* __restrict__ idxS,
                                                      float4* __restrict__ Ps, uint32_t N,
                                                      const uint32_t* __restrict__ uniq,
                                                      const uint32_t* __restrict__ count, Rows rows, SpanTable st,
                                                      sim::GridBounds g, Scenes sc) {
    const uint32_t t = blockIdx.x * kBS + threadIdx.x;
    if (t < N) {
        const uint32_t i = idxS[t];
        const float4 p = Xn[i];
        Ps[t] = make_float4(p.x, p.y, p.z, 0.0f);
        if (sc.sceneOfSlot) sc.sceneOfSlot[t] = sc.sceneOf[i];
    }
    const uint32_t M = *count;
    if constexpr (std::is_same<Rows, HashRowsDev>::value) {
        if (st.reach == 1) {
            const bool has = t < M;
            cell_spans_hashed_r1(has ? uniq[t] : 0u, t, has, HashRows{rows.slots, rows.mask, rows.scatter, *rows.epoch},
                                 st, g, sc);
            return;
        }
    }
    if (t < M) cell_spans(uniq[t], t, rows, st, g, sc);
}

struct CompactRowsCount {
    const uint32_t* __restrict__ uniq;
    const uint32_t* __restrict__ off;
    const uint32_t* __restrict__ count;
    __device__ __forceinline__ void span(uint32_t k0, uint32_t k1, uint32_t& lo, uint32_t& hi) const {
        CompactRows r{uniq, off, *count};
        r.span(k0, k1, lo, hi);
    }
};

struct HashRowsDev {
    const HashSlot* __restrict__ slots;
    uint32_t mask;
    uint32_t scatter;
    const uint32_t* __restrict__ epoch;
    __device__ __forceinline__ void span(uint32_t k0, uint32_t k1, uint32_t& lo, uint32_t& hi) const {
        HashRows r{slots, mask, scatter, *epoch};
        r.span(k0, k1, lo, hi);
    }
};

// Lane order of the neighbour passes.  Pair pops serve every lane of a wave that holds
// a queued pair, and a particle's neighbours arrive in bursts that follow its place
// inside its cell (one near a cell's top meets most of them in the rows above).  Each
// group of kLaneGroup slots is therefore dealt to its waves by that place (z, then y,
// in quarters of a cell), stable in slot order within a class, so the lanes of a wave
// burst together and a pop finds more of them busy.  perm[t] is the slot lane t of
// the neighbour passes takes; every slot of a group stays in the group.  Only the
// assignment of particles to lanes changes, never a particle's sums.  Groups of 128
// slots (two waves) measured best on the collapsed 1M dam break: 256 let the waves'
// candidate loads spread over more cells and slowed the lambda passes.
constexpr uint32_t kLaneGroup = 128;
template <uint32_t G>
__global__ __launch_bounds__(G) void k_lane_order(const float4* __restrict__ Ps, uint32_t N, sim::GridBounds g,
                                                  Scenes sc, uint32_t* __restrict__ perm) {
    constexpr uint32_t W = G / 64u;
    __shared__ uint32_t counts[W][16];
    const uint32_t t = blockIdx.x * G + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t cls = 16u;  // lanes past N sort after every class and are not placed
    if (t < N) {
        const float4 p = Ps[t];
        const sim::GridBounds& gb = sc.params ? sc.params[sc.sceneOfSlot[t]].grid : g;
        const float inv = 1.0f / gb.cellSize;
        const float fz = (p.z - gb.mins.z) * inv, fy = (p.y - gb.mins.y) * inv;
        const uint32_t qz = (uint32_t)((fz - floorf(fz)) * 4.0f) & 3u;
        const uint32_t qy = (uint32_t)((fy - floorf(fy)) * 4.0f) & 3u;
        cls = qz * 4u + qy;
    }
    uint32_t rank = 0;
    for (uint32_t c = 0; c < 16u; ++c) {
        const uint64_t b = __ballot(cls == c);
        if (cls == c) rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
        if (lane == 0) counts[w][c] = (uint32_t)__popcll(b);
    }
    __syncthreads();
    if (t >= N) return;
    uint32_t off = 0;
    for (uint32_t c = 0; c < cls; ++c)
        for (uint32_t v = 0; v < W; ++v) off += counts[v][c];
    for (uint32_t v = 0; v < w; ++v) off += counts[v][cls];
    perm[blockIdx.x * G + off + rank] = t;
}

// The slot a lane of the neighbour passes takes: its own, or perm's.
__device__ __forceinline__ uint32_t lane_slot(uint32_t t, uint32_t N, const uint32_t* __restrict__ perm) {
    return (perm && t < N) ? perm[t] : t;
}

// The candidate mask of every sorted slot, from the positions Ps.  Each lane takes a slot
// pair (two slots of one cell, which share the cell's rows, from build_slot_pairs): it
// walks the rows kMaskRound candidates a round, each load serving both slots (reads past
// the last particle's padding return zeros, and past the span are masked off), and
// records each slot's skin bit of every candidate in walk order, its own slot included
// (the walks that read the mask skip it), 32 to a word.  No pair arithmetic and no
// queue: the candidate test alone.  Built here rather than during the first delta pass
// (as it once was), it also serves the first lambda pass, which then skips the
// candidates outside the skin too.
// Eight loads a round kept more of them in flight than four (a 9% faster build), with no
// pair arithmetic to hide their latency behind.
constexpr uint32_t kMaskRound = 8;
template <int BS>
__global__ __launch_bounds__(BS) void k_mask_build(const float4* __restrict__ Ps, const uint32_t* __restrict__ segOf,
                                                   SpanTable st, CandMask mk, uint32_t N,
                                                   const uint2* __restrict__ pairs,
                                                   const uint32_t* __restrict__ pairCount) {
    const uint32_t np = *pairCount;
    const uint32_t nb = (np + (uint32_t)BS - 1u) / (uint32_t)BS;
    if (blockIdx.x >= nb) return;
    const uint32_t p = xcd_tile(blockIdx.x, nb) * BS + threadIdx.x;
    if (p >= np) return;
    const uint2 ab = pairs[p];
    const uint32_t a = ab.x;
    const bool hasB = ab.y != 0xFFFFFFFFu;
    const uint32_t b = hasB ? ab.y : a;
    const float4 pa = Ps[a], pb = Ps[b];
    const uint32_t m = segOf[a];
    const SlotBuf PS(Ps, N + kSlotPad);
    uint32_t cidx = 0, wa = 0, wb = 0;
    auto put = [&](uint32_t slot, uint32_t wi, uint32_t v) {
        if (wi < kMaskWords) mk.words[(uint64_t)wi * mk.stride + slot] = v;
    };
    for (int r = 0; r < st.rows; ++r) {
        const uint2 sp = st.spans[(uint64_t)r * st.stride + m];
        for (uint32_t k = sp.x; k < sp.y; k += kMaskRound) {
            const uint32_t cnt = min(kMaskRound, sp.y - k);
            float4 pj[kMaskRound];
#pragma unroll
            for (uint32_t u = 0; u < kMaskRound; ++u) pj[u] = PS.at(k, u);
            uint32_t sa = 0u, sb = 0u;
#pragma unroll
            for (uint32_t u = 0; u < kMaskRound; ++u) {
                sa |= in_range(pair_r2(pa.x - pj[u].x, pa.y - pj[u].y, pa.z - pj[u].z), mk.rs2) ? 1u << u : 0u;
                sb |= in_range(pair_r2(pb.x - pj[u].x, pb.y - pj[u].y, pb.z - pj[u].z), mk.rs2) ? 1u << u : 0u;
            }
            const uint32_t inSpan = (1u << cnt) - 1u;
            const uint32_t sh = cidx & 31u;
            uint64_t xa = (uint64_t)wa | ((uint64_t)(sa & inSpan) << sh);
            uint64_t xb = (uint64_t)wb | ((uint64_t)(sb & inSpan) << sh);
            if (sh + cnt >= 32u) {
                const uint32_t wi = cidx >> 5;
                put(a, wi, (uint32_t)xa);
                if (hasB) put(b, wi, (uint32_t)xb);
                xa >>= 32;
                xb >>= 32;
            }
            wa = (uint32_t)xa;
            wb = (uint32_t)xb;
            cidx += cnt;
        }
    }
    if ((cidx & 31u) != 0u) {
        put(a, cidx >> 5, wa);
        if (hasB) put(b, cidx >> 5, wb);
    }
}

// Lambda of a sorted slot: the full walk or the mask walk.
template <int BS>
__global__ __launch_bounds__(BS) void k_lambda_sorted(float4* Ps, const uint32_t* __restrict__ segOf, SpanTable st,
                                                       CandMask mk, int useMask, sim::DeviceParams dp, uint32_t N,
                                                       float4* __restrict__ recA, uint2* __restrict__ recB,
                                                       const uint32_t* __restrict__ perm,
                                                       const uint32_t* __restrict__ stepFlags) {
    __shared__ uint32_t sQ[kQD * BS];
    __shared__ float4 sBuf[BS];
    const uint32_t s = lane_slot(slot_of_thread<BS>(), N, perm);
    const bool valid = s < N;
    const uint32_t ss = valid ? s : N - 1;
    const float4 pi = Ps[ss];
    const uint32_t m = segOf[ss];
    const PairConsts c = make_pair_consts(dp);
    const SlotBuf PS(Ps, N + kSlotPad);
    LambdaAcc a;
    uint32_t cut = 0xFFFFFFFFu;  // recording: slot of the neighbour that reached the cap
    const int cap = dp.maxNeighbors;
    auto dist2 = [&](const float4& pj) {
        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        return pair_r2(rx, ry, rz);
    };
    with_pair_form(c, [&](auto form) {
        auto apply = [&](uint32_t k) {
            const float4 pj = PS[k];
            const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
            lambda_pair<decltype(form)::value>(a, rx, ry, rz, pair_r2(rx, ry, rz), c);
            if (recB && a.n == cap) cut = k;
        };
        if (useMask && !nan_position_seen(stepFlags))
            mask_walk<BS>(valid, s, m, st, mk, PS, sQ + threadIdx.x, dp.maxNeighbors, c.h2, dist2, apply);
        else
            neighbor_walk<BS>(valid, s, SpanRows(st, m), SortedFetch{PS}, sBuf + (threadIdx.x & ~63u),
                                     sQ + threadIdx.x, dp.maxNeighbors, c.h2, dist2, apply);
    });
    if (!valid) return;
    reinterpret_cast<float*>(Ps)[4ull * s + 3] = lambda_finish(a, dp);
    if (recA) {
        float b0, e0;
        lambda_adjoint_coefs(a, dp, b0, e0);
        recA[s] = make_float4(a.gx, a.gy, a.gz, b0);
        recB[s] = make_uint2(__float_as_uint(e0), cut);
    }
}

// Delta of a sorted slot: every candidate (a delta pass takes every neighbour; for it
// the candidate mask's walk measured slower than the full one).
template <int BS>
__global__ __launch_bounds__(BS) void k_delta_sorted(const float4* __restrict__ Pa, float4* __restrict__ Pb,
                                                      const uint32_t* __restrict__ segOf, SpanTable st,
                                                      sim::DeviceParams dp, uint32_t N,
                                                      int last, const float4* __restrict__ X,
                                                      const uint32_t* __restrict__ idxS, float4* __restrict__ Vs,
                                                      float invDt, Scenes sc, uint32_t* __restrict__ clampBits,
                                                      float4* __restrict__ recU, const uint32_t* __restrict__ perm,
                                                      uint32_t* __restrict__ stepFlags,
                                                      const RatioCoef* __restrict__ ratioCoef) {
    __shared__ uint32_t sQ[kQD * BS];
    __shared__ float4 sBuf[BS];
    const uint32_t s = lane_slot(slot_of_thread<BS>(), N, perm);
    const bool valid = s < N;
    const uint32_t ss = valid ? s : N - 1;
    const float4 pi = Pa[ss];
    const uint32_t m = segOf[ss];
    PairConsts c = make_pair_consts(dp);
    take_ratio_coef(c, ratioCoef);
    const SlotBuf PA(Pa, N + kSlotPad);
    float dx = 0.f, dy = 0.f, dz = 0.f;
    auto dist2 = [&](const float4& pj) {
        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        return pair_r2(rx, ry, rz);
    };
    float4* buf = sBuf + (threadIdx.x & ~63u);
    with_delta_form(c, [&](auto form) {
        auto apply = [&](uint32_t k) {
            const float4 pj = PA[k];
            const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
            delta_pair<decltype(form)::kFixed, decltype(form)::kRatio2>(dx, dy, dz, rx, ry, rz,
                                                                        pair_r2(rx, ry, rz), pi.w, pj.w, c);
        };
        neighbor_walk<BS>(valid, s, SpanRows(st, m), SortedFetch{PA}, buf, sQ + threadIdx.x, 0, c.h2, dist2, apply);
    });
    if (!valid) return;
    const float3 e = delta_scaled(dx, dy, dz, dp);
    if (recU) recU[s] = make_float4(e.x, e.y, e.z, 0.f);
    const float3 d = delta_clamp(e, dp);
    // reference: pbf_delta_apply.cu:232-233 (P += delta)
    float px = pi.x + d.x, py = pi.y + d.y, pz = pi.z + d.z;
    if (!last) {
        Pb[s] = make_float4(px, py, pz, 0.0f);
        if (stepFlags && (__builtin_isnan(px) || __builtin_isnan(py) || __builtin_isnan(pz)))
            atomicOr(stepFlags, kStepNanPosition);
        return;
    }
    // post-solve boundary (phases.cpp:94) and velocity (phases.cpp:102-105)
    float ux = 0.f, uy = 0.f, uz = 0.f;
    const sim::GridBounds& gb = sc.params ? sc.params[sc.sceneOfSlot[s]].grid : dp.grid;
    if (clampBits) clampBits[s] = clamp_bits(px, py, pz, gb);
    clamp_box(px, py, pz, ux, uy, uz, gb, 0.0f);
    Pb[s] = make_float4(px, py, pz, 1.0f);
    const float4 x = X[idxS[s]];
    const float vx = (px - x.x) * invDt, vy = (py - x.y) * invDt, vz = (pz - x.z) * invDt;
    Vs[s] = make_float4(vx, vy, vz, 0.0f);
    note_velocity(stepFlags, vx, vy, vz);
}

__global__ __launch_bounds__(kBS) void k_finalize_sorted(float4* __restrict__ Ps, const float4* __restrict__ X,
                                                         const uint32_t* __restrict__ idxS, float4* __restrict__ Vs,
                                                         sim::GridBounds g, float invDt, uint32_t N, Scenes sc,
                                                         uint32_t* __restrict__ stepFlags) {
    const uint32_t s = blockIdx.x * kBS + threadIdx.x;
    if (s >= N) return;
    if (sc.params) g = sc.params[sc.sceneOfSlot[s]].grid;
    const float4 p = Ps[s];
    float px = p.x, py = p.y, pz = p.z, ux = 0.f, uy = 0.f, uz = 0.f;
    clamp_box(px, py, pz, ux, uy, uz, g, 0.0f);
    Ps[s] = make_float4(px, py, pz, 1.0f);
    const float4 x = X[idxS[s]];
    const float vx = (px - x.x) * invDt, vy = (py - x.y) * invDt, vz = (pz - x.z) * invDt;
    Vs[s] = make_float4(vx, vy, vz, 0.0f);
    note_velocity(stepFlags, vx, vy, vz);
}

// XSPH's pair term is a few products, so its walk evaluates every candidate in place
// rather than queueing the neighbours for a later pass: the sums take the same
// neighbours in the same order, with no queue to fill and drain.  With one particle a
// lane the vector memory pipeline, not the VALU, bounded that walk, so each lane takes
// two particles of one cell (a slot pair: the cell's slots two at a time), which share
// the cell's neighbour rows: a round's four candidate positions are read once for both,
// and a candidate's velocity once when it is a neighbour of either (a non-neighbour's
// read is sent past the end of the buffer, where the hardware returns zeros without
// touching memory).
// Both arrays carry kSlotPad padding slots past the last particle.  The grid covers N
// pairs, the most there can be; blocks past the step's pair count return at once, and
// the XCD-aware order is taken over the blocks that have pairs.
//
// A candidate that is not a neighbour must leave the sums as they are.  In general each
// sum selects its old value (three selects a particle per candidate).  When the step's
// flags have kStepVelocityUnbounded clear (every velocity of the step has components below
// 2^126, so every difference v_j - v_i is finite), the weight is selected instead, W or +0,
// and the fused update runs on every candidate: fma(d, +0, s) == s for finite d, since the
// sums start at +0 and a round-to-nearest sum is -0 only when both of its terms are.
template <int BS>
__global__ __launch_bounds__(BS) void k_xsph_scatter(const float4* __restrict__ Ps, const float4* __restrict__ Vs,
                                                     const uint32_t* __restrict__ segOf, SpanTable st,
                                                     sim::DeviceParams dp, uint32_t N,
                                                     const uint2* __restrict__ pairs,
                                                     const uint32_t* __restrict__ pairCount,
                                                     const uint32_t* __restrict__ idxS, float4* __restrict__ Xn,
                                                     float4* __restrict__ V, Scenes sc, float* __restrict__ recGm,
                                                     const uint32_t* __restrict__ stepFlags) {
    const uint32_t np = *pairCount;
    const uint32_t nb = (np + (uint32_t)BS - 1u) / (uint32_t)BS;
    if (blockIdx.x >= nb) return;
    const uint32_t p = xcd_tile(blockIdx.x, nb) * BS + threadIdx.x;
    if (p >= np) return;
    const uint2 ab = pairs[p];
    const uint32_t a = ab.x;
    const bool hasB = ab.y != 0xFFFFFFFFu;
    const uint32_t b = hasB ? ab.y : a;
    const float4 pa = Ps[a], pb = Ps[b];
    const float4 va = Vs[a], vb = Vs[b];
    float vax = va.x, vay = va.y, vaz = va.z, vbx = vb.x, vby = vb.y, vbz = vb.z;
    float xc = dp.xsph_c, rest = dp.boundaryRestitution;
    sim::GridBounds grid = dp.grid;
    if (sc.params) {
        // a cell never spans two scenes
        const SceneParams& sp = sc.params[sc.sceneOfSlot[a]];
        xc = sp.xsph_c;
        rest = sp.restitution;
        grid = sp.grid;
    }
    if (xc > 0.f || recGm != nullptr) {
        const float h2 = dp.kernel.h2, poly6 = dp.kernel.poly6;
        const SlotBuf PS(Ps, N + kSlotPad), VS(Vs, N + kSlotPad);
        float sax = 0.f, say = 0.f, saz = 0.f, sbx = 0.f, sby = 0.f, sbz = 0.f;
        int na = 0, nb2 = 0;
        // one neighbour's term into the sums (sx, sy, sz), by the selected weight or
        // by selecting the old sums
        auto take = [&](auto weighted, float& sx, float& sy, float& sz, int& n, const float4& vj, const float4& vi,
                        float r2, bool in) {
            const float hr2 = h2 - r2;
            const float w = poly6 * hr2 * hr2 * hr2;
            if constexpr (decltype(weighted)::value) {
                xsph_pair(sx, sy, sz, vj, vi, in ? w : 0.0f);
            } else {
                float tx = sx, ty = sy, tz = sz;
                xsph_pair(tx, ty, tz, vj, vi, w);
                sx = in ? tx : sx;
                sy = in ? ty : sy;
                sz = in ? tz : sz;
            }
            n += in ? 1 : 0;
        };
        auto walk = [&](auto weighted) {
            inline_walk_pair(
                a, b, hasB, SpanRows(st, segOf[a]), h2, [&](uint32_t k, uint32_t u) { return PS.at(k, u); },
                [&](const float4& pj) { return pair_r2(pa.x - pj.x, pa.y - pj.y, pa.z - pj.z); },
                [&](const float4& pj) { return pair_r2(pb.x - pj.x, pb.y - pj.y, pb.z - pj.z); },
                [&](uint32_t k, uint32_t u, bool in) { return VS.at(in ? k : kSlotBufBeyond, u); },
                [&](const float4& vj, float r2, bool in) { take(weighted, sax, say, saz, na, vj, va, r2, in); },
                [&](const float4& vj, float r2, bool in) { take(weighted, sbx, sby, sbz, nb2, vj, vb, r2, in); });
        };
        if (stepFlags != nullptr && (*stepFlags & kStepVelocityUnbounded) == 0u)
            walk(std::true_type{});
        else
            walk(std::false_type{});
        if (recGm) {
            const float mor = (dp.restDensity > 0.f) ? (dp.particleMass / dp.restDensity) : 0.f;
            recGm[a] = xsph_gate(na, dp) * mor;
            if (hasB) recGm[b] = xsph_gate(nb2, dp) * mor;
        }
        if (xc > 0.f) {
            const float3 u = xsph_finish(va.x, va.y, va.z, sax, say, saz, na, dp, xc);
            vax = u.x; vay = u.y; vaz = u.z;
            const float3 w = xsph_finish(vb.x, vb.y, vb.z, sbx, sby, sbz, nb2, dp, xc);
            vbx = w.x; vby = w.y; vbz = w.z;
        }
    }
    // BoundaryOp with restitution (post_ops.cpp:25-29, boundary.cu:11-25)
    const float e = clamp_restitution(rest);
    float px = pa.x, py = pa.y, pz = pa.z;
    clamp_box(px, py, pz, vax, vay, vaz, grid, e);
    const uint32_t ia = idxS[a];
    Xn[ia] = make_float4(px, py, pz, 1.0f);
    V[ia] = make_float4(vax, vay, vaz, 0.0f);
    if (hasB) {
        px = pb.x; py = pb.y; pz = pb.z;
        clamp_box(px, py, pz, vbx, vby, vbz, grid, e);
        const uint32_t ib = idxS[b];
        Xn[ib] = make_float4(px, py, pz, 1.0f);
        V[ib] = make_float4(vbx, vby, vbz, 0.0f);
    }
}

inline uint32_t blocks_for(uint32_t n) { return ceil_div(n > 0 ? n : 1, kBS); }

}  // namespace

float mask_skin_radius2(const sim::SimParams& p) {
    if (!p.pbf.enable_disp_clamp || p.solverIters < 1 || !(p.pbf.disp_clamp_max_h >= 0.f)) return 0.f;
    const double h = p.kernel.h;
    // the mask serves the lambda passes of iterations 0 .. I-1, whose positions are at
    // most I-1 clamped corrections away from the ones it was built on (the delta passes
    // and XSPH walk without it)
    const double rs = h * (1.0 + 2.0 * (p.solverIters - 1) * (double)p.pbf.disp_clamp_max_h);
    if (rs > 2.0 * h) return 0.f;
    const double r = rs * (1.0 + 1e-3);  // margin for float rounding of positions and r^2
    return (float)(r * r);
}

void launch_predict_hash(const float4* X, const float4* V, float4* Xn, uint32_t* keys, uint32_t* epoch,
                         const sim::SimParams& p, const Scenes& sc, hipStream_t s, uint32_t* clampBits,
                         uint32_t* stepFlags) {
    const uint32_t N = p.numParticles;
    k_predict_hash<<<blocks_for(N), kBS, 0, s>>>(X, V, Xn, keys, epoch, p.gravity, p.dt, p.grid, N, sc,
                                                 clampBits, stepFlags);
    check_launch("predict_hash");
}

void launch_hash_insert(const HashTable& t, const uint32_t* uniq, const uint32_t* off, const uint32_t* count,
                        uint32_t maxCells, hipStream_t s) {
    k_hash_insert<<<blocks_for(maxCells), kBS, 0, s>>>(t, uniq, off, count);
    check_launch("hash_insert");
}

void launch_spans_gather(CellLookup lookup, const float4* Xn, const uint32_t* idxS, float4* Ps, uint32_t N,
                         const uint32_t* uniq, const uint32_t* off, const uint32_t* count, const uint32_t* cellStart,
                         const uint32_t* cellEnd, const HashTable& table, const SpanTable& st,
                         const sim::GridBounds& grid, const Scenes& sc, hipStream_t s) {
    if (N == 0) return;
    const uint32_t b = blocks_for(N);
    switch (lookup) {
        case CellLookup::Dense:
            k_spans_gather<<<b, kBS, 0, s>>>(Xn, idxS, Ps, N, uniq, count, DenseRows{cellStart, cellEnd}, st, grid, sc);
            break;
        case CellLookup::Compact:
            k_spans_gather<<<b, kBS, 0, s>>>(Xn, idxS, Ps, N, uniq, count, CompactRowsCount{uniq, off, count}, st,
                                             grid, sc);
            break;
        case CellLookup::Hashed:
            k_spans_gather<<<b, kBS, 0, s>>>(Xn, idxS, Ps, N, uniq, count,
                                             HashRowsDev{table.slots, table.mask, table.scatter, table.epoch}, st, grid, sc);
            break;
    }
    check_launch("spans_gather");
}

void launch_mask_build(const float4* Ps, const uint32_t* segOf, const SpanTable& st, const CandMask& mk, uint32_t N,
                       const uint2* pairs, const uint32_t* pairCount, hipStream_t s) {
    if (N == 0) return;
    k_mask_build<128><<<ceil_div(N, 128u), 128, 0, s>>>(Ps, segOf, st, mk, N, pairs, pairCount);
    check_launch("mask_build");
}

void launch_lambda_sorted(float4* Ps, const uint32_t* segOf, const SpanTable& st, const CandMask& mk,
                          bool useMask, const sim::DeviceParams& dp, uint32_t N, hipStream_t s, float4* recA,
                          uint2* recB, const uint32_t* perm, const uint32_t* stepFlags) {
    if (N == 0) return;
    if ((recA == nullptr) != (recB == nullptr)) {
        set_error(PBFX_E_INVALID_ARGUMENT, "launch_lambda_sorted: recA and recB go together");
        return;
    }
    const int um = useMask ? 1 : 0;
    if (st.block == 128)
        k_lambda_sorted<128><<<ceil_div(N, 128u), 128, 0, s>>>(Ps, segOf, st, mk, um, dp, N, recA, recB, perm,
                                                               stepFlags);
    else
        k_lambda_sorted<256><<<ceil_div(N, 256u), 256, 0, s>>>(Ps, segOf, st, mk, um, dp, N, recA, recB, perm,
                                                               stepFlags);
    check_launch("lambda_sorted");
}

void launch_delta_sorted(const float4* Pa, float4* Pb, const uint32_t* segOf, const SpanTable& st,
                         const sim::DeviceParams& dp, uint32_t N, bool last,
                         const float4* X, const uint32_t* idxS, float4* Vs, float invDt, const Scenes& sc,
                         hipStream_t s, uint32_t* clampBits, float4* recU, const uint32_t* perm,
                         uint32_t* stepFlags, const RatioCoef* ratioCoef) {
    if (N == 0) return;
    const int lastPass = last ? 1 : 0;
    if (st.block == 128)
        k_delta_sorted<128><<<ceil_div(N, 128u), 128, 0, s>>>(Pa, Pb, segOf, st, dp, N, lastPass, X,
                                                              idxS, Vs, invDt, sc, clampBits, recU, perm, stepFlags,
                                                              ratioCoef);
    else
        k_delta_sorted<256><<<ceil_div(N, 256u), 256, 0, s>>>(Pa, Pb, segOf, st, dp, N, lastPass, X,
                                                              idxS, Vs, invDt, sc, clampBits, recU, perm, stepFlags,
                                                              ratioCoef);
    check_launch("delta_sorted");
}

void launch_ratio_check(const sim::DeviceParams& dp, RatioCoef* coef, uint32_t* bad, hipStream_t s) {
    k_clear_word<<<1, 1, 0, s>>>(bad);
    k_ratio_check<<<8192, kBS, 0, s>>>(dp, bad);
    k_ratio_record<<<1, 1, 0, s>>>(dp, bad, coef);
    check_launch("ratio_check");
}

void launch_lane_order(const float4* Ps, uint32_t N, const sim::GridBounds& grid, const Scenes& sc, uint32_t* perm,
                       hipStream_t s) {
    if (N == 0) return;
    k_lane_order<kLaneGroup><<<ceil_div(N, kLaneGroup), kLaneGroup, 0, s>>>(Ps, N, grid, sc, perm);
    check_launch("lane_order");
}

void launch_finalize_sorted(float4* Ps, const float4* X, const uint32_t* idxS, float4* Vs,
                            const sim::GridBounds& grid, float invDt, uint32_t N, const Scenes& sc,
                            hipStream_t s, uint32_t* stepFlags) {
    if (N == 0) return;
    k_finalize_sorted<<<blocks_for(N), kBS, 0, s>>>(Ps, X, idxS, Vs, grid, invDt, N, sc, stepFlags);
    check_launch("finalize_sorted");
}

void launch_xsph_scatter(const float4* Ps, const float4* Vs, const uint32_t* segOf, const SpanTable& st,
                         const sim::DeviceParams& dp, uint32_t N, const uint2* pairs, const uint32_t* pairCount,
                         const uint32_t* idxS, float4* Xn, float4* V, const Scenes& sc, hipStream_t s,
                         float* recGm, const uint32_t* stepFlags) {
    if (N == 0) return;
    if (st.block == 128)
        k_xsph_scatter<128><<<ceil_div(N, 128u), 128, 0, s>>>(Ps, Vs, segOf, st, dp, N, pairs, pairCount, idxS, Xn,
                                                              V, sc, recGm, stepFlags);
    else
        k_xsph_scatter<256><<<ceil_div(N, 256u), 256, 0, s>>>(Ps, Vs, segOf, st, dp, N, pairs, pairCount, idxS, Xn,
                                                              V, sc, recGm, stepFlags);
    check_launch("xsph_scatter");
}

}  // namespace pbfx
