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
#include "pbf_common.hpp"
#include "pbf_hip.h"
#include "pbf_math.hpp"
#include "sorted_path.hpp"

namespace pbfx {
namespace {

constexpr int kBS = 256;

__device__ __forceinline__ uint32_t slot_of_thread() {
    return xcd_tile(blockIdx.x, gridDim.x) * kBS + threadIdx.x;
}

__global__ __launch_bounds__(kBS) void k_predict_hash(const float4* __restrict__ X, const float4* __restrict__ V,
                                                      float4* __restrict__ Xn, uint32_t* __restrict__ keys,
                                                      uint32_t* __restrict__ epoch, float3 g, float dt,
                                                      sim::GridBounds grid, uint32_t N, Scenes sc,
                                                      uint32_t* __restrict__ clampBits) {
    const uint32_t i = blockIdx.x * kBS + threadIdx.x;
    if (epoch && i == 0) epoch[0] = epoch[0] + 1u;
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

// Home slot of a cell key.  While the key space fits the table the home is the key
// itself, so consecutive keys (the cells of an x-window) share cache lines.  Keys past
// the table size are shifted by a multiplicative hash of their high bits, so key blocks
// that wrap around the table do not all land on the same slots.
__device__ __forceinline__ uint32_t table_home(uint32_t key, uint32_t mask) {
    const int L = __popc(mask);  // mask = 2^L - 1
    const uint32_t hi = L >= 32 ? 0u : (key >> L);
    return (key + hi * 0x9E3779B1u) & mask;
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
    uint32_t slot = table_home(key, t.mask);
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
    uint32_t ep;
    __device__ __forceinline__ bool cell(uint32_t key, uint32_t& b, uint32_t& e) const {
        uint32_t slot = table_home(key, mask);
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
    const uint32_t* __restrict__ epoch;
    __device__ __forceinline__ void span(uint32_t k0, uint32_t k1, uint32_t& lo, uint32_t& hi) const {
        HashRows r{slots, mask, *epoch};
        r.span(k0, k1, lo, hi);
    }
};

// ------------------------------------------------------------------------
// Neighbour walk of one wave (64 consecutive sorted slots, one particle per lane).
//
// Two walk modes per neighbour row, chosen wave-uniformly.  Stream: when the union
// of the lanes' slot spans is at most SpanTable::streamRatio times the longest
// lane span, the wave streams the union through a per-wave LDS buffer, 64
// candidates at a time loaded coalesced, and every lane tests each candidate (an
// LDS broadcast) against its own span, itself and the distance h; with chunkBits
// it tests a whole chunk into a bit mask first.  Per lane: otherwise each lane
// walks its own span from global memory, kUnroll candidates a round so their
// loads are in flight together.  When a wave spans several cells the union carries
// candidates most lanes do not need, and the per-lane walk is faster.
//
// Deferred pairs.  In-range slots go to a per-lane FIFO in LDS; the expensive
// per-pair work (sqrt, IEEE divisions, the s_corr power) runs when some lane's
// FIFO is full, for one queued pair on every lane that has one, and at the end.
// Otherwise a wave would execute the per-pair path whenever any of its 64 lanes
// finds a neighbour, at low utilisation.
//
// Order.  Each lane still sees its candidates in reference order (rows dz-major
// then dy, slots ascending within a row) and consumes its FIFO in order, so
// every sum is formed in the reference's order.
constexpr int kQD = 16;          // FIFO depth per lane
constexpr uint32_t kUnroll = 4;  // candidates per round of a lane's own walk
static_assert(kSlotPad + 1 >= kUnroll, "a round may load kUnroll - 1 slots past the last particle");

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o));
    return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
    return v;
}

// Candidate mask: bit c of particle s is set when candidate c of its walk (the
// concatenated neighbour-row spans, in walk order) was within the skin radius
// h + 2 * I * maxDisp when the first delta pass scanned the predicted positions.
// Every later pass of the step moves a particle by at most I * maxDisp (each
// delta is clamped to maxDisp; the box clamp is a projection and cannot increase
// distances), so a pair outside the skin cannot come within h later in the step.
// Later passes visit only the set bits, in walk order, and still test r2 <= h2.
// The mask covers the first 32 * kMaskWords candidates; candidates past that are
// tested directly.
constexpr uint32_t kMaskWords = kCandMaskWords;

// Bits [0, n) of a 64-bit word (n <= 64).
__device__ __forceinline__ uint64_t low_bits64(uint32_t n) { return n >= 64u ? ~0ull : ((1ull << n) - 1ull); }

template <bool kBuild, class Dist2, class Apply>
__device__ __forceinline__ void neighbor_walk(bool valid, uint32_t s, uint32_t m, const SpanTable& st,
                                              const CandMask& mk, const float4* P, float4* buf, uint32_t* q,
                                              int cap, float h2, Dist2&& dist2, Apply&& apply) {
    const uint32_t lane = threadIdx.x & 63;
    int qh = 0, qn = 0, found = 0;
    uint32_t cidx = 0, word = 0;  // mask building: own-candidate index and pending word
    auto push = [&](uint32_t k) {
        q[((qh + qn) & (kQD - 1)) * kBS] = k;
        ++qn;
        ++found;
    };
    auto pop_one = [&]() {
        if (qn > 0) {
            const uint32_t k = q[qh * kBS];
            qh = (qh + 1) & (kQD - 1);
            --qn;
            apply(k);
        }
    };
    // pops until no lane holds more than `keep` pairs (written as a do-while: the
    // compiler then keeps the loop-carried sums in place instead of copying them
    // on every pop)
    auto pop_above = [&](int keep) {
        if (__any(qn > keep)) do {
                pop_one();
            } while (__any(qn > keep));
    };
    // one own candidate: record its skin bit, queue it when in range
    auto visit = [&](uint32_t k, float r2, uint32_t& hi, uint32_t lo) {
        if (kBuild) {
            if (r2 <= mk.rs2) word |= 1u << (cidx & 31);
            ++cidx;
            if ((cidx & 31) == 0) {
                const uint32_t w = (cidx >> 5) - 1;
                if (w < kMaskWords) mk.words[(uint64_t)w * mk.stride + s] = word;
                word = 0;
            }
        }
        if (k != s && r2 <= h2) {
            push(k);
            if (cap > 0 && found >= cap) hi = lo;
        }
    };
    for (int r = 0; r < st.rows; ++r) {
        uint32_t lo = 0, hi = 0;
        if (valid) {
            const uint2 sp = st.spans[(uint64_t)r * st.stride + m];
            lo = sp.x;
            hi = sp.y;
        }
        if (cap > 0 && found >= cap) hi = lo;  // neighbour cap reached: nothing more to take
        const bool has = lo < hi;
        const uint32_t ulo = wave_min_u32(has ? lo : 0xFFFFFFFFu);
        const uint32_t uhi = wave_max_u32(has ? hi : 0u);
        if (uhi <= ulo) continue;
        const uint32_t lmax = wave_max_u32(hi - lo);
        if (uhi - ulo > st.streamRatio * lmax) {
            // sparse row: each lane walks its own span, kUnroll candidates a round so
            // that their loads are in flight together; in-range ones are queued in order
            uint32_t k = lo;
            while (__any(k < hi)) {
                const uint32_t cnt = k < hi ? min(kUnroll, hi - k) : 0u;
                // k <= hi <= N, so the loads reach at most slot N + kUnroll - 2, inside
                // the kSlotPad slots past the last particle; tests past cnt are masked off
                float4 pj[kUnroll];
#pragma unroll
                for (uint32_t u = 0; u < kUnroll; ++u) pj[u] = P[k + u];
                const uint32_t inSpan = (1u << cnt) - 1u;
                const uint32_t self = s - k < kUnroll ? 1u << (s - k) : 0u;
                uint32_t bits = 0u, skin = 0u;
#pragma unroll
                for (uint32_t u = 0; u < kUnroll; ++u) {
                    const float r2 = dist2(pj[u]);
                    bits |= (r2 <= h2) ? 1u << u : 0u;
                    if (kBuild) skin |= (r2 <= mk.rs2) ? 1u << u : 0u;
                }
                bits &= inSpan & ~self;
                skin &= inSpan;
                if (kBuild && cnt > 0) {
                    const uint32_t sh = cidx & 31u;
                    uint64_t w = (uint64_t)word | ((uint64_t)skin << sh);
                    if (sh + cnt >= 32u) {
                        const uint32_t wi = cidx >> 5;
                        if (wi < kMaskWords) mk.words[(uint64_t)wi * mk.stride + s] = (uint32_t)w;
                        w >>= 32;
                    }
                    word = (uint32_t)w;
                    cidx += cnt;
                }
                if (cap > 0) {
                    const int room = cap - found;
                    while (__popc(bits) > room) bits &= ~(1u << (31 - __clz(bits)));
                    found += __popc(bits);
                    if (found >= cap) hi = k;  // nothing more to take in this row
                }
                pop_above(kQD - (int)kUnroll);
#pragma unroll
                for (uint32_t u = 0; u < kUnroll; ++u) {  // room for kUnroll pushes was made above
                    if ((bits >> u) & 1u) {
                        q[((qh + qn) & (kQD - 1)) * kBS] = k + u;
                        ++qn;
                    }
                }
                k += cnt;
            }
        } else if (kBuild || st.chunkBits == 0) {
            // dense row, candidate by candidate (and building the candidate mask):
            // stream the union through LDS
            for (uint32_t base = ulo; base < uhi; base += 64) {
                const uint32_t c = base + lane;
                if (c < uhi) buf[lane] = P[c];
                __builtin_amdgcn_wave_barrier();
                const uint32_t n = min(64u, uhi - base);
                for (uint32_t t = 0; t < n; ++t) {
                    if (__any(qn == kQD)) pop_one();
                    const uint32_t k = base + t;
                    if (k >= lo && k < hi) visit(k, dist2(buf[t]), hi, lo);
                }
                __builtin_amdgcn_wave_barrier();
            }
        } else {
            // dense row: stream the union through LDS, 64 candidates at a time.  Every
            // lane tests every candidate of a chunk in a branch-free wave-uniform loop
            // and keeps the in-range ones as bits; then only its own span's bits, minus
            // its own slot and anything past the neighbour cap, enter its FIFO in order.
            for (uint32_t base = ulo; base < uhi; base += 64) {
                const uint32_t c = base + lane;
                if (c < uhi) buf[lane] = P[c];
                __builtin_amdgcn_wave_barrier();
                const uint32_t n = min(64u, uhi - base);
                const uint32_t n0 = min(n, 32u);
                uint32_t m0 = 0u, m1 = 0u;
                for (uint32_t t = 0; t < n0; ++t) m0 |= (dist2(buf[t]) <= h2) ? (1u << t) : 0u;
                for (uint32_t t = 32; t < n; ++t) m1 |= (dist2(buf[t]) <= h2) ? (1u << (t - 32)) : 0u;
                __builtin_amdgcn_wave_barrier();
                const uint32_t a = lo > base ? min(lo - base, n) : 0u;
                const uint32_t b = hi > base ? min(hi - base, n) : 0u;
                uint64_t bits = (((uint64_t)m1 << 32) | m0) & low_bits64(b) & ~low_bits64(a);
                if (s - base < n) bits &= ~(1ull << (s - base));
                if (cap > 0) {
                    const int room = cap - found;
                    while (__popcll(bits) > room) bits &= ~(1ull << (63 - __clzll(bits)));
                    found += __popcll(bits);
                }
                while (__any(bits != 0ull)) {
                    if (__any(qn == kQD)) pop_one();
                    if (bits != 0ull && qn < kQD) {
                        const uint32_t t = (uint32_t)__ffsll((long long)bits) - 1u;
                        bits &= bits - 1ull;
                        q[((qh + qn) & (kQD - 1)) * kBS] = base + t;
                        ++qn;
                    }
                }
                // stop the row once no lane needs more of it
                if (!__any((cap <= 0 || found < cap) && hi > base + 64)) break;
            }
            if (cap > 0 && found >= cap) hi = lo;
        }
    }
    if (kBuild && valid && (cidx & 31) != 0) {
        const uint32_t w = cidx >> 5;
        if (w < kMaskWords) mk.words[(uint64_t)w * mk.stride + s] = word;
    }
    pop_above(0);
}

// Walk of the set mask bits (then the unmasked tail) in walk order, per lane.
// Each round takes up to kUnroll candidates, so their position loads are in flight
// together; the next mask word is loaded while the current one is consumed.
template <class Dist2, class Apply>
__device__ __forceinline__ void mask_walk(bool valid, uint32_t s, uint32_t m, const SpanTable& st,
                                          const CandMask& mk, const float4* P, uint32_t* q, int cap, float h2,
                                          Dist2&& dist2, Apply&& apply) {
    int qh = 0, qn = 0, found = 0;
    auto pop_one = [&]() {
        if (qn > 0) {
            const uint32_t k = q[qh * kBS];
            qh = (qh + 1) & (kQD - 1);
            --qn;
            apply(k);
        }
    };
    // pops until no lane holds more than `keep` pairs (written as a do-while: the
    // compiler then keeps the loop-carried sums in place instead of copying them
    // on every pop)
    auto pop_above = [&](int keep) {
        if (__any(qn > keep)) do {
                pop_one();
            } while (__any(qn > keep));
    };
    uint32_t total = 0;
    if (valid) {
        for (int r = 0; r < st.rows; ++r) {
            const uint2 sp = st.spans[(uint64_t)r * st.stride + m];
            total += sp.y - sp.x;
        }
    }
    // row cursor: candidate indices [rowStart, rowStart + rowLen) are slots rowLo + (ci - rowStart);
    // candidate indices only grow, so the cursor only moves forward
    int r = -1;
    uint32_t rowStart = 0, rowLen = 0, rowLo = 0;
    auto slot_of = [&](uint32_t ci) {
        while (ci >= rowStart + rowLen) {
            rowStart += rowLen;
            ++r;
            const uint2 sp = st.spans[(uint64_t)r * st.stride + m];
            rowLo = sp.x;
            rowLen = sp.y - sp.x;
        }
        return rowLo + (ci - rowStart);
    };
    const uint32_t nw = valid ? min((total + 31u) >> 5, kMaskWords) : 0u;
    const uint32_t covered = kMaskWords * 32u;
    uint32_t w = 0;
    uint32_t word = nw > 0 ? mk.words[s] : 0u;
    uint32_t next = nw > 1 ? mk.words[(uint64_t)mk.stride + s] : 0u;
    uint32_t tail = covered;  // next unmasked candidate index past the mask
    bool active = valid && total > 0;
    while (__any(active)) {
        pop_above(kQD - (int)kUnroll);
        // ks[u] for u >= cnt repeats the last taken slot (slot 0 when none was taken), so
        // every load below is in bounds and branch-free
        uint32_t ks[kUnroll];
        uint32_t cnt = 0, prev = 0u;
        bool more = active;
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) {
            if (more) {
                while (word == 0u && w + 1 < nw) {
                    ++w;
                    word = next;
                    next = (w + 1 < nw) ? mk.words[(uint64_t)(w + 1) * mk.stride + s] : 0u;
                }
                uint32_t ci = 0u;
                if (word != 0u) {
                    ci = w * 32u + (uint32_t)__ffs((int)word) - 1u;
                    word &= word - 1u;
                } else if (tail < total) {
                    ci = tail++;
                } else {
                    more = false;
                }
                if (more) {
                    prev = slot_of(ci);
                    ++cnt;
                }
            }
            ks[u] = prev;
        }
        if (cnt < kUnroll) active = false;  // every candidate taken
        float4 pj[kUnroll];
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) pj[u] = P[ks[u]];
        uint32_t bits = 0u;
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u)
            bits |= ((uint32_t)(u < cnt) & (uint32_t)(ks[u] != s) & (uint32_t)(dist2(pj[u]) <= h2)) << u;
        if (cap > 0) {
            const int room = cap - found;
            while (__popc(bits) > room) bits &= ~(1u << (31 - __clz(bits)));
            found += __popc(bits);
            if (found >= cap) active = false;
        }
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) {  // room for kUnroll pushes was made above
            if ((bits >> u) & 1u) {
                q[((qh + qn) & (kQD - 1)) * kBS] = ks[u];
                ++qn;
            }
        }
    }
    pop_above(0);
}

// Lambda of a sorted slot: the full walk (first iteration) or the mask walk.
__global__ __launch_bounds__(kBS) void k_lambda_sorted(float4* Ps, const uint32_t* __restrict__ segOf, SpanTable st,
                                                       CandMask mk, int useMask, sim::DeviceParams dp, uint32_t N,
                                                       float4* __restrict__ recA, uint2* __restrict__ recB) {
    __shared__ uint32_t sQ[kQD * kBS];
    __shared__ float4 sBuf[kBS];
    const uint32_t s = slot_of_thread();
    const bool valid = s < N;
    const uint32_t ss = valid ? s : N - 1;
    const float4 pi = Ps[ss];
    const uint32_t m = segOf[ss];
    const PairConsts c = make_pair_consts(dp);
    LambdaAcc a;
    uint32_t cut = 0xFFFFFFFFu;  // recording: slot of the neighbour that reached the cap
    const int cap = dp.maxNeighbors;
    auto dist2 = [&](const float4& pj) {
        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        return rx * rx + ry * ry + rz * rz;
    };
    auto apply = [&](uint32_t k) {
        const float4 pj = Ps[k];
        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        lambda_pair(a, rx, ry, rz, rx * rx + ry * ry + rz * rz, c);
        if (recB && a.n == cap) cut = k;
    };
    if (useMask)
        mask_walk(valid, s, m, st, mk, Ps, sQ + threadIdx.x, dp.maxNeighbors, c.h2, dist2, apply);
    else
        neighbor_walk<false>(valid, s, m, st, mk, Ps, sBuf + (threadIdx.x & ~63u), sQ + threadIdx.x,
                             dp.maxNeighbors, c.h2, dist2, apply);
    if (!valid) return;
    reinterpret_cast<float*>(Ps)[4ull * s + 3] = lambda_finish(a, dp);
    if (recA) {
        float b0, e0;
        lambda_adjoint_coefs(a, dp, b0, e0);
        recA[s] = make_float4(a.gx, a.gy, a.gz, b0);
        recB[s] = make_uint2(__float_as_uint(e0), cut);
    }
}

// Delta of a sorted slot.  maskMode: 0 full walk, 1 full walk that builds the
// candidate mask, 2 mask walk.
__global__ __launch_bounds__(kBS) void k_delta_sorted(const float4* __restrict__ Pa, float4* __restrict__ Pb,
                                                      const uint32_t* __restrict__ segOf, SpanTable st,
                                                      CandMask mk, int maskMode, sim::DeviceParams dp, uint32_t N,
                                                      int last, const float4* __restrict__ X,
                                                      const uint32_t* __restrict__ idxS, float4* __restrict__ Vs,
                                                      float invDt, Scenes sc, uint32_t* __restrict__ clampBits,
                                                      float4* __restrict__ recU) {
    __shared__ uint32_t sQ[kQD * kBS];
    __shared__ float4 sBuf[kBS];
    const uint32_t s = slot_of_thread();
    const bool valid = s < N;
    const uint32_t ss = valid ? s : N - 1;
    const float4 pi = Pa[ss];
    const uint32_t m = segOf[ss];
    const PairConsts c = make_pair_consts(dp);
    float dx = 0.f, dy = 0.f, dz = 0.f;
    auto dist2 = [&](const float4& pj) {
        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        return rx * rx + ry * ry + rz * rz;
    };
    auto apply = [&](uint32_t k) {
        const float4 pj = Pa[k];
        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        delta_pair(dx, dy, dz, rx, ry, rz, rx * rx + ry * ry + rz * rz, pi.w, pj.w, c);
    };
    float4* buf = sBuf + (threadIdx.x & ~63u);
    if (maskMode == 2)
        mask_walk(valid, s, m, st, mk, Pa, sQ + threadIdx.x, 0, c.h2, dist2, apply);
    else if (maskMode == 1)
        neighbor_walk<true>(valid, s, m, st, mk, Pa, buf, sQ + threadIdx.x, 0, c.h2, dist2, apply);
    else
        neighbor_walk<false>(valid, s, m, st, mk, Pa, buf, sQ + threadIdx.x, 0, c.h2, dist2, apply);
    if (!valid) return;
    const float3 e = delta_scaled(dx, dy, dz, dp);
    if (recU) recU[s] = make_float4(e.x, e.y, e.z, 0.f);
    const float3 d = delta_clamp(e, dp);
    // reference: pbf_delta_apply.cu:232-233 (P += delta)
    float px = pi.x + d.x, py = pi.y + d.y, pz = pi.z + d.z;
    if (!last) {
        Pb[s] = make_float4(px, py, pz, 0.0f);
        return;
    }
    // post-solve boundary (phases.cpp:94) and velocity (phases.cpp:102-105)
    float ux = 0.f, uy = 0.f, uz = 0.f;
    const sim::GridBounds& gb = sc.params ? sc.params[sc.sceneOfSlot[s]].grid : dp.grid;
    if (clampBits) clampBits[s] = clamp_bits(px, py, pz, gb);
    clamp_box(px, py, pz, ux, uy, uz, gb, 0.0f);
    Pb[s] = make_float4(px, py, pz, 1.0f);
    const float4 x = X[idxS[s]];
    Vs[s] = make_float4((px - x.x) * invDt, (py - x.y) * invDt, (pz - x.z) * invDt, 0.0f);
}

__global__ __launch_bounds__(kBS) void k_finalize_sorted(float4* __restrict__ Ps, const float4* __restrict__ X,
                                                         const uint32_t* __restrict__ idxS, float4* __restrict__ Vs,
                                                         sim::GridBounds g, float invDt, uint32_t N, Scenes sc) {
    const uint32_t s = blockIdx.x * kBS + threadIdx.x;
    if (s >= N) return;
    if (sc.params) g = sc.params[sc.sceneOfSlot[s]].grid;
    const float4 p = Ps[s];
    float px = p.x, py = p.y, pz = p.z, ux = 0.f, uy = 0.f, uz = 0.f;
    clamp_box(px, py, pz, ux, uy, uz, g, 0.0f);
    Ps[s] = make_float4(px, py, pz, 1.0f);
    const float4 x = X[idxS[s]];
    Vs[s] = make_float4((px - x.x) * invDt, (py - x.y) * invDt, (pz - x.z) * invDt, 0.0f);
}

// Eight waves per SIMD (64 VGPRs): the walk is latency-bound on its candidate and
// velocity loads, and the extra waves hide more of it than the few spilled registers
// cost (measured on the collapsed 1M dam break).
__global__ __launch_bounds__(kBS) __attribute__((amdgpu_waves_per_eu(8))) void k_xsph_scatter(const float4* __restrict__ Ps, const float4* __restrict__ Vs,
                                                      const uint32_t* __restrict__ segOf, SpanTable st,
                                                      CandMask mk, int useMask, sim::DeviceParams dp, uint32_t N,
                                                      const uint32_t* __restrict__ idxS, float4* __restrict__ Xn,
                                                      float4* __restrict__ V, Scenes sc, float* __restrict__ recGm) {
    __shared__ uint32_t sQ[kQD * kBS];
    __shared__ float4 sBuf[kBS];
    const uint32_t s = slot_of_thread();
    const bool valid = s < N;
    const uint32_t ss = valid ? s : N - 1;
    const float4 pi = Ps[ss];
    const float4 vi = Vs[ss];
    float vx = vi.x, vy = vi.y, vz = vi.z;
    float xc = dp.xsph_c, rest = dp.boundaryRestitution;
    sim::GridBounds grid = dp.grid;
    if (sc.params) {
        const SceneParams& sp = sc.params[sc.sceneOfSlot[ss]];
        xc = sp.xsph_c;
        rest = sp.restitution;
        grid = sp.grid;
    }
    // the walk is a wave-level operation: every lane enters it when any lane smooths
    const bool walk = xc > 0.f || recGm != nullptr;
    if (__any(walk)) {
        const uint32_t m = segOf[ss];
        const float h2 = dp.kernel.h2, poly6 = dp.kernel.poly6;
        float sx = 0.f, sy = 0.f, sz = 0.f;
        int n = 0;
        auto dist2 = [&](const float4& pj) {
            const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
            return rx * rx + ry * ry + rz * rz;
        };
        auto apply = [&](uint32_t k) {
            const float4 pj = Ps[k];
            const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
            const float hr2 = h2 - (rx * rx + ry * ry + rz * rz);
            const float w = poly6 * hr2 * hr2 * hr2;
            const float4 vj = Vs[k];
            sx += (vj.x - vi.x) * w;
            sy += (vj.y - vi.y) * w;
            sz += (vj.z - vi.z) * w;
            ++n;
        };
        if (useMask)
            mask_walk(valid && walk, s, m, st, mk, Ps, sQ + threadIdx.x, 0, h2, dist2, apply);
        else
            neighbor_walk<false>(valid && walk, s, m, st, mk, Ps, sBuf + (threadIdx.x & ~63u),
                                 sQ + threadIdx.x, 0, h2, dist2, apply);
        if (recGm && valid)
            recGm[s] = xsph_gate(n, dp) * ((dp.restDensity > 0.f) ? (dp.particleMass / dp.restDensity) : 0.f);
        if (xc > 0.f) {
            const float3 v = xsph_finish(vi.x, vi.y, vi.z, sx, sy, sz, n, dp, xc);
            vx = v.x; vy = v.y; vz = v.z;
        }
    }
    if (!valid) return;
    // BoundaryOp with restitution (post_ops.cpp:25-29, boundary.cu:11-25)
    float px = pi.x, py = pi.y, pz = pi.z;
    clamp_box(px, py, pz, vx, vy, vz, grid, clamp_restitution(rest));
    const uint32_t i = idxS[s];
    Xn[i] = make_float4(px, py, pz, 1.0f);
    V[i] = make_float4(vx, vy, vz, 0.0f);
}

inline uint32_t blocks_for(uint32_t n) { return ceil_div(n > 0 ? n : 1, kBS); }

}  // namespace

float mask_skin_radius2(const sim::SimParams& p) {
    if (!p.pbf.enable_disp_clamp || p.solverIters < 1 || !(p.pbf.disp_clamp_max_h >= 0.f)) return 0.f;
    const double h = p.kernel.h;
    const double rs = h * (1.0 + 2.0 * p.solverIters * (double)p.pbf.disp_clamp_max_h);
    if (rs > 2.0 * h) return 0.f;
    const double r = rs * (1.0 + 1e-3);  // margin for float rounding of positions and r^2
    return (float)(r * r);
}

void launch_predict_hash(const float4* X, const float4* V, float4* Xn, uint32_t* keys, uint32_t* epoch,
                         const sim::SimParams& p, const Scenes& sc, hipStream_t s, uint32_t* clampBits) {
    const uint32_t N = p.numParticles;
    k_predict_hash<<<blocks_for(N), kBS, 0, s>>>(X, V, Xn, keys, epoch, p.gravity, p.dt, p.grid, N, sc,
                                                 clampBits);
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
                                             HashRowsDev{table.slots, table.mask, table.epoch}, st, grid, sc);
            break;
    }
    check_launch("spans_gather");
}

void launch_lambda_sorted(float4* Ps, const uint32_t* segOf, const SpanTable& st, const CandMask& mk,
                          bool useMask, const sim::DeviceParams& dp, uint32_t N, hipStream_t s, float4* recA,
                          uint2* recB) {
    if (N == 0) return;
    if ((recA == nullptr) != (recB == nullptr)) {
        set_error(PBFX_E_INVALID_ARGUMENT, "launch_lambda_sorted: recA and recB go together");
        return;
    }
    k_lambda_sorted<<<blocks_for(N), kBS, 0, s>>>(Ps, segOf, st, mk, useMask ? 1 : 0, dp, N, recA, recB);
    check_launch("lambda_sorted");
}

void launch_delta_sorted(const float4* Pa, float4* Pb, const uint32_t* segOf, const SpanTable& st,
                         const CandMask& mk, int maskMode, const sim::DeviceParams& dp, uint32_t N, bool last,
                         const float4* X, const uint32_t* idxS, float4* Vs, float invDt, const Scenes& sc,
                         hipStream_t s, uint32_t* clampBits, float4* recU) {
    if (N == 0) return;
    k_delta_sorted<<<blocks_for(N), kBS, 0, s>>>(Pa, Pb, segOf, st, mk, maskMode, dp, N, last ? 1 : 0, X, idxS, Vs,
                                                 invDt, sc, clampBits, recU);
    check_launch("delta_sorted");
}

void launch_finalize_sorted(float4* Ps, const float4* X, const uint32_t* idxS, float4* Vs,
                            const sim::GridBounds& grid, float invDt, uint32_t N, const Scenes& sc,
                            hipStream_t s) {
    if (N == 0) return;
    k_finalize_sorted<<<blocks_for(N), kBS, 0, s>>>(Ps, X, idxS, Vs, grid, invDt, N, sc);
    check_launch("finalize_sorted");
}

void launch_xsph_scatter(const float4* Ps, const float4* Vs, const uint32_t* segOf, const SpanTable& st,
                         const CandMask& mk, bool useMask, const sim::DeviceParams& dp, uint32_t N,
                         const uint32_t* idxS, float4* Xn, float4* V, const Scenes& sc, hipStream_t s,
                         float* recGm) {
    if (N == 0) return;
    k_xsph_scatter<<<blocks_for(N), kBS, 0, s>>>(Ps, Vs, segOf, st, mk, useMask ? 1 : 0, dp, N, idxS, Xn, V, sc,
                                                 recGm);
    check_launch("xsph_scatter");
}

}  // namespace pbfx
