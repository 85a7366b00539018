// walk.hpp — the neighbour walk of one wave, shared by the cell-reordered kernels
// (k_solve_sorted.hip), the drop-in launchers over the reference layout
// (k_solve_ref.hip) and the reverse-pass kernels (k_adjoint.hip), and the walk of a
// candidate mask.  Every walk visits a particle's candidates in the reference's order.
#pragma once

#include "pbf_common.hpp"
#include "sorted_path.hpp"

namespace pbfx {
namespace walk {

constexpr int kBS = 256;  // lanes per block of the kernels that walk (default)

// A slot-indexed float4 array read through a buffer resource: each read is a 32-bit
// byte offset (slot * 16) from the descriptor's base, so the walk's candidate and pair
// loads need no 64-bit address arithmetic per candidate, and a round's consecutive
// slots become immediate offsets of one address.  Reads at or past `slots` return
// zeros.  Arrays of up to kMaxSlotBufSlots slots (the offsets are 32-bit).
constexpr uint32_t kMaxSlotBufSlots = 1u << 27;
// A slot index whose reads fall past the end of any SlotBuf (its byte offset lies
// beyond every array of at most kMaxSlotBufSlots slots).
constexpr uint32_t kSlotBufBeyond = 0x0FFFFFF0u;
typedef int SlotBufV4 __attribute__((ext_vector_type(4)));
struct SlotBuf {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ SlotBuf(const float4* p, uint32_t slots)
        // word 3 of a gfx9 buffer descriptor: 32-bit data format, raw addressing
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<float4*>(p), (short)0, (int)(slots * 16u), 0x00020000)) {}
    __device__ __forceinline__ float4 operator[](uint32_t k) const { return at(k, 0); }
    // slot k + u with u a compile-time constant in [0, 4): the u * 16 bytes ride in the
    // load's scalar offset as an inline constant, so a round of four shares one address
    __device__ __forceinline__ float4 at(uint32_t k, uint32_t u) const {
        const SlotBufV4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)(k << 4), (int)(u << 4), 0);
        return make_float4(__int_as_float(v.x), __int_as_float(v.y), __int_as_float(v.z), __int_as_float(v.w));
    }
};

// Neighbour walk of one wave (64 consecutive sorted slots, one particle per lane).
//
// Rows.  A row source gives the lane's particle its neighbour rows: count(), and
// span(r, lo, hi), the [lo, hi) of sorted slots of row r (rows dz-major, then dy).
// The cell-reordered kernels read them from the per-step span table; the drop-in
// launchers compute them from the cell lookup they are handed.  fetch.id(k) names the
// particle in sorted slot k (the slot itself, or the sorted index array's entry) and
// fetch.at(id) returns its position; the FIFO carries ids, so apply(id) reads the
// particle's data with one load.  When Fetch::kPadded, the slots past the last
// particle may be read.
//
// Two walk modes per neighbour row, chosen wave-uniformly.  Stream: when the union
// of the lanes' slot spans is at most rows.streamRatio times the longest lane span,
// the wave streams the union through a per-wave LDS buffer, 64 candidates at a time
// loaded coalesced, and every lane tests each candidate (an LDS broadcast) against
// its own span, itself and the distance h; with rows.chunkBits it tests a whole
// chunk into a bit mask first.  Per lane: otherwise each lane walks its own span
// from global memory, kUnroll candidates a round so their loads are in flight
// together.  When a wave spans several cells the union carries candidates most
// lanes do not need, and the per-lane walk is faster.
//
// Deferred pairs.  In-range candidates go to a per-lane FIFO in LDS; the expensive
// per-pair work (sqrt, IEEE divisions, the s_corr power) runs when some lane's
// FIFO is full (in the per-lane walk: when it has no room for a round's kUnroll
// pushes), for one queued pair on every lane that has one, and at the end.
// Otherwise a wave would execute the per-pair path whenever any of its 64 lanes
// finds a neighbour, at low utilisation.
//
// Order.  Each lane still sees its candidates in reference order (rows dz-major
// then dy, slots ascending within a row) and consumes its FIFO in order, so
// every sum is formed in the reference's order.
constexpr int kQD = 16;          // FIFO depth per lane
constexpr uint32_t kUnroll = 4;  // candidates per round of a lane's own walk
static_assert(kSlotPad + 1 >= kUnroll, "a round may load kUnroll - 1 slots past the last particle");

// The reference's range test: a candidate is skipped only when r2 > h2
// (pbf_lambda.cu:69, pbf_delta_apply.cu:68, xsph.cu:64), so a NaN distance (a NaN
// position) counts as in range and its NaN enters the sums.  Every range test of the
// walks, and the candidate mask's skin test, takes this form.
__device__ __forceinline__ bool in_range(float r2, float h2) { return !(r2 > h2); }

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
// h + 2 * (I - 1) * maxDisp of the predicted positions (k_mask_build, before the first
// iteration).  The passes that read the mask (the lambda passes of iterations 0 .. I-1)
// see positions at most I - 1 corrections later, each clamped to maxDisp, so a pair
// outside the skin cannot have come within h.  Those passes visit only the set bits, in
// walk order, and still test the range.  (The delta passes, which take every neighbour,
// and XSPH, after the last correction and the box clamp, walk their candidates
// directly.)
// The mask covers the first 32 * kMaskWords candidates; candidates past that are
// tested directly.
constexpr uint32_t kMaskWords = kCandMaskWords;

// Candidate u of a per-lane round starting at slot k: with a padded slot-order fetch,
// slot k + u read at an immediate offset from slot k's address; otherwise the fetch of
// the candidate's id.
template <class Fetch>
__device__ __forceinline__ auto fetch_round(const Fetch& f, uint32_t k, uint32_t u, uint32_t id)
    -> decltype(f.at(id)) {
    if constexpr (Fetch::kPadded)
        return f.at_next(k, u);
    else
        return f.at(id);
}

// Bits [0, n) of a 64-bit word (n <= 64).
__device__ __forceinline__ uint64_t low_bits64(uint32_t n) { return n >= 64u ? ~0ull : ((1ull << n) - 1ull); }

// kLanes: lanes per block of the calling kernel (the FIFO's stride in LDS).
template <int kLanes, class Rows, class Fetch, class Dist2, class Apply>
__device__ __forceinline__ void neighbor_walk(bool valid, uint32_t s, const Rows& rows, const Fetch& fetch, float4* buf,
                                              uint32_t* q, int cap, float h2, Dist2&& dist2, Apply&& apply) {
    const uint32_t lane = threadIdx.x & 63;
    int qh = 0, qn = 0, found = 0;
    auto push = [&](uint32_t k) {
        q[((qh + qn) & (kQD - 1)) * kLanes] = k;
        ++qn;
        ++found;
    };
    auto pop_one = [&]() {
        if (qn > 0) {
            const uint32_t k = q[qh * kLanes];
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
    // one own candidate: queue it when in range
    auto visit = [&](uint32_t k, float r2, uint32_t& hi, uint32_t lo) {
        if (k != s && in_range(r2, h2)) {
            push(fetch.id(k));
            if (cap > 0 && found >= cap) hi = lo;
        }
    };
    const int nrows = rows.count();
    for (int r = 0; r < nrows; ++r) {
        uint32_t lo = 0, hi = 0;
        if (valid) rows.span(r, lo, hi);
        if (cap > 0 && found >= cap) hi = lo;  // neighbour cap reached: nothing more to take
        const bool has = lo < hi;
        const uint32_t ulo = wave_min_u32(has ? lo : 0xFFFFFFFFu);
        const uint32_t uhi = wave_max_u32(has ? hi : 0u);
        if (uhi <= ulo) continue;
        const uint32_t lmax = wave_max_u32(hi - lo);
        if (uhi - ulo > rows.streamRatio * lmax) {
            // sparse row: each lane walks its own span, kUnroll candidates a round so
            // that their loads are in flight together; in-range ones are queued in order
            uint32_t k = lo;
            while (__any(k < hi)) {
                const uint32_t cnt = k < hi ? min(kUnroll, hi - k) : 0u;
                // padded: k <= hi <= N, so the loads reach at most slot N + kUnroll - 2,
                // inside the kSlotPad slots past the last particle; otherwise loads past
                // the span read its last slot.  Tests past cnt are masked off.
                const uint32_t last = hi > 0u ? hi - 1u : 0u;
                uint32_t id[kUnroll];
                float4 pj[kUnroll];
#pragma unroll
                for (uint32_t u = 0; u < kUnroll; ++u) id[u] = fetch.id(Fetch::kPadded ? k + u : min(k + u, last));
#pragma unroll
                for (uint32_t u = 0; u < kUnroll; ++u) pj[u] = fetch_round(fetch, k, u, id[u]);
                const uint32_t inSpan = (1u << cnt) - 1u;
                const uint32_t self = s - k < kUnroll ? 1u << (s - k) : 0u;
                uint32_t bits = 0u;
#pragma unroll
                for (uint32_t u = 0; u < kUnroll; ++u) bits |= in_range(dist2(pj[u]), h2) ? 1u << u : 0u;
                bits &= inSpan & ~self;
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
                        q[((qh + qn) & (kQD - 1)) * kLanes] = id[u];
                        ++qn;
                    }
                }
                k += cnt;
            }
        } else if (rows.chunkBits == 0) {
            // dense row, candidate by candidate: stream the union through LDS
            for (uint32_t base = ulo; base < uhi; base += 64) {
                const uint32_t c = base + lane;
                if (c < uhi) buf[lane] = fetch.at(fetch.id(c));
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
                if (c < uhi) buf[lane] = fetch.at(fetch.id(c));
                __builtin_amdgcn_wave_barrier();
                const uint32_t n = min(64u, uhi - base);
                const uint32_t n0 = min(n, 32u);
                uint32_t m0 = 0u, m1 = 0u;
                for (uint32_t t = 0; t < n0; ++t) m0 |= in_range(dist2(buf[t]), h2) ? (1u << t) : 0u;
                for (uint32_t t = 32; t < n; ++t) m1 |= in_range(dist2(buf[t]), h2) ? (1u << (t - 32)) : 0u;
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
                        q[((qh + qn) & (kQD - 1)) * kLanes] = fetch.id(base + t);
                        ++qn;
                    }
                }
                // stop the row once no lane needs more of it
                if (!__any((cap <= 0 || found < cap) && hi > base + 64)) break;
            }
            if (cap > 0 && found >= cap) hi = lo;
        }
    }
    pop_above(0);
}

// Walk of the set mask bits (then the unmasked tail) in walk order, per lane.
// Each round takes up to kUnroll candidates, so their position loads are in flight
// together; the next mask word is loaded while the current one is consumed.
template <int kLanes, class Dist2, class Apply>
__device__ __forceinline__ void mask_walk(bool valid, uint32_t s, uint32_t m, const SpanTable& st,
                                          const CandMask& mk, const SlotBuf& P, uint32_t* q, int cap, float h2,
                                          Dist2&& dist2, Apply&& apply) {
    int qh = 0, qn = 0, found = 0;
    auto pop_one = [&]() {
        if (qn > 0) {
            const uint32_t k = q[qh * kLanes];
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
    // Span-table entries are read at 32-bit indices r * stride + m: the table holds fewer
    // than 2^32 entries (GridBuffers::allocate refuses larger ones).
    uint32_t total = 0;
    if (valid) {
        uint32_t e = m;
        for (int r = 0; r < st.rows; ++r, e += st.stride) {
            const uint2 sp = st.spans[e];
            total += sp.y - sp.x;
        }
    }
    // Row cursor: the current row's candidate indices end at rowEnd, and its candidate ci
    // is slot ci + delta; spanAt indexes the next row's span.  A round takes candidates of
    // one row only, so the cursor moves before a round's candidates are taken, once per
    // round at most, instead of being checked for each of them.  A round that reaches
    // the end of its row takes fewer than kUnroll candidates.
    uint32_t rowEnd = 0, delta = 0, spanAt = m;
    const uint32_t nw = valid ? min((total + 31u) >> 5, kMaskWords) : 0u;
    const uint32_t covered = kMaskWords * 32u;
    uint32_t w = 0;
    uint32_t word = nw > 0 ? mk.words[s] : 0u;
    uint32_t next = nw > 1 ? mk.words[(uint64_t)mk.stride + s] : 0u;
    uint32_t tail = covered;  // next unmasked candidate index past the mask
    bool active = valid && total > 0;
    while (__any(active)) {
        pop_above(kQD - (int)kUnroll);
        // A round takes up to kUnroll candidates from one mask word (its lowest set bits),
        // or from the unmasked tail once the words are used up; empty words are skipped
        // once per round rather than once per candidate.
        if (active) {
            while (word == 0u && w + 1 < nw) {
                ++w;
                word = next;
                next = (w + 1 < nw) ? mk.words[(uint64_t)(w + 1) * mk.stride + s] : 0u;
            }
            // move the cursor to the row of the round's first candidate (every candidate
            // index is below total, the end of the last row, so the cursor stops in a row)
            const uint32_t first = word != 0u ? w * 32u + (uint32_t)__ffs((int)word) - 1u : tail;
            if (first < total) {
                while (first >= rowEnd) {
                    const uint2 sp = st.spans[spanAt];
                    spanAt += st.stride;
                    delta = sp.x - rowEnd;
                    rowEnd += sp.y - sp.x;
                }
            }
        }
        const bool fromTail = word == 0u;
        const uint32_t wbase = w * 32u;
        const uint32_t tailEnd = min(total, rowEnd);
        // ks[u] for u >= cnt repeats the last taken slot (slot 0 when none was taken), so
        // every load below is in bounds and branch-free; taken candidates come first
        uint32_t ks[kUnroll];
        uint32_t cnt = 0, prev = 0u;
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) {
            bool take;
            uint32_t ci;
            if (!fromTail) {
                ci = wbase + (uint32_t)__ffs((int)word) - 1u;
                take = (word != 0u) & (ci < rowEnd);
                if (take) word &= word - 1u;
            } else {
                ci = tail;
                take = tail < tailEnd;
                tail += take ? 1u : 0u;
            }
            if (active && take) {
                prev = ci + delta;
                ++cnt;
            }
            ks[u] = prev;
        }
        if (word == 0u && w + 1 >= nw && tail >= total) active = false;  // every candidate taken
        float4 pj[kUnroll];
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) pj[u] = P[ks[u]];
        uint32_t bits = 0u;
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u)
            bits |= ((uint32_t)(u < cnt) & (uint32_t)(ks[u] != s) & (uint32_t)in_range(dist2(pj[u]), h2)) << u;
        if (cap > 0) {
            const int room = cap - found;
            while (__popc(bits) > room) bits &= ~(1u << (31 - __clz(bits)));
            found += __popc(bits);
            if (found >= cap) active = false;
        }
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) {  // room for kUnroll pushes was made above
            if ((bits >> u) & 1u) {
                q[((qh + qn) & (kQD - 1)) * kLanes] = ks[u];
                ++qn;
            }
        }
    }
    pop_above(0);
}

// Walk without the pair queue, for a pair term cheap enough to evaluate on every
// candidate (XSPH's: a few products against the neighbour's velocity).  Each lane
// walks its own spans row by row, kUnroll candidates a round: load(k, u) reads the
// positions of slots k + u together, then gather(k, u, in) reads what the term needs
// besides, for the neighbours only (in: in the span, not the lane's own slot,
// in_range), and take() gets every candidate in walk order with its flag and selects
// the neighbours' terms into its sums.  No lane waits on another, so a lane whose
// spans end early simply stops.  Slot k + u may lie up to kUnroll - 1 slots past the
// span and past the last particle.
template <class Rows, class Load, class Dist2, class Gather, class Take>
__device__ __forceinline__ void inline_walk(bool valid, uint32_t s, const Rows& rows, float h2, Load&& load,
                                            Dist2&& dist2, Gather&& gather, Take&& take) {
    if (!valid) return;
    const int nrows = rows.count();
    for (int r = 0; r < nrows; ++r) {
        uint32_t lo = 0, hi = 0;
        rows.span(r, lo, hi);
        for (uint32_t k = lo; k < hi; k += kUnroll) {
            const uint32_t cnt = min(kUnroll, hi - k);
            decltype(load(k, 0u)) c[kUnroll];
            decltype(gather(k, 0u, true)) g[kUnroll];
            float r2[kUnroll];
            bool in[kUnroll];
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) c[u] = load(k, u);
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) {
                r2[u] = dist2(c[u]);
                in[u] = (u < cnt) & (k + u != s) & in_range(r2[u], h2);
            }
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) g[u] = gather(k, u, in[u]);
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) take(c[u], g[u], r2[u], in[u]);
        }
    }
}

// inline_walk for two particles of one cell (slots a and b; b valid only when hasB):
// they share the cell's neighbour rows, so each round's positions are read once for
// both, and a candidate's gather() once when it is a neighbour of either.  Each
// particle's take gets every candidate in walk order with its own flag.
template <class Rows, class Load, class Dist2A, class Dist2B, class Gather, class TakeA, class TakeB>
__device__ __forceinline__ void inline_walk_pair(uint32_t a, uint32_t b, bool hasB, const Rows& rows, float h2,
                                                 Load&& load, Dist2A&& dist2a, Dist2B&& dist2b, Gather&& gather,
                                                 TakeA&& takeA, TakeB&& takeB) {
    const int nrows = rows.count();
    for (int r = 0; r < nrows; ++r) {
        uint32_t lo = 0, hi = 0;
        rows.span(r, lo, hi);
        for (uint32_t k = lo; k < hi; k += kUnroll) {
            const uint32_t cnt = min(kUnroll, hi - k);
            decltype(load(k, 0u)) c[kUnroll];
            decltype(gather(k, 0u, true)) g[kUnroll];
            float ra[kUnroll], rb[kUnroll];
            bool ia[kUnroll], ib[kUnroll];
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) c[u] = load(k, u);
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) {
                ra[u] = dist2a(c[u]);
                rb[u] = dist2b(c[u]);
                ia[u] = (u < cnt) & (k + u != a) & in_range(ra[u], h2);
                ib[u] = hasB & (u < cnt) & (k + u != b) & in_range(rb[u], h2);
            }
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) g[u] = gather(k, u, ia[u] | ib[u]);
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) {
                takeA(g[u], ra[u], ia[u]);
                takeB(g[u], rb[u], ib[u]);
            }
        }
    }
}

// Rows of a sorted slot's neighbourhood from the per-step span table.
struct SpanRows {
    const SpanTable& st;
    uint32_t m;  // occupied-cell index of the slot
    uint32_t streamRatio, chunkBits;
    __device__ __forceinline__ SpanRows(const SpanTable& t, uint32_t cell)
        : st(t), m(cell), streamRatio(t.streamRatio), chunkBits(t.chunkBits) {}
    __device__ __forceinline__ int count() const { return st.rows; }
    __device__ __forceinline__ void span(int r, uint32_t& lo, uint32_t& hi) const {
        const uint2 sp = st.spans[(uint64_t)r * st.stride + m];
        lo = sp.x;
        hi = sp.y;
    }
};

// Sorted positions, padded by kSlotPad slots past the last particle; a particle's id
// is its slot.
struct SortedFetch {
    static constexpr bool kPadded = true;
    SlotBuf P;
    __device__ __forceinline__ uint32_t id(uint32_t k) const { return k; }
    __device__ __forceinline__ float4 at(uint32_t k) const { return P[k]; }
    // the particle in slot k + u (u < kUnroll, a compile-time constant)
    __device__ __forceinline__ float4 at_next(uint32_t k, uint32_t u) const { return P.at(k, u); }
};

}  // namespace walk
}  // namespace pbfx
