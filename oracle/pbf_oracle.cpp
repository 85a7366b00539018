// pbf_oracle.cpp — CPU oracle of the PBF solver step.  TEST INFRASTRUCTURE ONLY.
//
// A single-precision restatement of the reference's CUDA kernels, used as the
// checker for the HIP path.  It is built with -ffp-contract=off so every
// expression rounds exactly as written; expression association and neighbour
// order follow the cited reference lines, because a 1e-5*h position tolerance
// leaves room for about one ulp.  The product never links or calls this file.
//
// Deliberate semantic choices (documented where the reference is racy or
// headless-incomplete):
//   * XSPH reads a velocity snapshot (Jacobi); the reference runs it in place.
//   * lambda and delta take positions/lambda by pointer instead of device
//     globals (reference: pbf_lambda.cu:28,65 read g_pos_pred).
//   * after each step the predicted buffer becomes current (ping-pong swap,
//     reference: simulator.cpp:144-159).
//   * float->int conversions saturate (NaN -> 0), as GPU conversion does.
//
// Pins (tests/test_oracle_*.py): closed-form kernel constants, brute-force
// neighbour sums, lexicographic sort, linear-scan ranges, isolated-particle
// ballistic steps, two-particle s_corr closed form, lattice symmetry, uniform
// XSPH invariance, neighbour-cap ordering.
#include "pbf_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

static_assert(sizeof(OKernelCoeffs) == 24, "layout");
static_assert(sizeof(OGridBounds) == 40, "layout");
static_assert(sizeof(OPbfTuning) == 88, "layout");
static_assert(sizeof(ODeviceParams) == 192, "layout");
static_assert(offsetof(ODeviceParams, pbf) == 100, "layout");
static_assert(offsetof(ODeviceParams, xsph_c) == 188, "layout");
static_assert(sizeof(OSimParams) == 212, "layout");

namespace {

// Saturating float->int32 conversion with NaN -> 0 (the GPU's conversion rule).
inline int sat_f2i(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return std::numeric_limits<int>::max();
    if (f < -2147483648.0f) return std::numeric_limits<int>::min();
    return (int)f;
}

inline bool in_bounds(int x, int y, int z, const oi3& d) {
    return x >= 0 && x < d.x && y >= 0 && y < d.y && z >= 0 && z < d.z;
}

// Row-major linear cell index (reference: cuda_grid_utils.cuh:21-24).
inline uint32_t lin_idx(int x, int y, int z, const oi3& d) {
    return (uint32_t)((z * d.y + y) * d.x + x);
}

// Neighbour reach in cells (reference: pbf_lambda.cu:45-47).
inline int reach_of(float h, float cs) {
    return (cs > 0.f) ? std::max(1, sat_f2i(std::ceil(h / cs))) : 1;
}

// Key -> cell decode (reference: pbf_lambda.cu:31-36).
inline void decode_key(uint32_t key, const oi3& dim, int& x, int& y, int& z) {
    x = (int)(key % (uint32_t)dim.x);
    uint32_t kdx = key / (uint32_t)dim.x;
    y = (int)(kdx % (uint32_t)dim.y);
    z = (int)(kdx / (uint32_t)dim.y);
}

// Squared distance of a neighbour pair (reference: pbf_lambda.cu:68,
// pbf_delta_apply.cu:68, xsph.cu:62), with the products fused into the adds in this
// fixed association, as the reference's nvcc build contracts them; the GPU kernels
// form the same expression.  (The density statistics keep the unfused form of the
// statistics kernel they restate.)
inline float pair_r2(float rx, float ry, float rz) { return std::fma(rz, rz, std::fma(ry, ry, rx * rx)); }

// s_corr power (reference: pbf_delta_apply.cu:81 calls powf, documented by CUDA to
// 2 ulp).  Exponents 1 to 4 are float products by squaring, x, x*x, x*(x*x),
// (x*x)*(x*x): within 1.5 ulp of the power.  Other integral exponents up to 16 are the
// same binary exponentiation in double, rounded once to float; the rest call powf.
// Pinned against float64 powers and powf in the tests.
inline float pow_scorr(float x, float n) {
    if (n == 1.0f) return x;
    if (n == 2.0f) return x * x;
    if (n == 3.0f) return x * (x * x);
    if (n == 4.0f) {
        const float x2 = x * x;
        return x2 * x2;
    }
    if (n >= 1.0f && n <= 16.0f && (float)(int)n == n) {
        double b = (double)x, r = 1.0;
        int e = (int)n;
        while (true) {
            if (e & 1) r *= b;
            e >>= 1;
            if (!e) break;
            b *= b;
        }
        return (float)r;
    }
    return std::pow(x, n);
}

// Cell range providers: dense start/end arrays or compact unique-key search.
struct DenseRanges {
    const uint32_t* start;
    const uint32_t* end;
    bool get(uint32_t c, uint32_t& b, uint32_t& e) const {
        b = start[c];
        e = end[c];
        return !(b == 0xFFFFFFFFu || b >= e);  // reference: pbf_lambda.cu:56-58
    }
};

struct CompactRanges {
    const uint32_t* uniq;
    const uint32_t* off;
    uint32_t M;
    bool get(uint32_t c, uint32_t& b, uint32_t& e) const {
        return oracle_compact_cell_range(uniq, off, M, c, &b, &e) != 0;
    }
};

// Visits neighbours j != i of sorted slot s in reference order: dz, dy, dx in
// [-reach, reach] (z outermost), cells in bounds, slots ascending
// (reference: pbf_lambda.cu:49-84).  fn returns false to stop the whole walk.
template <class Ranges, class Fn>
inline void for_each_neighbor(uint32_t s, const uint32_t* idxS, const uint32_t* keyS, const OGridBounds& grid,
                              float h, const Ranges& rg, Fn&& fn) {
    const uint32_t i = idxS[s];
    int cx, cy, cz;
    decode_key(keyS[s], grid.dim, cx, cy, cz);
    const int reach = reach_of(h, grid.cellSize);
    for (int dz = -reach; dz <= reach; ++dz)
        for (int dy = -reach; dy <= reach; ++dy)
            for (int dx = -reach; dx <= reach; ++dx) {
                const int x = cx + dx, y = cy + dy, z = cz + dz;
                if (!in_bounds(x, y, z, grid.dim)) continue;
                uint32_t b, e;
                if (!rg.get(lin_idx(x, y, z, grid.dim), b, e)) continue;
                for (uint32_t k = b; k < e; ++k) {
                    const uint32_t j = idxS[k];
                    if (j == i) continue;
                    if (!fn(j)) return;
                }
            }
}

// Lambda for one sorted slot (reference: pbf_lambda.cu:18-107).
template <class Ranges>
float lambda_one(uint32_t s, const of4* P, const uint32_t* idxS, const uint32_t* keyS, const Ranges& rg,
                 const ODeviceParams& dp) {
    const OKernelCoeffs& kc = dp.kernel;
    const OPbfTuning& pbf = dp.pbf;
    const float m = dp.particleMass;
    const uint32_t i = idxS[s];
    const of4 pi = P[i];
    float density = 0.f, gx = 0.f, gy = 0.f, gz = 0.f, sumGrad2 = 0.f;
    int n = 0;
    const int cap = dp.maxNeighbors;
    const bool hasCap = cap > 0;
    if (!(hasCap && n >= cap)) {
        for_each_neighbor(s, idxS, keyS, dp.grid, kc.h, rg, [&](uint32_t j) -> bool {
            const of4 pj = P[j];
            const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
            const float r2 = pair_r2(rx, ry, rz);
            if (r2 > kc.h2) return true;
            n++;
            const float hr2 = kc.h2 - r2;
            density = std::fma(m * kc.poly6 * hr2 * hr2, hr2, density);
            const float r = std::sqrt(r2);
            const float r_safe = std::fmax(r, pbf.grad_r_eps);
            const float t = kc.h - r;
            const float coeff = (-3.0f * kc.spiky * (t * t) / r_safe);
            const float ax = coeff * rx, ay = coeff * ry, az = coeff * rz;
            gx += ax; gy += ay; gz += az;
            sumGrad2 += std::fma(az, az, std::fma(ay, ay, ax * ax));
            // cap: the reference stops at the next candidate or cell once n >= cap
            return !(hasCap && n >= cap);
        });
    }
    const float gradSum2 = gx * gx + gy * gy + gz * gz;
    const float scale = (dp.restDensity > 0.f) ? (m / dp.restDensity) : 0.f;
    float denom = (scale * scale) * (sumGrad2 + gradSum2) + pbf.lambda_denom_eps;
    if (pbf.compliance > 0.f && dp.dt > 0.f) denom += pbf.compliance / (dp.dt * dp.dt);
    denom = std::fmax(denom, pbf.lambda_denom_eps);
    if (n < 2 || (sumGrad2 + gradSum2) < 1e-12f) return 0.0f;
    const float C = (dp.restDensity > 0.f) ? (density / dp.restDensity - 1.f) : 0.f;
    float lam = -C / denom;
    if (!std::isfinite(lam)) lam = 0.f;
    if (pbf.enable_lambda_clamp) {
        if (std::fabs(lam) > pbf.lambda_max_abs) lam = std::copysign(pbf.lambda_max_abs, lam);
    }
    return lam;
}

// Delta for one sorted slot (reference: pbf_delta_apply.cu:21-112).
template <class Ranges>
of4 delta_one(uint32_t s, const of4* P, const float* lambda, const uint32_t* idxS, const uint32_t* keyS,
              const Ranges& rg, const ODeviceParams& dp) {
    const OKernelCoeffs& kc = dp.kernel;
    const OPbfTuning& pbf = dp.pbf;
    const uint32_t i = idxS[s];
    const of4 pi = P[i];
    const float dq = pbf.scorr_dq_h * kc.h;
    const float hr2_q = kc.h2 - dq * dq;
    const float w_q = (hr2_q > 0.f) ? (kc.poly6 * hr2_q * hr2_q * hr2_q) : pbf.wq_min;
    float dx = 0.f, dy = 0.f, dz = 0.f;
    for_each_neighbor(s, idxS, keyS, dp.grid, kc.h, rg, [&](uint32_t j) -> bool {
        const of4 pj = P[j];
        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        const float r2 = pair_r2(rx, ry, rz);
        if (r2 > kc.h2) return true;
        const float r = std::sqrt(r2);
        const float t = kc.h - r;
        const float r_safe = std::fmax(r, pbf.grad_r_eps);
        const float coeff = (-3.0f * kc.spiky * (t * t) / r_safe);
        const float ax = coeff * rx, ay = coeff * ry, az = coeff * rz;
        float scorr = 0.0f;
        if (pbf.scorr_enable) {
            const float hr2 = kc.h2 - r2;
            const float w = kc.poly6 * hr2 * hr2 * hr2;
            const float ratio = (w_q > 0.f) ? (w / w_q) : 0.0f;
            scorr = -pbf.scorr_k * pow_scorr(ratio, pbf.scorr_n);
            if (pbf.scorr_min < 0.f) scorr = std::fmax(scorr, pbf.scorr_min);
        }
        const float lij = (lambda[i] + lambda[j] + scorr);
        dx = std::fma(lij, ax, dx);
        dy = std::fma(lij, ay, dy);
        dz = std::fma(lij, az, dz);
        return true;
    });
    const float invRest = (dp.restDensity > 0.f) ? (1.0f / dp.restDensity) : 0.f;
    float ex = invRest * dx, ey = invRest * dy, ez = invRest * dz;
    if (pbf.enable_relax) {
        ex *= pbf.relax_omega;
        ey *= pbf.relax_omega;
        ez *= pbf.relax_omega;
    }
    if (pbf.enable_disp_clamp) {
        const float maxDisp = pbf.disp_clamp_max_h * kc.h;
        const float maxDisp2 = maxDisp * maxDisp;
        const float len2 = ex * ex + ey * ey + ez * ez;
        if (len2 > maxDisp2) {
            const float len = std::sqrt(len2);
            const float sc = maxDisp / (len + 1e-20f);
            ex *= sc; ey *= sc; ez *= sc;
        }
    }
    return of4{ex, ey, ez, 0.0f};
}

// XSPH for one sorted slot, Jacobi read of velIn (reference: xsph.cu:20-95).
template <class Ranges>
of4 xsph_one(uint32_t s, const of4* V, const of4* P, const uint32_t* idxS, const uint32_t* keyS,
             const Ranges& rg, const ODeviceParams& dp) {
    const OKernelCoeffs& kc = dp.kernel;
    const uint32_t i = idxS[s];
    const of4 pi = P[i];
    const of4 vi = V[i];
    float sx = 0.f, sy = 0.f, sz = 0.f;
    int n = 0;
    for_each_neighbor(s, idxS, keyS, dp.grid, kc.h, rg, [&](uint32_t j) -> bool {
        const of4 pj = P[j];
        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        const float r2 = pair_r2(rx, ry, rz);
        if (r2 > kc.h2) return true;
        const float hr2 = kc.h2 - r2;
        const float w = kc.poly6 * hr2 * hr2 * hr2;
        const of4 vj = V[j];
        sx = std::fma(vj.x - vi.x, w, sx);
        sy = std::fma(vj.y - vi.y, w, sy);
        sz = std::fma(vj.z - vi.z, w, sz);
        ++n;
        return true;
    });
    float gate = 1.0f;
    if (dp.pbf.xsph_gate_enable) {
        const int nMin = std::max(0, dp.pbf.xsph_n_min);
        const int nMax = std::max(nMin + 1, dp.pbf.xsph_n_max);
        const float t = ((float)n - (float)nMin) / (float)(nMax - nMin);
        gate = std::fmin(1.f, std::fmax(0.f, t));
    }
    const float mor = (dp.restDensity > 0.f) ? (dp.particleMass / dp.restDensity) : 0.f;
    const float c = dp.xsph_c;
    return of4{vi.x + (c * gate) * mor * sx, vi.y + (c * gate) * mor * sy, vi.z + (c * gate) * mor * sz, 0.0f};
}

template <class Ranges>
void lambda_all(float* lambda, const of4* P, const uint32_t* idxS, const uint32_t* keyS, const Ranges& rg,
                const ODeviceParams& dp, uint32_t N) {
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < (int64_t)N; ++s) lambda[idxS[s]] = lambda_one((uint32_t)s, P, idxS, keyS, rg, dp);
}

// Delta compute for all slots, then the separate apply pass
// (reference: pbf_delta_apply.cu:223-235, launched after KDeltaCompute at :264-266).
template <class Ranges>
void delta_apply_all(of4* P, of4* delta, const float* lambda, const uint32_t* idxS, const uint32_t* keyS,
                     const Ranges& rg, const ODeviceParams& dp, uint32_t N) {
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < (int64_t)N; ++s) delta[idxS[s]] = delta_one((uint32_t)s, P, lambda, idxS, keyS, rg, dp);
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < (int64_t)N; ++s) {
        const uint32_t i = idxS[s];
        P[i].x = P[i].x + delta[i].x;
        P[i].y = P[i].y + delta[i].y;
        P[i].z = P[i].z + delta[i].z;
    }
}

template <class Ranges>
void xsph_all(of4* velOut, const of4* velIn, const of4* P, const uint32_t* idxS, const uint32_t* keyS,
              const Ranges& rg, const ODeviceParams& dp, uint32_t N) {
    if (N == 0 || dp.xsph_c <= 0.f) return;  // reference: xsph.cu:201
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < (int64_t)N; ++s) velOut[idxS[s]] = xsph_one((uint32_t)s, velIn, P, idxS, keyS, rg, dp);
}

// Grid scratch for one step.
struct Grid {
    std::vector<uint32_t> keys, idx, keyS, idxS, start, end, uniq, off;
    uint32_t count = 0;
};

void build_grid(Grid& g, const of4* P, const OGridBounds& grid, uint32_t N, bool compact) {
    g.keys.resize(N); g.idx.resize(N); g.keyS.resize(N); g.idxS.resize(N);
    oracle_hash_keys(g.keys.data(), g.idx.data(), P, &grid, N);
    oracle_sort_pairs(g.keys.data(), g.keyS.data(), g.idx.data(), g.idxS.data(), N);
    if (compact) {
        g.uniq.resize(std::max<uint32_t>(N, 1)); g.off.resize(std::max<uint32_t>(N, 1) + 1);
        oracle_cell_compact(g.uniq.data(), g.off.data(), &g.count, g.keyS.data(), N);
    } else {
        const uint32_t C = (uint32_t)grid.dim.x * (uint32_t)grid.dim.y * (uint32_t)grid.dim.z;
        g.start.resize(C); g.end.resize(C);
        oracle_cell_ranges(g.start.data(), g.end.data(), g.keyS.data(), N, C);
    }
}

}  // namespace

extern "C" {

void oracle_make_kernel_coeffs(float h, OKernelCoeffs* kc) {
    // reference: parameters.h:94-107
    const float pi = 3.14159265358979323846f;
    kc->h = h;
    kc->inv_h = 1.0f / h;
    kc->h2 = h * h;
    const float h3 = h * kc->h2;
    const float h6 = h3 * h3;
    const float h9 = h6 * h3;
    kc->poly6 = 315.0f / (64.0f * pi * h9);
    kc->spiky = 15.0f / (pi * h6);
    kc->visc = 15.0f / (2.0f * pi * h3);
}

void oracle_make_device_params(const OSimParams* sp, ODeviceParams* dp) {
    // reference: parameters.h:109-123
    std::memset(dp, 0, sizeof(*dp));
    dp->kernel = sp->kernel;
    dp->grid = sp->grid;
    dp->gravity = sp->gravity;
    dp->restDensity = sp->restDensity;
    dp->dt = sp->dt;
    dp->inv_dt = (sp->dt > 0.f) ? (1.0f / sp->dt) : 0.0f;
    dp->boundaryRestitution = sp->boundaryRestitution;
    dp->maxNeighbors = sp->maxNeighbors;
    dp->particleMass = sp->particleMass;
    dp->pbf = sp->pbf;
    dp->xsph_c = sp->xsph_c;
}

void oracle_compute_dims(const OGridBounds* g, oi3* d) {
    // reference: grid_system.h:9-17
    const float cs = (g->cellSize > 0.f) ? g->cellSize : 1.f;
    d->x = sat_f2i(std::ceil((g->maxs.x - g->mins.x) / cs));
    d->y = sat_f2i(std::ceil((g->maxs.y - g->mins.y) / cs));
    d->z = sat_f2i(std::ceil((g->maxs.z - g->mins.z) / cs));
    if (d->x < 0) d->x = 0;
    if (d->y < 0) d->y = 0;
    if (d->z < 0) d->z = 0;
}

void oracle_integrate_pred(const of4* X, const of4* V, of4* P, float gx, float gy, float gz, float dt,
                           uint32_t N) {
    // reference: integrate.cu:37-51
    if (!X || !V || !P || N == 0) return;
    for (uint32_t i = 0; i < N; ++i) {
        const of4 p4 = X[i], v4 = V[i];
        const float vx = v4.x + gx * dt;
        const float vy = v4.y + gy * dt;
        const float vz = v4.z + gz * dt;
        P[i] = of4{p4.x + vx * dt, p4.y + vy * dt, p4.z + vz * dt, p4.w};
    }
}

void oracle_boundary(of4* P, of4* V, const OGridBounds* g, float restitution, uint32_t N) {
    // reference: boundary.cu:11-25
    const float e = std::fmin(std::fmax(restitution, 0.0f), 1.0f);
    for (uint32_t i = 0; i < N; ++i) {
        float px = P[i].x, py = P[i].y, pz = P[i].z;
        float vx = V[i].x, vy = V[i].y, vz = V[i].z;
        if (px < g->mins.x) { px = g->mins.x; vx *= -e; }
        if (py < g->mins.y) { py = g->mins.y; vy *= -e; }
        if (pz < g->mins.z) { pz = g->mins.z; vz *= -e; }
        if (px > g->maxs.x) { px = g->maxs.x; vx *= -e; }
        if (py > g->maxs.y) { py = g->maxs.y; vy *= -e; }
        if (pz > g->maxs.z) { pz = g->maxs.z; vz *= -e; }
        P[i] = of4{px, py, pz, 1.0f};
        V[i] = of4{vx, vy, vz, 0.0f};
    }
}

void oracle_velocity(of4* V, const of4* X, const of4* P, float invDt, uint32_t N) {
    // reference: boundary.cu:28-34
    for (uint32_t i = 0; i < N; ++i)
        V[i] = of4{(P[i].x - X[i].x) * invDt, (P[i].y - X[i].y) * invDt, (P[i].z - X[i].z) * invDt, 0.0f};
}

void oracle_hash_keys(uint32_t* keys, uint32_t* idx, const of4* P, const OGridBounds* g, uint32_t N) {
    // reference: hash_keys.cu:6-27 (cell clamped to [0, dim-1] per axis)
    for (uint32_t i = 0; i < N; ++i) {
        const float rx = (P[i].x - g->mins.x) / g->cellSize;
        const float ry = (P[i].y - g->mins.y) / g->cellSize;
        const float rz = (P[i].z - g->mins.z) / g->cellSize;
        int cx = sat_f2i(std::floor(rx)), cy = sat_f2i(std::floor(ry)), cz = sat_f2i(std::floor(rz));
        cx = std::max(0, std::min(cx, g->dim.x - 1));
        cy = std::max(0, std::min(cy, g->dim.y - 1));
        cz = std::max(0, std::min(cz, g->dim.z - 1));
        keys[i] = lin_idx(cx, cy, cz, g->dim);
        idx[i] = i;
    }
}

void oracle_sort_pairs(const uint32_t* kin, uint32_t* kout, const uint32_t* vin, uint32_t* vout, uint32_t N) {
    // reference: sort_pairs.cu:18-27 — CUB DeviceRadixSort::SortPairs over bits [0,32):
    // a stable sort by key.
    std::vector<uint32_t> perm(N);
    std::iota(perm.begin(), perm.end(), 0u);
    std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return kin[a] < kin[b]; });
    std::vector<uint32_t> k(N), v(N);
    for (uint32_t s = 0; s < N; ++s) { k[s] = kin[perm[s]]; v[s] = vin[perm[s]]; }
    std::copy(k.begin(), k.end(), kout);
    std::copy(v.begin(), v.end(), vout);
}

void oracle_cell_ranges(uint32_t* start, uint32_t* end, const uint32_t* keyS, uint32_t N, uint32_t numCells) {
    // reference: cell_ranges.cu:62-63 (memsets) and :21-51 (boundary detection)
    for (uint32_t c = 0; c < numCells; ++c) { start[c] = 0xFFFFFFFFu; end[c] = 0u; }
    if (N == 0 || numCells == 0) return;
    for (uint32_t i = 0; i < N; ++i) {
        const uint32_t k = keyS[i];
        if (k >= numCells) continue;
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
}

void oracle_cell_compact(uint32_t* uniq, uint32_t* off, uint32_t* count, const uint32_t* keyS, uint32_t N) {
    // reference: cell_compact.cu:187-199 (N == 0) and the segment semantics of :96-184 / :8-36
    if (N == 0) {
        off[0] = 0;
        *count = 0;
        return;
    }
    uint32_t m = 0;
    for (uint32_t i = 0; i < N; ++i) {
        if (i == 0 || keyS[i] != keyS[i - 1]) {
            uniq[m] = keyS[i];
            off[m] = i;
            ++m;
        }
    }
    off[m] = N;
    *count = m;
}

int oracle_compact_cell_range(const uint32_t* uniq, const uint32_t* off, uint32_t M, uint32_t key, uint32_t* beg,
                              uint32_t* end) {
    // reference: cuda_grid_utils.cuh:34-67
    if (M == 0u) { *beg = *end = 0u; return 0; }
    int lo = 0, hi = (int)M;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (uniq[mid] < key) lo = mid + 1; else hi = mid;
    }
    if (lo >= (int)M || uniq[lo] != key) { *beg = *end = 0u; return 0; }
    *beg = off[lo];
    *end = off[lo + 1];
    return *beg < *end ? 1 : 0;
}

void oracle_lambda_dense(float* lambda, const of4* P, const uint32_t* idxS, const uint32_t* keyS,
                         const uint32_t* cs, const uint32_t* ce, const ODeviceParams* dp, uint32_t N) {
    if (N == 0) return;
    lambda_all(lambda, P, idxS, keyS, DenseRanges{cs, ce}, *dp, N);
}

void oracle_lambda_compact(float* lambda, const of4* P, const uint32_t* idxS, const uint32_t* keyS,
                           const uint32_t* uniq, const uint32_t* off, const uint32_t* count,
                           const ODeviceParams* dp, uint32_t N) {
    if (N == 0) return;
    lambda_all(lambda, P, idxS, keyS, CompactRanges{uniq, off, *count}, *dp, N);
}

void oracle_delta_apply_dense(of4* P, of4* delta, const float* lambda, const uint32_t* idxS, const uint32_t* keyS,
                              const uint32_t* cs, const uint32_t* ce, const ODeviceParams* dp, uint32_t N) {
    if (N == 0) return;
    delta_apply_all(P, delta, lambda, idxS, keyS, DenseRanges{cs, ce}, *dp, N);
}

void oracle_delta_apply_compact(of4* P, of4* delta, const float* lambda, const uint32_t* idxS, const uint32_t* keyS,
                                const uint32_t* uniq, const uint32_t* off, const uint32_t* count,
                                const ODeviceParams* dp, uint32_t N) {
    if (N == 0) return;
    delta_apply_all(P, delta, lambda, idxS, keyS, CompactRanges{uniq, off, *count}, *dp, N);
}

void oracle_xsph_dense(of4* vo, const of4* vi, const of4* P, const uint32_t* idxS, const uint32_t* keyS,
                       const uint32_t* cs, const uint32_t* ce, const ODeviceParams* dp, uint32_t N) {
    xsph_all(vo, vi, P, idxS, keyS, DenseRanges{cs, ce}, *dp, N);
}

void oracle_xsph_compact(of4* vo, const of4* vi, const of4* P, const uint32_t* idxS, const uint32_t* keyS,
                         const uint32_t* uniq, const uint32_t* off, const uint32_t* count, const ODeviceParams* dp,
                         uint32_t N) {
    xsph_all(vo, vi, P, idxS, keyS, CompactRanges{uniq, off, *count}, *dp, N);
}

int oracle_run(of4* X, of4* V, const OSimParams* sp, int compact, int steps) {
    // One step = phases.cpp:108-117 order, then post ops (post_ops.cpp:239-251), then swap.
    const uint32_t N = sp->numParticles;
    if (N == 0) return 0;
    ODeviceParams dp;
    oracle_make_device_params(sp, &dp);
    std::vector<of4> P(N), delta(N), Vsnap(N);
    std::vector<float> lambda(N);
    Grid g;
    for (int step = 0; step < steps; ++step) {
        // PhaseIntegrate (phases.cpp:33-37)
        oracle_integrate_pred(X, V, P.data(), sp->gravity.x, sp->gravity.y, sp->gravity.z, sp->dt, N);
        oracle_boundary(P.data(), V, &sp->grid, 0.0f, N);
        // PhaseGridBuildFull (phases.cpp:44-63)
        build_grid(g, P.data(), sp->grid, N, compact != 0);
        // PhaseSolveIterations (phases.cpp:70-95)
        for (int it = 0; it < sp->solverIters; ++it) {
            if (compact) {
                CompactRanges rg{g.uniq.data(), g.off.data(), g.count};
                lambda_all(lambda.data(), P.data(), g.idxS.data(), g.keyS.data(), rg, dp, N);
                delta_apply_all(P.data(), delta.data(), lambda.data(), g.idxS.data(), g.keyS.data(), rg, dp, N);
            } else {
                DenseRanges rg{g.start.data(), g.end.data()};
                lambda_all(lambda.data(), P.data(), g.idxS.data(), g.keyS.data(), rg, dp, N);
                delta_apply_all(P.data(), delta.data(), lambda.data(), g.idxS.data(), g.keyS.data(), rg, dp, N);
            }
        }
        oracle_boundary(P.data(), V, &sp->grid, 0.0f, N);
        // PhaseVelocity (phases.cpp:102-105): inv_dt computed on the host as 1.0f / dt
        oracle_velocity(V, X, P.data(), 1.0f / sp->dt, N);
        // XsphOp (post_ops.cpp:54-71), Jacobi; then BoundaryOp with restitution (post_ops.cpp:25-29)
        if (sp->xsph_c > 0.f) {
            std::copy(V, V + N, Vsnap.begin());
            if (compact) {
                xsph_all(V, Vsnap.data(), P.data(), g.idxS.data(), g.keyS.data(),
                         CompactRanges{g.uniq.data(), g.off.data(), g.count}, dp, N);
            } else {
                xsph_all(V, Vsnap.data(), P.data(), g.idxS.data(), g.keyS.data(),
                         DenseRanges{g.start.data(), g.end.data()}, dp, N);
            }
        }
        oracle_boundary(P.data(), V, &sp->grid, sp->boundaryRestitution, N);
        // ping-pong swap (simulator.cpp:144-159): predicted becomes current
        std::copy(P.begin(), P.end(), X);
    }
    return 0;
}

void oracle_density_stats(const of4* X, const of4* V, const OSimParams* sp, double* out) {
    const uint32_t N = sp->numParticles;
    for (int k = 0; k < 5; ++k) out[k] = 0.0;
    if (N == 0) return;
    ODeviceParams dp;
    oracle_make_device_params(sp, &dp);
    Grid g;
    build_grid(g, X, sp->grid, N, false);
    DenseRanges rg{g.start.data(), g.end.data()};
    const OKernelCoeffs& kc = dp.kernel;
    double sumErr = 0.0, maxErr = -1e300, sumN = 0.0, sumSpeed = 0.0, sumPos = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : sumErr, sumN, sumSpeed, sumPos) reduction(max : maxErr)
    for (int64_t s = 0; s < (int64_t)N; ++s) {
        const uint32_t i = g.idxS[s];
        const of4 pi = X[i];
        float density = 0.f;
        int n = 0;
        const int cap = dp.maxNeighbors;
        for_each_neighbor((uint32_t)s, g.idxS.data(), g.keyS.data(), dp.grid, kc.h, rg, [&](uint32_t j) -> bool {
            const of4 pj = X[j];
            const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
            const float r2 = rx * rx + ry * ry + rz * rz;
            if (r2 > kc.h2) return true;
            n++;
            const float hr2 = kc.h2 - r2;
            density += dp.particleMass * kc.poly6 * hr2 * hr2 * hr2;
            return !(cap > 0 && n >= cap);
        });
        const double err = (double)density / (double)dp.restDensity - 1.0;
        sumErr += err;
        sumPos += err > 0.0 ? err : 0.0;
        maxErr = std::max(maxErr, err);
        sumN += n;
        const of4 v = V[i];
        sumSpeed += std::sqrt((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z);
    }
    out[0] = sumErr / N;
    out[1] = maxErr;
    out[2] = sumN / N;
    out[3] = sumSpeed / N;
    out[4] = sumPos / N;
}

float oracle_jitter_uniform(uint64_t seed, uint64_t particle, uint32_t axis) {
    // SplitMix64 finaliser (Steele, Lea, Flood 2014) over (seed, 3*particle+axis+1).
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (particle * 3ull + (uint64_t)axis + 1ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (float)(uint32_t)(z >> 40) * (1.0f / 16777216.0f);
}

void oracle_seed_lattice(of4* X, uint32_t nx, uint32_t ny, uint32_t nz, float ox, float oy, float oz,
                         float spacing, const OGridBounds* g, float amp, uint64_t seed) {
    // reference: simulator.cpp:826-840 (lattice, x fastest, clamp to box) and :843-868 (jitter, re-clamp)
    uint32_t idx = 0;
    for (uint32_t iz = 0; iz < nz; ++iz)
        for (uint32_t iy = 0; iy < ny; ++iy)
            for (uint32_t ix = 0; ix < nx; ++ix) {
                float x = ox + (float)ix * spacing;
                float y = oy + (float)iy * spacing;
                float z = oz + (float)iz * spacing;
                x = std::fmin(std::fmax(x, g->mins.x), g->maxs.x);
                y = std::fmin(std::fmax(y, g->mins.y), g->maxs.y);
                z = std::fmin(std::fmax(z, g->mins.z), g->maxs.z);
                X[idx++] = of4{x, y, z, 1.0f};
            }
    if (amp > 0.f) {
        const uint32_t N = nx * ny * nz;
        for (uint32_t i = 0; i < N; ++i) {
            const float jx = (2.0f * oracle_jitter_uniform(seed, i, 0) - 1.0f) * amp;
            const float jy = (2.0f * oracle_jitter_uniform(seed, i, 1) - 1.0f) * amp;
            const float jz = (2.0f * oracle_jitter_uniform(seed, i, 2) - 1.0f) * amp;
            X[i].x = std::fmin(std::fmax(X[i].x + jx, g->mins.x), g->maxs.x);
            X[i].y = std::fmin(std::fmax(X[i].y + jy, g->mins.y), g->maxs.y);
            X[i].z = std::fmin(std::fmax(X[i].z + jz, g->mins.z), g->maxs.z);
        }
    }
}

float oracle_pow_scorr(float x, float n) { return pow_scorr(x, n); }

void oracle_set_num_threads(int n) {
#ifdef _OPENMP
    if (n >= 1) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
