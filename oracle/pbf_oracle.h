/* pbf_oracle.h — CPU oracle of the PBF solver step.  TEST INFRASTRUCTURE ONLY.
 *
 * This is the checker the GPU path is compared against: a single-precision,
 * no-FMA restatement of the reference's CUDA kernels.  Only tests/, the smoke
 * check and bench.py's cpu_baseline leg may load it; the product never does.
 *
 * Plain C API over POD mirrors of the reference's parameter structs
 * (reference: sim/parameters.h:9-92).  Layouts equal the product's
 * include/pbf_params.h: KernelCoeffs 24 B, GridBounds 40 B, PbfTuning 88 B,
 * DeviceParams 192 B, SimParams 212 B.  Particle arrays are float4 AoS
 * (x, y, z, w) in original particle order, as in the reference.
 */
#ifndef PBF_ORACLE_H
#define PBF_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float x, y, z; } of3;
typedef struct { int x, y, z; } oi3;
typedef struct { float x, y, z, w; } of4;

typedef struct { float h, inv_h, h2, poly6, spiky, visc; } OKernelCoeffs;
typedef struct { of3 mins; of3 maxs; float cellSize; oi3 dim; } OGridBounds;
typedef struct {
    int scorr_enable; float scorr_k, scorr_n, scorr_dq_h, wq_min, scorr_min;
    float grad_r_eps, lambda_denom_eps;
    float compliance; int xpbd_enable;
    int enable_lambda_clamp; float lambda_max_abs;
    int enable_disp_clamp; float disp_clamp_max_h;
    int enable_relax; float relax_omega;
    int xsph_gate_enable, xsph_n_min, xsph_n_max;
    int lambda_warm_start_enable; float lambda_warm_start_decay;
    int semi_implicit_integration_enable;
} OPbfTuning;
typedef struct {
    uint32_t numParticles, maxParticles; of3 gravity; float dt, cfl, restDensity;
    int solverIters, maxNeighbors, sortEveryN; float boundaryRestitution, particleMass;
    OPbfTuning pbf; float xsph_c; OKernelCoeffs kernel; OGridBounds grid; uint32_t ghostParticleCount;
} OSimParams;
typedef struct {
    OKernelCoeffs kernel; OGridBounds grid; of3 gravity;
    float restDensity, dt, inv_dt, boundaryRestitution; int maxNeighbors; float particleMass;
    OPbfTuning pbf; float xsph_c;
} ODeviceParams;

/* --- parameters (reference: sim/parameters.h:94-123, grid_system.h:9-17) --- */
void oracle_make_kernel_coeffs(float h, OKernelCoeffs* out);
void oracle_make_device_params(const OSimParams* sp, ODeviceParams* out);
void oracle_compute_dims(const OGridBounds* g, oi3* out);

/* --- integrate / boundary / velocity (reference: integrate.cu:29-52, boundary.cu:11-34) --- */
void oracle_integrate_pred(const of4* posCurr, const of4* vel, of4* posPred,
                           float gx, float gy, float gz, float dt, uint32_t N);
void oracle_boundary(of4* posPred, of4* vel, const OGridBounds* grid, float restitution, uint32_t N);
void oracle_velocity(of4* vel, const of4* pos, const of4* posPred, float invDt, uint32_t N);

/* --- grid build (reference: hash_keys.cu:6-27, sort_pairs.cu:18-27,
 *     cell_ranges.cu:21-80, cell_compact.cu:8-36,96-231) --- */
void oracle_hash_keys(uint32_t* keys, uint32_t* idx, const of4* posPred, const OGridBounds* grid, uint32_t N);
void oracle_sort_pairs(const uint32_t* keysIn, uint32_t* keysOut, const uint32_t* valsIn, uint32_t* valsOut,
                       uint32_t N);
void oracle_cell_ranges(uint32_t* cellStart, uint32_t* cellEnd, const uint32_t* keysSorted, uint32_t N,
                        uint32_t numCells);
void oracle_cell_compact(uint32_t* uniqueKeys, uint32_t* offsets, uint32_t* count, const uint32_t* keysSorted,
                         uint32_t N);
/* lower_bound range lookup on the compact arrays (reference: cuda_grid_utils.cuh:34-67).
 * Returns 1 and [beg,end) when the cell is present and non-empty. */
int oracle_compact_cell_range(const uint32_t* uniqueKeys, const uint32_t* offsets, uint32_t M, uint32_t key,
                              uint32_t* beg, uint32_t* end);

/* --- neighbour solves (reference: pbf_lambda.cu:10-108, pbf_delta_apply.cu:10-113,223-235,
 *     xsph.cu:9-96).  *_dense take cellStart/cellEnd; *_compact take uniqueKeys/offsets/count. --- */
void oracle_lambda_dense(float* lambda, const of4* posPred, const uint32_t* idxS, const uint32_t* keyS,
                         const uint32_t* cellStart, const uint32_t* cellEnd, const ODeviceParams* dp, uint32_t N);
void oracle_lambda_compact(float* lambda, const of4* posPred, const uint32_t* idxS, const uint32_t* keyS,
                           const uint32_t* uniq, const uint32_t* offsets, const uint32_t* count,
                           const ODeviceParams* dp, uint32_t N);
void oracle_delta_apply_dense(of4* posPred, of4* delta, const float* lambda, const uint32_t* idxS,
                              const uint32_t* keyS, const uint32_t* cellStart, const uint32_t* cellEnd,
                              const ODeviceParams* dp, uint32_t N);
void oracle_delta_apply_compact(of4* posPred, of4* delta, const float* lambda, const uint32_t* idxS,
                                const uint32_t* keyS, const uint32_t* uniq, const uint32_t* offsets,
                                const uint32_t* count, const ODeviceParams* dp, uint32_t N);
/* XSPH with Jacobi semantics: velIn is a snapshot, velOut must not alias it. */
void oracle_xsph_dense(of4* velOut, const of4* velIn, const of4* posPred, const uint32_t* idxS,
                       const uint32_t* keyS, const uint32_t* cellStart, const uint32_t* cellEnd,
                       const ODeviceParams* dp, uint32_t N);
void oracle_xsph_compact(of4* velOut, const of4* velIn, const of4* posPred, const uint32_t* idxS,
                         const uint32_t* keyS, const uint32_t* uniq, const uint32_t* offsets,
                         const uint32_t* count, const ODeviceParams* dp, uint32_t N);

/* --- whole step (reference: phases.cpp:30-117, post_ops.cpp:25-85, simulator.cpp:144-159) ---
 * X (current positions) and V are advanced in place by `steps` steps; after each step the
 * predicted buffer becomes current (the ping-pong swap).  compact != 0 selects the compact
 * ("hashed") grid path.  Returns 0 on success. */
int oracle_run(of4* X, of4* V, const OSimParams* sp, int compact, int steps);

/* Density-constraint statistics of a state, using the lambda kernel's density sum
 * (neighbour cap included, self excluded).  out[0] = mean(rho/rho0 - 1),
 * out[1] = max(rho/rho0 - 1), out[2] = mean in-range neighbours, out[3] = mean |v|,
 * out[4] = mean(max(rho/rho0 - 1, 0)). */
void oracle_density_stats(const of4* X, const of4* V, const OSimParams* sp, double* out);

/* --- scene seeding (reference: simulator.cpp:808-878) ---
 * Lattice x fastest, then y, then z; origin + i*spacing; clamped to the box; then a
 * per-axis jitter (2u-1)*jitterAmp with u from SplitMix64(seed, particle, axis); clamped again.
 * w = 1.  Velocities are not touched. */
void oracle_seed_lattice(of4* X, uint32_t nx, uint32_t ny, uint32_t nz, float ox, float oy, float oz,
                         float spacing, const OGridBounds* grid, float jitterAmp, uint64_t seed);
/* The uniform variate used by the seeder: 24-bit mantissa float in [0,1). */
float oracle_jitter_uniform(uint64_t seed, uint64_t particle, uint32_t axis);

/* The s_corr power (W/W(dq))^n as the solver evaluates it. */
float oracle_pow_scorr(float x, float n);

int oracle_num_threads(void);
/* Threads the OpenMP loops use from now on (n >= 1). */
void oracle_set_num_threads(int n);

#ifdef __cplusplus
}
#endif

#endif /* PBF_ORACLE_H */
