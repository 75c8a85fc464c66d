// Internal types shared by the host runtime (host.cpp) and the HIP kernels
// (kernels.hip).  Plain structs only: the kernels take them by value.
#pragma once

#include <cstdint>

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

namespace dsim {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kFiberNone = 65535u;

// One 64-byte record per triangle for each collision test.
//   reference record: h0..h11 of spinSystem.cpp:612-648, then the first
//                     vertex P1 (x,y,z) and info (fibre | membrane<<16)
//   MT record:        n.x n.y n.z n.v0 | v0.x v0.y v0.z info | e1 - | e2 -
// Labelling uses 32-byte AABB records: min.xyz info | max.xyz 0.
// Spins per tile of the trajectory buffer (StepParams::traj): each tile's rows
// are contiguous, so a phase GEMM spin tile reads one block; within a tile
// the k (3 * step + axis) of a spin come in 16-byte chunks of kTrajChunk.
constexpr uint32_t kTrajTile = 128;
constexpr uint32_t kTrajChunk = 4;

struct StepParams {
    // substrate (device pointers)
    const float4 *tri_ref;
    const float4 *tri_mt;
    const uint32_t *cell_start;
    const uint32_t *cell_tri;
    const float4 *cell_plane;  // plane (n, n.v0) of cell_tri[k], cell-major
    const float4 *cell_mt;     // [3] per entry: v0 (w: triangle index), e1, e2
    uint32_t n_entries;        // length of cell_tri and cell_plane
    const float *cell_clear;   // per-cell distance lower bound, or null (no skipping)
    const uint8_t *fiber_myelin;
    uint32_t n_fibers;
    uint32_t dims[3];
    float lo[3];
    float cell[3];
    float inv_cell[3];
    // physics
    float std_dev[4];
    float t2[4];
    float decay[4];
    float dt;
    float perm;
    uint32_t reflection_type;
    float clamp_hi, clamp_lo;
    // periodic voxel [wrap_lo, wrap_hi) per axis, period wrap_len = hi - lo
    float wrap_lo[3], wrap_hi[3], wrap_len[3];
    // spin state (device pointers, SoA)
    uint32_t n;
    float *x, *y, *z, *mag, *phase;
    uint32_t *label;  // compartment | fibre << 8
    const uint32_t *spin_id;  // identity of the spin held by each slot (Philox counter)
    uint32_t *seed;   // MWC, 2 per spin
    float *disp;      // [3][n] unwrapped displacement
    float *mom;       // [3][n] moment
    float *momc;      // [3][n] Kahan compensation
    uint32_t *hits;   // membrane hits per spin since the last reorder, or null
    uint32_t *wrap;   // periodic voxel crossings per spin (three signed 10-bit counts), used
                      // for the phase of the unwrapped position when moments are off
    uint32_t hit_window;  // steps those counts cover (0: no history)
    float *traj;          // displacement after each step of the launch by spin id, in
                          // tiles of kTrajTile spins: k = 3 * step + axis of spin s at
                          // [s / kTrajTile][k / kTrajChunk][s % kTrajTile][k % kTrajChunk]
                          // (moment mode), or null
    uint32_t traj_stride;  // floats per tile: 3 * iters * kTrajTile
    // per call
    float gx, gy, gz, pc, weight;
    uint32_t key0, key1;
    uint64_t step0;
    const float4 *inject;  // [iters][n]
    uint32_t *trace;       // [iters][n][5]
    uint32_t *flags;       // status word: bit 0 collision loop guard hit, bit 2 entry out of range
    unsigned long long *stats;  // [kWalkStats] work counters (counting build of the walk)
    uint32_t *queue;       // [8] slot counters of the persistent walk, or null (fixed blocks)
    uint32_t queue_run;    // slots per queue run (persistent walk; set at launch)
};

struct LabelParams {
    const float4 *tri_ref;
    const float4 *tri_mt;
    uint32_t use_mt;  // 1: Moller-Trumbore crossing test (physical mode)
    const float4 *tri_aabb;
    uint32_t n_tri;
    uint32_t n;
    const float *x, *y, *z;
    uint32_t *label;
    uint32_t *tries;
    float xmax, ymax, zmax;
    uint32_t max_tries;
    uint32_t ray_axis;  // 1, 2, 3: ray along +x, +y, +z; 0: dominant axis (reference)
    // First ray through an AABB or exact-membership grid (launch_label_row):
    // the row of cells the ray runs along gives the candidates of the
    // triangles lying inside the grid box [lo, hi); `outside` lists the others
    // (n_outside on the device).  Spins left for the tiled pass go to `retry`
    // as (spin, first try), counted in *n_retry.
    const uint32_t *cell_start, *cell_tri;
    uint32_t dims[3];
    float lo[3], hi[3], cell[3], inv_cell[3];
    uint32_t exact_grid;  // 1: exact-membership grid (outside list: boxes reaching past the face)
    uint32_t *outside, *n_outside;
    uint2 *retry;
    uint32_t *n_retry;
};

enum Rng { kRngInject = 0, kRngMwc = 1, kRngPhilox = 2 };

// What a step launch records beside the step: nothing, the per-step trace,
// or the walk's work counters (the DSIM_STAT_* indices of dsim.h).
enum Instr { kInstrNone = 0, kInstrTrace = 1, kInstrStats = 2 };
constexpr uint32_t kWalkStats = 16;
enum WalkStat {
    kStatSteps = 0, kStatTrips, kStatSearchLanes, kStatCells, kStatPlane, kStatPasses,
    kStatTriangle, kStatTriPasses, kStatGraze, kStatHits, kStatShortcut, kStatDeferred,
    kStatWaves, kStatLiveLaneTrips
};

// Job table of the batched walk (physical mode, Philox, moment accumulation):
// job j advances the simulation of job[j].p by job[j].iters steps on blocks
// [block_start[j], block_start[j+1]).
constexpr uint32_t kMaxJobs = 256;
// Spins per workgroup of the physical-mode step kernels (k_walk and
// k_walk_batch): one wave.
constexpr uint32_t kWalkBlockSpins = 64;

struct WalkJob {
    StepParams p;
    uint32_t iters;
};
struct WalkTable {
    uint32_t n_jobs;
    uint32_t total_blocks;
    uint32_t block_start[kMaxJobs + 1];
    WalkJob job[kMaxJobs];
};

// Launchers (kernels.hip).  Return hipError_t of the launch.
hipError_t launch_step(const StepParams &p, uint32_t iters, int mode, int periodic, int rng,
                       int instr, bool moment, hipStream_t s);
hipError_t launch_signal(uint32_t n, const float *mag, const float *phase, const uint32_t *label,
                         double *partial, uint32_t n_blocks, double *out15, hipStream_t s);
hipError_t launch_moment_signal(uint32_t n, const float *mag, const float *mom, const float *x0,
                                float wsum, const uint32_t *label, uint32_t ndir,
                                const float *grads, float pc, double *partial, uint32_t n_blocks,
                                double *out, hipStream_t s);
hipError_t launch_label(const LabelParams &p, hipStream_t s);
// Grid-row labelling of the first +x ray, then the tiled pass over the spins
// it leaves (ambiguous rays, spins outside the grid).  The buffers of the
// grid fields of p must be set; *n_outside and *n_retry must be zero.
hipError_t launch_label_row(const LabelParams &p, hipStream_t s);
// Batched physical-mode walk over a device-resident job table (periodic or
// clamped boundary for every job of the launch).  recs: where the jobs' grid
// entries keep their records (kRecsCellMajor: every job's grid has cell-major
// copies; kRecsIndirect: none has; kRecsEither: mixed, decided per job).
enum { kRecsEither = 0, kRecsCellMajor = 1, kRecsIndirect = 2 };
hipError_t launch_walk_batch(const WalkTable *table, uint32_t grid_blocks, bool periodic,
                             int recs, hipStream_t s);
// MWC seed pairs of n spins from Philox stream 3 (both words nonzero).
hipError_t launch_init_seeds(uint32_t n, uint32_t *seed, uint32_t key0, uint32_t key1,
                             hipStream_t s);
hipError_t launch_init_uniform(uint32_t n, float *x, float *y, float *z, float *mag, float *phase,
                               uint32_t *label, uint32_t *id, uint32_t *wrap, const float lo[3],
                               const float hi[3], uint32_t key0, uint32_t key1, uint64_t draw,
                               hipStream_t s);
hipError_t launch_moment_reset(uint32_t n, const float *x, const float *y, const float *z,
                               float *x0, float *disp, float *mom, float *momc, float *mag,
                               hipStream_t s);

constexpr uint32_t kSignalBlocks = 1024;  // partial-sum blocks of the readout kernels

// Phase of arbitrary gradient waveforms (see kernels.hip, k_phase_gemm):
// phi[s][w] += sum_k traj[k][s] * G[w][k] over one chunk of K = 3*steps,
// with G given as bf16 hi and lo parts [n_wave][k_total] (k contiguous),
// phase[s][w] in FP32 ([n_pad][n_wave]) and its Kahan compensation as bf16,
// spins 2m and 2m + 1 in the low and high half of comp[m][w]
// ([n_pad / 2][n_wave]).  traj in the tiled layout of
// StepParams::traj with k_len rows (tiles of k_len * kTrajTile floats),
// spins padded to kPhaseTileM.
constexpr uint32_t kPhaseTileM = 128;  // spins per workgroup tile
static_assert(kPhaseTileM == kTrajTile, "one trajectory tile per GEMM spin tile");
constexpr uint32_t kPhaseTileN = 128;  // waveforms per workgroup tile
constexpr uint32_t kPhaseSliceK = 32;  // k per LDS stage
// k_len trajectory rows in Kahan folds of k_fold rows (both multiples of
// kPhaseSliceK; the last fold may be shorter).
hipError_t launch_phase_gemm(const float *traj, uint32_t n, uint32_t n_pad, uint32_t k_len,
                             uint32_t k_fold, const uint16_t *g_hi, const uint16_t *g_lo,
                             uint32_t k_total, uint32_t k0, uint32_t n_wave, float *phase,
                             uint32_t *comp, hipStream_t s);
// Signal of every waveform: per block of state slots and waveform the FP64
// sums of m cos(phase), m sin(phase) and counts for the total and
// compartments 0-3, phase = phase[id[s]][w] + gsum[w] . x0[s] (in FP64; mag,
// label, x0 and id in slot order, phase by spin id); then a fixed-order
// final pass into out[n_wave][15].
hipError_t launch_phase_signal(uint32_t n, const float *phase, uint32_t n_wave,
                               const double *gsum, const float *mag, const uint32_t *label,
                               const float *x0, const uint32_t *id, double *partial,
                               uint32_t n_blocks, double *out, hipStream_t s);

// Spatial reordering of spins: Morton key of each slot's cell, radix sort of
// (key, slot), then a gather of every per-spin field into its new slot.
constexpr uint32_t kMaxFields = 24;
struct Fields {
    uint32_t *src[kMaxFields];
    uint32_t *dst[kMaxFields];
    uint32_t count;
};
hipError_t sort_temp_bytes(uint32_t n, size_t *bytes);
// Sort keys of the spins (hit band, then cell Morton index) and the slot
// permutation that orders them, stably; *perm points at the permutation (one
// of the two value buffers).
hipError_t launch_sort_perm(const StepParams &p, uint32_t *keys_a, uint32_t *keys_b,
                            uint32_t *vals_a, uint32_t *vals_b, void *temp, size_t temp_bytes,
                            hipStream_t s, const uint32_t **perm);
hipError_t launch_gather(uint32_t n, const uint32_t *perm, const Fields &f, hipStream_t s);

}  // namespace dsim
