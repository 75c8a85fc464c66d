/*
 * dsim: MI355X-native diffusion-MRI Monte Carlo simulate path (C ABI).
 *
 * This header is the drop-in boundary that replaces the reference's
 * `extern "C"` CUDA host glue (dSim3/src/spinSystem.cuh:1-59, defined in
 * dSim3/src/spinSystem.cu:18-299) and the SpinSystem methods that feed it.
 * Every entry point names the reference interface it replaces.
 *
 * Conventions
 *   - Plain C types only: pointers, sizes, PODs.  No device pointers cross
 *     the boundary unless a function says so.
 *   - Every function returns a status: DSIM_OK (0) or a negative DSIM_ERR_*;
 *     dsim_last_error() returns a message for the calling thread.  Nothing
 *     calls exit() (the reference's CUDA_SAFE_CALL does, spinSystem.cu:38-98).
 *   - Units at the boundary are the reference's: coordinates normalised to
 *     [-1,1] (micrometres / spaceScale, spinSystem.cpp:422-424), gradients as
 *     gradScaled = G[mT/m] * 1e-9 * spaceScale (spinSystem.cpp:1931) and
 *     phaseConstant = 2*pi*gamma*dt (spinSystem.cpp:1938).
 *   - A dsim_sim owns its device buffers and one HIP stream; calls on
 *     different handles may run from different host threads.  Calls on one
 *     handle must not overlap.
 */
#ifndef DSIM_H
#define DSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DSIM_ABI_VERSION 1

#define DSIM_OK 0
#define DSIM_ERR_ARG (-1)      /* bad argument or size */
#define DSIM_ERR_HIP (-2)      /* HIP runtime error (message has the HIP string) */
#define DSIM_ERR_NOMEM (-3)    /* device or host allocation failed */
#define DSIM_ERR_STATE (-4)    /* call not valid in the handle's current mode */
#define DSIM_ERR_OVERFLOW (-5) /* a device-side bound was exceeded (flagged, not faulted) */
#define DSIM_ERR_IO (-6)       /* file could not be opened */
#define DSIM_ERR_FORMAT (-7)   /* input file rejected (same rules as the reference) */
#define DSIM_ERR_NODEV (-8)    /* no HIP device visible */

#define DSIM_NO_TRIANGLE 0xFFFFFFFFu
#define DSIM_NO_FIBER 65535u

/* spinData, dSim3/src/dSimDataTypes.h:9-24 (12 bytes with padding). */
typedef struct {
    float signalMagnitude;
    float signalPhase;
    uint8_t compartmentType; /* 0 extra, 1 intra-axonal, 2 myelin, 3 glia interior */
    uint16_t insideFiber;    /* DSIM_NO_FIBER when outside every fibre */
} dsim_spin_data;

/* Mesh arrays as SpinSystem::initFibers leaves them
 * (m_vertices, m_trgls, m_hTriInfo; spinSystem.cpp:290-560). */
typedef struct {
    uint32_t n_vertices;
    const float *vertices;    /* [n_vertices*3], normalised */
    uint32_t n_triangles;
    const uint32_t *triangles; /* [n_triangles*3], 0-based vertex indices */
    const uint32_t *tri_info;  /* [n_triangles*3]: fibre, membrane (0 A, 1 M, 2 G), first vertex */
} dsim_mesh;

enum {
    DSIM_GRID_REFERENCE = 0, /* n^3 cubes over [-1,1]^3, reference membership (populateCubes) */
    DSIM_GRID_AABB = 1,      /* any box, triangle-AABB membership (conservative) */
    DSIM_GRID_TRIBOX = 2     /* any box, exact triangle/box membership within eps (separating
                              * axes): for meshes whose long triangles do not run along a
                              * grid axis; entries keep only the triangle index */
};

typedef struct {
    uint32_t kind;
    uint32_t dims[3]; /* REFERENCE: dims[0]=n (the reference hard-codes 50); AABB: cells per axis */
    float lo[3];      /* AABB: grid origin (normalised units) */
    float cell[3];    /* AABB: cell edge per axis */
    float eps;        /* AABB: membership margin */
    uint32_t clearance_rings; /* AABB and exact grids: cells at ring distance at most
                               * clearance_rings - 2 from a listed triangle refine their
                               * clearance with exact distances, farther ones use the
                               * ring distance alone (1: none refine; 0: the default,
                               * rings 0-3 for AABB grids and rings 0-1 for exact
                               * grids) */
} dsim_grid_params;

enum {
    DSIM_MODE_REFERENCE = 0, /* the reference's CPU semantics (spinKernelCpu.cpp) */
    DSIM_MODE_PHYSICAL = 1   /* Moller-Trumbore, nearest hit, physical compartment table */
};
enum { DSIM_BOUNDARY_CLAMP = 0, DSIM_BOUNDARY_PERIODIC = 1 };
enum { DSIM_RNG_MWC = 1, DSIM_RNG_PHILOX = 2 };

typedef struct {
    float dt;                 /* ms (sim.timeStep) */
    float std_dev[4];         /* per compartment, normalised (spinSystem.cpp:52-56) */
    float t2[4];              /* ms per compartment, INFINITY = no decay (spinSystem.cpp:47-51) */
    float permeability;       /* transmission probability per hit (sim.permeability) */
    uint32_t reflection_type; /* 1 = specular (diffusion.cpp:96), 0 = back to start */
    uint32_t mode;            /* DSIM_MODE_* */
    uint32_t boundary;        /* DSIM_BOUNDARY_*; the reference mode always clamps */
    uint32_t rng;             /* DSIM_RNG_*; MWC is the reference's per-spin stream (its seed
                               * pairs start as a draw keyed by `seed`, both words nonzero;
                               * dsim_set_state_aos sets the reference's own) */
    uint64_t seed;            /* Philox key */
    float spin_radius;        /* clamp margin (spinSystem.cpp:61: 0.005) */
    uint32_t sort_interval;   /* physical mode with Philox: reorder spins in device memory
                               * every this many steps (0 = never), by recent membrane-hit
                               * rate, then cell.  Results do not depend on it. */
    uint32_t steps_per_launch; /* steps fused per kernel launch (0 = 1024; with reordering at
                                  most sort_interval) */
    float voxel_lo[3];        /* periodic voxel [voxel_lo, voxel_hi) per axis, normalised;     */
    float voxel_hi[3];        /* all six zero means [-1, 1) on every axis (the cubic voxel)    */
    uint32_t fixed_blocks;    /* physical mode with Philox: 0 lets a persistent walk take over
                               * once every resident workgroup has two or more spins (each lane
                               * takes its next spin as soon as one is done); 1 keeps one
                               * 64-spin workgroup per block.  Results do not depend on it. */
} dsim_params;

typedef struct dsim_substrate dsim_substrate;
typedef struct dsim_sim dsim_sim;

/* ---- library ------------------------------------------------------------ */
const char *dsim_last_error(void);
int dsim_abi_version(void);
/* Number of visible HIP devices (0 on a host without a GPU). */
int dsim_device_count(int *n);

/* Device-memory support entries of the reference's C ABI
 * (dSim3/src/spinSystem.cuh:6-13, defined in spinSystem.cu:35-88), for
 * callers that keep their own device arrays (as SpinSystem does): */
/* allocateArray (spinSystem.cu:35-38): device allocation of `bytes`. */
int dsim_alloc_array(void **dev_ptr, size_t bytes);
/* freeArray (spinSystem.cu:46-49). */
int dsim_free_array(void *dev_ptr);
/* threadSync (spinSystem.cu:57-60): waits for all work on the device. */
int dsim_thread_sync(void);
/* copyArrayFromDevice (spinSystem.cu:68-75) without the GL buffer mapping. */
int dsim_copy_from_device(void *host, const void *device, size_t bytes);
/* copyArrayToDevice (spinSystem.cu:83-86): device + offset <- host. */
int dsim_copy_to_device(void *device, const void *host, size_t offset, size_t bytes);

/* ---- substrate ---------------------------------------------------------- */
/* Replaces the device half of SpinSystem::initFibers (triangle helpers and
 * their texture binds, spinSystem.cpp:612-665) and SpinSystem::populateCubes
 * with its cube-array copies (spinSystem.cpp:1094-1352; bindVertices,
 * bindTriangleHelpers, bindTriInfo, bindCubeCounter, bindTrianglesInCubes in
 * spinSystem.cu:127-215).  The grid is stored as CSR (cell_start, cell_tri)
 * in place of the reference's dense padded table; lists keep the reference's
 * ascending triangle order. */
int dsim_substrate_create(const dsim_mesh *mesh, const dsim_grid_params *grid,
                          dsim_substrate **out);
/* The same substrate built on the device: triangle records, the grid (the
 * reference's populateCubes membership, spinSystem.cpp:1094-1352, or the AABB
 * membership), its cell-major copies and the AABB clearance are computed by
 * GPU kernels with the host build's arithmetic, so every array equals
 * dsim_substrate_create's.  Needs a HIP device; the host keeps no copy of the
 * grid (the getters below read it back). */
int dsim_substrate_create_device(const dsim_mesh *mesh, const dsim_grid_params *grid,
                                 dsim_substrate **out);
int dsim_substrate_destroy(dsim_substrate *s);
/* Grid sizes: cells, list entries, and the reference's m_maxTrianglesPerCube. */
int dsim_substrate_grid_info(const dsim_substrate *s, uint32_t *n_cells, uint32_t *n_entries,
                             uint32_t *max_per_cell);
/* Copies the CSR grid to host arrays (cell_start[n_cells+1], cell_tri[n_entries]). */
int dsim_substrate_get_grid(const dsim_substrate *s, uint32_t *cell_start, uint32_t *cell_tri);
/* Copies m_triangleHelpers (12 floats per triangle, spinSystem.cpp:612-648). */
int dsim_substrate_get_helpers(const dsim_substrate *s, float *helpers);
/* AABB grids only (no reference counterpart): per-cell lower bound of the
 * distance from any point of the cell to any triangle registered in the grid
 * [n_cells] (1e30 when there is none).  The physical step skips collision
 * detection while a spin that was inside the grid at its last check has
 * walked a path shorter than this clearance. */
int dsim_substrate_get_clearance(const dsim_substrate *s, float *clearance);
/* No reference counterpart: the axis of the physical mode's labelling rays
 * (0 +x, 1 +y, 2 +z), chosen from the mesh so that the rays run across its
 * long walls (findSpinsInFibers casts along the spin's dominant axis,
 * spinSystem.cpp:1677-1683, which misses an open tube along that axis and
 * can leave a cut tube through its end): the axis along which the triangles'
 * extents sum least, +x unless that sum exceeds the smaller of the others by
 * more than 5%. */
int dsim_substrate_label_axis(const dsim_substrate *s, uint32_t *axis);
/* Device bytes held by the substrate. */
int dsim_substrate_bytes(const dsim_substrate *s, uint64_t *bytes);

/* ---- simulation handle --------------------------------------------------- */
/* Replaces SpinSystem::build's device allocations and constant uploads
 * (spinSystem.cpp:178-281: allocateArray, copyConstantToDevice). */
int dsim_sim_create(dsim_substrate *s, const dsim_params *p, uint32_t n_spins, dsim_sim **out);
int dsim_sim_destroy(dsim_sim *sim);
/* setArray + setSpinArray (spinSystem.cpp:1584-1606): host AoS positions
 * [n*3], spinData [n], and MWC seeds [n*2] (may be NULL unless rng==MWC). */
int dsim_set_state_aos(dsim_sim *sim, const float *pos, const dsim_spin_data *spins,
                       const uint32_t *seeds);
/* getArray + getSpinArray (spinSystem.cpp:1544-1571); any pointer may be NULL. */
int dsim_get_state_aos(dsim_sim *sim, float *pos, dsim_spin_data *spins, uint32_t *seeds);
/* Sets the Philox step counter (the next integrated step uses this value). */
int dsim_set_step(dsim_sim *sim, uint64_t step);
int dsim_get_step(dsim_sim *sim, uint64_t *step);

/* resetSpins' position draw (spinSystem.cpp:1386-1398), on the device:
 * positions uniform in [-start_box, start_box]^3 from the Philox stream,
 * magnitude 1, phase 0, labels cleared.  The draw is keyed by the handle's
 * seed and its step counter (dsim_set_step), so draws at different counter
 * values differ.  Follow with dsim_label_spins. */
int dsim_init_spins_uniform(dsim_sim *sim, float start_box);
/* The same draw over a box [lo, hi] per axis (a non-cubic periodic voxel):
 * coordinate = (hi - lo) * (u - 0.5) + (hi + lo) / 2, u from the same Philox
 * stream, so the cube [-b, b]^3 gives dsim_init_spins_uniform(sim, b). */
int dsim_init_spins_box(dsim_sim *sim, const float lo[3], const float hi[3]);

/* findSpinsInFibers (spinSystem.cpp:1615-1852), on the device.  The reference
 * mode casts its ray along the spin's dominant axis to 1.1*max (as the
 * reference does); the physical mode casts it along +x so that open prisms
 * are counted.  Ambiguous rays (edge or vertex hits, or an odd crossing count
 * on two fibres of one membrane type) are re-cast with a reproducible jitter
 * up to max_tries times; spins still ambiguous keep their last label and are
 * counted in *n_ambiguous. */
int dsim_label_spins(dsim_sim *sim, float xmax, float ymax, float zmax, uint32_t max_tries,
                     uint32_t *n_ambiguous);

/* integrateSystem (spinSystem.cuh:34-45, spinSystem.cu:221-270): advance all
 * spins `iterations` steps under a constant gradient, phase += g.pos *
 * phaseConstant per step with pos the position after the step (in a periodic
 * voxel the unwrapped position: pos + crossings * period per axis), state
 * left on the device. */
int dsim_integrate(dsim_sim *sim, const float grad_scaled[3], float phase_constant,
                   uint32_t iterations);
/* Test entry: as dsim_integrate but the normals and uniform of every step come
 * from n3u (host [iterations][n][4]) and, if trace is not NULL, the per-step
 * outcome is written to trace (host [iterations][n][5]: first hit triangle,
 * last hit triangle, hits, cell of the end position, compartment|fibre<<8). */
int dsim_integrate_injected(dsim_sim *sim, const float grad_scaled[3], float phase_constant,
                            const float *n3u, uint32_t iterations, uint32_t *trace);

/* getMrSignal(double[]) (spinSystem.cpp:1452-1500): out[0] total, out[1+c]
 * compartment c, |sum m e^{i phase}| / count, FP64 sum. */
int dsim_signal(dsim_sim *sim, double out[5]);

/* ---- shared-walk gradient moments (one walk serves every direction) ------ */
/* Starts an acquisition (the per-direction resetSpins of diffusion.cpp:565
 * without moving the spins): records x0, zeroes the displacement and moment
 * accumulators and sets every magnitude to 1.  Positions and labels stay. */
int dsim_moment_reset(dsim_sim *sim);
/* Advance `iterations` steps accumulating M += weight * (x - x0) per step,
 * weight in {+1, 0, -1} for the PGSE lobes of diffusion.cpp:564-585. */
int dsim_integrate_moment(dsim_sim *sim, float weight, uint32_t iterations);
/* Same, with injected normals/uniforms and optional trace (as above). */
int dsim_integrate_moment_injected(dsim_sim *sim, float weight, const float *n3u,
                                   uint32_t iterations, uint32_t *trace);
/* Signal per direction j: phase = phase_constant * g_j . (M + W x0), FP64
 * sum; out[j*5 + k] as in dsim_signal.  grads_scaled is [ndir*3]. */
int dsim_signal_moment(dsim_sim *sim, uint32_t ndir, const float *grads_scaled,
                       float phase_constant, double *out);
/* Host copy of the moment accumulators [n*3] and x0 [n*3] (either may be NULL). */
int dsim_get_moment(dsim_sim *sim, float *moment, float *x0);

/* ---- allocation-free readout ------------------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): dsim_signal_moment over device arrays the caller owns, with its
 * partial sums in a fixed device array of the library (7.9 MB per GPU, kept
 * until the process ends) instead of allocated memory.  Inputs are device
 * pointers in the handle's field-major layout: mag [n], moment [3][n], x0
 * [3][n] (may be NULL when wsum is 0), label [n] (compartment | fibre << 8),
 * grads_scaled [ndir*3]; out is a device array [ndir*5] laid out as
 * dsim_signal_moment's.  Work is queued on `stream` (a hipStream_t, NULL for
 * the null stream) and not waited for.  At most DSIM_NALLOC_MAX_DIRECTIONS
 * directions per call (DSIM_ERR_ARG beyond).  Every call on one GPU shares
 * the array: calls that overlap on one GPU race and must be serialised by
 * the caller (one stream, or events). */
#define DSIM_NALLOC_MAX_DIRECTIONS 256
int dsim_signal_moment_nalloc(uint32_t n, const float *mag, const float *moment, const float *x0,
                              float wsum, const uint32_t *label, uint32_t ndir,
                              const float *grads_scaled, float phase_constant, double *out,
                              void *stream);

/* ---- timing ------------------------------------------------------------- */
int dsim_sync(dsim_sim *sim);
/* Milliseconds spent in step-kernel launches since the last reset, measured
 * with a HIP event pair around each launch on the handle's stream (spin
 * reorders between launches are not included); launches counts them. */
int dsim_step_timing(dsim_sim *sim, double *ms, uint64_t *launches);
int dsim_step_timing_reset(dsim_sim *sim);

/* No reference counterpart: work counters of the physical-mode walk.
 * While counting is on, the integrate calls run a counting build of the walk
 * (same results, slower); the counts add up over calls until counting is
 * switched on again (which zeroes them).  Counting on a handle in the
 * reference mode or with MWC draws fails with DSIM_ERR_ARG at the integrate
 * call; Philox and injected draws are counted. */
#define DSIM_WALK_STATS 16
enum {
    DSIM_STAT_STEPS = 0,         /* spin-steps completed */
    DSIM_STAT_TRIPS = 1,         /* trips round a wave's loop (one search round each) */
    DSIM_STAT_SEARCH_LANES = 2,  /* lanes searching, summed over trips */
    DSIM_STAT_CELLS = 3,         /* grid cells listed by searching lanes */
    DSIM_STAT_PLANE_TESTS = 4,   /* (segment, grid entry) plane tests */
    DSIM_STAT_PLANE_PASSES = 5,  /* wave passes of the plane test */
    DSIM_STAT_TRI_TESTS = 6,     /* Moller-Trumbore tests (plane crossings) */
    DSIM_STAT_TRI_PASSES = 7,    /* wave passes of the triangle test */
    DSIM_STAT_FP64_TESTS = 8,    /* near-grazing FP64 re-tests */
    DSIM_STAT_HITS = 9,          /* membrane hits */
    DSIM_STAT_SHORTCUTS = 10,    /* steps completed by the clearance shortcut */
    DSIM_STAT_DEFERRED = 11,     /* lanes whose cell list waited a round, summed over trips */
    DSIM_STAT_WAVES = 12,        /* waves run */
    DSIM_STAT_LIVE_LANES = 13    /* lanes with steps left, summed over trips */
};
int dsim_walk_stats_enable(dsim_sim *sim, int enable);
int dsim_walk_stats_read(dsim_sim *sim, uint64_t counts[DSIM_WALK_STATS]);

#ifdef __cplusplus
}
#endif
#endif
