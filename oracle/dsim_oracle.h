/*
 * dsim oracle: plain-C restatement of dSim3's CPU simulate path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP product
 * path (tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg are
 * the only users).  The product library never links, loads or calls it.
 *
 * Each function cites the reference lines it restates (paths relative to the
 * reference repository root).  Functions marked "build-defined" have no
 * counterpart in the reference; they restate the production semantics of the
 * HIP path (periodic voxel, Moller-Trumbore test, nearest-hit traversal,
 * Philox4x32-10 stream, gradient-moment accumulation) so the GPU can be
 * checked bit-exactly against them as well.
 *
 * Pins (how each function is checked independently of this file):
 *   orc_mwc_next ............ closed-form integer recurrence, hand vectors
 *   orc_philox4x32_10 ....... Random123 known-answer vectors
 *   orc_tri_coll_detect ..... collisions01.m known answer + closed-form
 *                             plane/barycentric cases + exact-rational brute force
 *   orc_reflect_pos ......... Householder invariants (involution, isometry)
 *   orc_cube_pos/hash ....... closed form
 *   orc_populate_cubes ...... brute-force triangle/box overlap (exact rationals)
 *   orc_coll_detect_grid .... hand-built multi-hit cases; grid-independence of
 *                             the nearest-hit mode vs brute force over all tris
 *   orc_integrate ........... invariants (impermeable closed mesh keeps labels,
 *                             free diffusion variance, T2 closed form); the
 *                             per-step phase by a one-step known answer
 *                             (post-step position) and by PGSE against
 *                             exp(-bD) in the clamp box and the periodic voxel
 *   orc_find_spins .......... analytic point-in-prism on the sample meshes
 *   orc_mr_signal ........... closed form
 *   orc_parse_mesh .......... the reference's sample mesh files
 *   orc_pgse_schedule ....... closed-form timing
 */
#ifndef DSIM_ORACLE_H
#define DSIM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NONE 0xFFFFFFFFu
#define ORC_FIBER_NONE 65535u

typedef struct { float x, y, z; } orc_f3;

/* collResult, reference dSim3/src/spinSystem.h:36-43 */
typedef struct {
    uint32_t type;      /* 0 none, 1 inside, 2 edge, 3 vertex */
    orc_f3 pt;
    uint32_t idx;
    float dist_sq;      /* softsurfer: squared distance; MT: parametric t */
} orc_coll;

/* Everything the step needs.  Mirrors the SpinSystem members read by
 * dSim3/src/spinKernelCpu.cpp (m_triangleHelpers, m_vertices, m_hTriInfo,
 * m_hCubeCounter, m_hTrianglesInCubes, m_hStdDevs, m_hT2Values, ...). */
typedef struct {
    uint32_t n_tri;
    const float *helpers;       /* [n_tri*12] spinSystem.cpp:612-648 layout */
    const float *vertices;      /* [n_vert*3] */
    const uint32_t *tri_info;   /* [n_tri*3] fiber, membrane, first vertex */
    const uint32_t *trgls;      /* [n_tri*3] vertex indices (labelling AABBs) */
    const float *tri_mt;        /* [n_tri*16] n.xyz, n.v0, v0.xyz, -, e1.xyz, -, e2.xyz, - (MT test) */
    /* dense padded grid, spinSystem.cpp:1094-1352 layout */
    uint32_t dims[3];
    float lo[3];
    float cell[3];              /* cell edge per axis */
    float inv_cell[3];          /* production cell math uses reciprocal */
    uint32_t ref_cell_math;     /* 1: floor((p-lo)/cell) (reference), 0: floor((p-lo)*inv_cell) */
    uint32_t max_per_cell;
    const uint32_t *cell_count; /* [ncells] */
    const uint32_t *cell_list;  /* [ncells*max_per_cell] */
    /* physics */
    float std_dev[4];
    float t2[4];
    float decay[4];             /* production: dt/T2 per compartment */
    float dt;
    float permeability;
    uint32_t reflection_type;   /* diffusion.cpp:96 hard-codes 1 */
    float spin_radius;          /* clamp margin, spinSystem.cpp:61 */
    uint32_t boundary;          /* 0 clamp (reference), 1 periodic [-1,1) */
    uint32_t test;              /* 0 softsurfer (reference), 1 Moller-Trumbore */
    uint32_t traversal;         /* 0 first cell x->y->z (reference), 1 nearest hit */
    uint32_t transitions;       /* 0 reference table, 1 physical table */
    const uint8_t *fiber_has_myelin;
    uint32_t n_fibers;
    /* build-defined periodic voxel [wrap_lo, wrap_hi) per axis (boundary 1);
     * all zero means [-1, 1) on every axis */
    float wrap_lo[3], wrap_hi[3];
} orc_world;

/* Per-spin state, host layout of the reference: pos AoS xyz (m_hPos),
 * spinData fields split out (dSimDataTypes.h:9-24), seeds AoS (m_hSeed). */
typedef struct {
    uint32_t n;
    float *pos;         /* [n*3] */
    float *mag;         /* [n] */
    float *phase;       /* [n] or NULL */
    uint8_t *comp;      /* [n] */
    uint16_t *fiber;    /* [n] */
    uint32_t *seed;     /* [n*2] MWC state, or NULL */
    float *disp;        /* [n*3] unwrapped displacement (moment mode) or NULL */
    float *moment;      /* [n*3] sum of w*disp (moment mode) or NULL */
    float *moment_c;    /* [n*3] Kahan compensation or NULL */
    int32_t *wrap;      /* [n*3] periodic voxel crossings (+1 per wrap back from hi), or
                         * NULL; with it the phase uses the unwrapped position */
} orc_state;

enum { ORC_RNG_INJECT = 0, ORC_RNG_MWC = 1, ORC_RNG_PHILOX = 2 };

/* ---- RNG ---------------------------------------------------------------- */
uint32_t orc_mwc_next(uint32_t seed[2]);
void orc_box_muller(float *u1, float *u2);
void orc_randn(uint32_t seed[2], float *n1, float *n2, float *n3, float *u);
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float orc_u32_to_unit(uint32_t x);
void orc_philox_step_normals(const uint32_t key[2], uint32_t spin, uint64_t step,
                             float n[3]);
float orc_philox_step_uniform(const uint32_t key[2], uint32_t spin, uint64_t step);

/* ---- geometry helpers ----------------------------------------------------- */
void orc_tri_helpers(const float *vertices, const uint32_t *trgls, uint32_t n_tri,
                     float *helpers);
void orc_tri_mt_records(const float *vertices, const uint32_t *trgls, uint32_t n_tri,
                        float *rec);
void orc_cube_pos(const orc_world *w, orc_f3 p, uint32_t cpos[3]);
uint32_t orc_cube_hash(const orc_world *w, const uint32_t cpos[3]);
int orc_point_in_cube(orc_f3 p, orc_f3 mn, orc_f3 mx);
int orc_ray_tri_intersect(orc_f3 r1, orc_f3 r2, orc_f3 t1, orc_f3 t2, orc_f3 t3);
/* Reference grid build.  count[n_cells]; list may be NULL on the sizing pass.
 * Returns max_per_cell. */
uint32_t orc_populate_cubes(const float *vertices, const uint32_t *trgls, uint32_t n_tri,
                            uint32_t n_cubes, uint32_t *count, uint32_t *list,
                            uint32_t max_per_cell);
/* Build-defined conservative grid: triangle AABB (grown by eps) vs cell box. */
uint32_t orc_populate_aabb(const float *vertices, const uint32_t *trgls, uint32_t n_tri,
                           const uint32_t dims[3], const float lo[3], const float cell[3],
                           float eps, uint32_t *count, uint32_t *list,
                           uint32_t max_per_cell);

/* The same with the triangle's plane required to pass within eps of the cell
 * (for long triangles that do not run along an axis). */
uint32_t orc_populate_aabb_plane(const float *vertices, const uint32_t *trgls, uint32_t n_tri,
                                 const uint32_t dims[3], const float lo[3], const float cell[3],
                                 float eps, uint32_t *count, uint32_t *list,
                                 uint32_t max_per_cell);

/* ---- collision ------------------------------------------------------------ */
orc_coll orc_tri_coll_detect(const orc_world *w, orc_f3 o, orc_f3 p, uint32_t tri);
orc_coll orc_mt_coll_detect(const orc_world *w, orc_f3 o, orc_f3 p, uint32_t tri);
orc_coll orc_cube_coll_detect(const orc_world *w, orc_f3 o, orc_f3 p, uint32_t cube,
                              uint32_t excluded);
orc_f3 orc_reflect_pos(const orc_world *w, orc_f3 start, orc_f3 target, orc_f3 coll,
                       uint32_t tri, uint32_t type);
typedef struct { uint32_t first, last, ncoll; } orc_hits;
orc_f3 orc_coll_detect_grid(const orc_world *w, orc_f3 o, orc_f3 p, float u,
                            uint8_t *comp, uint16_t *fiber, orc_hits *hits);
/* Brute force over all triangles, nearest-hit semantics (grid-free check). */
orc_f3 orc_coll_detect_brute(const orc_world *w, orc_f3 o, orc_f3 p, float u,
                             uint8_t *comp, uint16_t *fiber, orc_hits *hits);

/* ---- stepping ------------------------------------------------------------- */
typedef struct {
    uint32_t rng;               /* ORC_RNG_* */
    uint32_t philox_key[2];
    uint64_t step0;             /* global step counter of the first step */
    const float *inject;        /* [iters][n][4] normals xyz + uniform */
    orc_f3 grad;                /* gradScaled, spinSystem.cpp:1931 */
    float phase_constant;       /* spinSystem.cpp:1938 */
    float weight;               /* moment mode lobe weight (+1, 0, -1) */
    uint32_t *trace;            /* [iters][n][5] first,last,ncoll,cell,label or NULL */
    int n_threads;              /* OpenMP threads (<=0: default) */
    uint64_t *stats;            /* NULL or [3]: steps, triangles whose box overlaps the
                                 * step segment's box, membrane hits (work model counters) */
} orc_step_args;

void orc_integrate(const orc_world *w, orc_state *s, uint32_t iters,
                   const orc_step_args *a);

/* ---- labelling / readout --------------------------------------------------- */
/* findSpinsInFibers with a deterministic retry jitter.  Returns the number of
 * spins whose labels stayed ambiguous after max_tries.  plus_x: 0 the
 * reference's dominant-axis ray, 1/2/3 a ray along +x/+y/+z. */
uint32_t orc_find_spins(const orc_world *w, const float *pos, uint32_t n,
                        float xmax, float ymax, float zmax, uint32_t max_tries,
                        uint32_t plus_x, uint8_t *comp, uint16_t *fiber, uint32_t *tries_out);
void orc_mr_signal(uint32_t n, const float *mag, const float *phase,
                   const uint8_t *comp, double out[5]);
/* Build-defined: per-direction signal from gradient moments. */
void orc_moment_signal(uint32_t n, const float *mag, const float *moment,
                       const float *x0, float wsum, const uint8_t *comp,
                       uint32_t ndir, const float *grad_scaled, float phase_constant,
                       double *out);

/* ---- formats --------------------------------------------------------------- */
typedef struct {
    uint32_t n_vert, n_tri, n_axon, n_myelin, n_glia, n_fibers, max_tri_on_surface;
    float xmax, ymax, zmax;
    float *vertices;    /* malloc'd, scaled by 1/spaceScale */
    uint32_t *trgls;    /* malloc'd */
    uint32_t *tri_info; /* malloc'd */
} orc_mesh;
/* 0 ok, -1 file missing, -2 axon/myelin count mismatch (spinSystem.cpp:593-596).
 * On -2 the arrays are still filled; release them with orc_free_mesh after
 * any return. */
int orc_parse_mesh(const char *path, float space_scale, orc_mesh *m);
void orc_free_mesh(orc_mesh *m);
/* Returns number of gradient lines, fills up to cap rows of 6 floats. */
int orc_parse_grads(const char *path, float *rows, int cap);
/* Per-update gradient sign schedule for one line: sign[k] in {+1,0,-1} for
 * updates k=0..n-1, readout at update n.  Returns n = L+B+R; lbr (4 ints,
 * or NULL) receives L, B, R and whether the readout happens. */
int orc_pgse_schedule(float ldelta, float bdelta, float readout, float dt,
                      int steps_per_update, int *sign, int cap, int *lbr);

#ifdef __cplusplus
}
#endif
#endif
