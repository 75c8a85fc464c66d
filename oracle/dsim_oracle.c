/*
 * dsim oracle: plain-C restatement of dSim3's CPU simulate path.
 * TEST INFRASTRUCTURE ONLY (see dsim_oracle.h).  Never linked into the
 * product library.
 *
 * Arithmetic contract: build with -ffp-contract=off.  Reference-semantics
 * functions use plain float expressions in the reference's evaluation order;
 * build-defined production functions use explicit fmaf() so that the HIP
 * kernels (also built with contraction off, using the same fmaf calls) give
 * bit-identical results on identical inputs.
 */
#include "dsim_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TWOPI_F 6.28318530717959f   /* spinSystem.cpp:20 */

static inline orc_f3 f3(float x, float y, float z) { orc_f3 p = {x, y, z}; return p; }

/* ======================================================================== */
/* RNG                                                                       */
/* ======================================================================== */

/* Marsaglia multiply-with-carry, dSim3/src/spinKernelCpu.cpp:41-45 */
uint32_t orc_mwc_next(uint32_t seed[2])
{
    seed[0] = 36969u * (seed[0] & 65535u) + (seed[0] >> 16);
    seed[1] = 18000u * (seed[1] & 65535u) + (seed[1] >> 16);
    return (seed[0] << 16) + seed[1];
}

/* Basic Box-Muller, dSim3/src/spinKernelCpu.cpp:27-32 */
void orc_box_muller(float *u1, float *u2)
{
    float r = sqrtf(-2.0f * logf(*u1));
    float phi = TWOPI_F * *u2;
    *u1 = r * cosf(phi);
    *u2 = r * sinf(phi);
}

/* Three normals plus the untransformed fourth uniform,
 * dSim3/src/spinKernelCpu.cpp:53-63 */
void orc_randn(uint32_t seed[2], float *n1, float *n2, float *n3, float *u)
{
    *n1 = ((float)orc_mwc_next(seed) + 1.0f) / 4294967296.0f;
    *n2 = ((float)orc_mwc_next(seed) + 1.0f) / 4294967296.0f;
    *n3 = ((float)orc_mwc_next(seed) + 1.0f) / 4294967296.0f;
    *u  = ((float)orc_mwc_next(seed) + 1.0f) / 4294967296.0f;
    float n4 = *u;
    orc_box_muller(n1, n2);
    orc_box_muller(n3, &n4);
}

/* Build-defined.  Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11): 10 rounds of the 4x32 S-box with
 * multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl key increments 0x9E3779B9 /
 * 0xBB67AE85. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Build-defined.  23 random bits centred in their bin: exactly representable,
 * strictly inside (0,1). */
float orc_u32_to_unit(uint32_t x)
{
    return ((float)(x >> 9) + 0.5f) * 1.1920928955078125e-07f; /* 2^-23 */
}

/* Build-defined.  Counter = (spin, step lo, step hi, stream 0).  Two
 * Box-Muller pairs give x,y from the first pair and z from the cosine branch
 * of the second. */
void orc_philox_step_normals(const uint32_t key[2], uint32_t spin, uint64_t step, float n[3])
{
    uint32_t ctr[4] = {spin, (uint32_t)step, (uint32_t)(step >> 32), 0u};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    float u1 = orc_u32_to_unit(o[0]), u2 = orc_u32_to_unit(o[1]);
    float u3 = orc_u32_to_unit(o[2]), u4 = orc_u32_to_unit(o[3]);
    float r1 = sqrtf(-2.0f * logf(u1));
    float r2 = sqrtf(-2.0f * logf(u3));
    n[0] = r1 * cosf(TWOPI_F * u2);
    n[1] = r1 * sinf(TWOPI_F * u2);
    n[2] = r2 * cosf(TWOPI_F * u4);
}

/* Build-defined.  Permeation uniform of the step: the normals use the top 23
 * bits of each word of the stream-0 block, so the 9 low bits of words 0 and 1
 * and 5 low bits of word 2 make 23 further bits, independent of the normals. */
float orc_philox_step_uniform(const uint32_t key[2], uint32_t spin, uint64_t step)
{
    uint32_t ctr[4] = {spin, (uint32_t)step, (uint32_t)(step >> 32), 0u};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    uint32_t m = (o[0] & 511u) | ((o[1] & 511u) << 9) | ((o[2] & 31u) << 18);
    return ((float)m + 0.5f) * 1.1920928955078125e-07f; /* 2^-23 */
}

/* ======================================================================== */
/* Geometry helpers                                                          */
/* ======================================================================== */

/* m_triangleHelpers, dSim3/src/spinSystem.cpp:612-648:
 * {n_hat(3), u=p3-p1 (3), v=p2-p1 (3), u.v, u.u, v.v} */
void orc_tri_helpers(const float *V, const uint32_t *T, uint32_t n_tri, float *H)
{
    for (uint32_t i = 0; i < n_tri; ++i) {
        uint32_t p1 = T[i * 3 + 0], p2 = T[i * 3 + 1], p3 = T[i * 3 + 2];
        float x1 = V[p1 * 3 + 0], y1 = V[p1 * 3 + 1], z1 = V[p1 * 3 + 2];
        float x2 = V[p2 * 3 + 0], y2 = V[p2 * 3 + 1], z2 = V[p2 * 3 + 2];
        float x3 = V[p3 * 3 + 0], y3 = V[p3 * 3 + 1], z3 = V[p3 * 3 + 2];
        float nx = (y2 - y1) * (z3 - z1) - (z2 - z1) * (y3 - y1);
        float ny = (z2 - z1) * (x3 - x1) - (x2 - x1) * (z3 - z1);
        float nz = (x2 - x1) * (y3 - y1) - (y2 - y1) * (x3 - x1);
        float len = sqrtf(nx * nx + ny * ny + nz * nz);
        float *h = H + (size_t)i * 12;
        h[0] = nx / len; h[1] = ny / len; h[2] = nz / len;
        h[3] = x3 - x1; h[4] = y3 - y1; h[5] = z3 - z1;
        h[6] = x2 - x1; h[7] = y2 - y1; h[8] = z2 - z1;
        h[9] = (x3 - x1) * (x2 - x1) + (y3 - y1) * (y2 - y1) + (z3 - z1) * (z2 - z1);
        h[10] = (x3 - x1) * (x3 - x1) + (y3 - y1) * (y3 - y1) + (z3 - z1) * (z3 - z1);
        h[11] = (x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1) + (z2 - z1) * (z2 - z1);
    }
}

/* Build-defined: segments within about 1e-3 rad of a triangle's plane
 * (det^2 < ORC_GRAZE_SQ |e1 x e2|^2 |d|^2) are decided in FP64
 * (orc_mt_coll_detect). */
#define ORC_GRAZE_SQ 1e-6f

/* Build-defined 64-byte Moller-Trumbore record: plane (n, n.v0), v0, e1=v1-v0
 * with ORC_GRAZE_SQ |e1 x e2|^2 in the fourth slot, e2=v2-v0.  The normal is
 * the reference helper normal. */
void orc_tri_mt_records(const float *V, const uint32_t *T, uint32_t n_tri, float *R)
{
    for (uint32_t i = 0; i < n_tri; ++i) {
        float h[12];
        orc_tri_helpers(V, T + (size_t)i * 3, 1, h);
        uint32_t a = T[i * 3 + 0], b = T[i * 3 + 1], c = T[i * 3 + 2];
        float *r = R + (size_t)i * 16;
        float v0x = V[a * 3 + 0], v0y = V[a * 3 + 1], v0z = V[a * 3 + 2];
        r[0] = h[0]; r[1] = h[1]; r[2] = h[2];
        r[3] = fmaf(h[0], v0x, fmaf(h[1], v0y, h[2] * v0z));
        r[4] = v0x; r[5] = v0y; r[6] = v0z; r[7] = 0.0f;
        r[8] = V[b * 3 + 0] - v0x; r[9] = V[b * 3 + 1] - v0y; r[10] = V[b * 3 + 2] - v0z;
        r[12] = V[c * 3 + 0] - v0x; r[13] = V[c * 3 + 1] - v0y; r[14] = V[c * 3 + 2] - v0z; r[15] = 0.0f;
        /* ORC_GRAZE_SQ |e1 x e2|^2, the near-grazing scale of det^2 / |d|^2 */
        float cx = fmaf(r[9], r[14], -(r[10] * r[13]));
        float cy = fmaf(r[10], r[12], -(r[8] * r[14]));
        float cz = fmaf(r[8], r[13], -(r[9] * r[12]));
        r[11] = ORC_GRAZE_SQ * fmaf(cx, cx, fmaf(cy, cy, cz * cz));
    }
}

/* calcCubePos, dSim3/src/spinSystem.cpp:691-717 (std::max/std::min
 * semantics kept, so NaN maps to cell 0).  With ref_cell_math the cell index
 * is floor((p-lo)/cell); lo=-1 gives the reference's (p+1.0f)/m_cubeLength
 * exactly. */
static inline uint32_t cell_axis(float p, float lo, float cell, float inv, uint32_t n, int ref)
{
    float q = ref ? floorf((p - lo) / cell) : floorf((p - lo) * inv);
    float hi = (float)(n - 1);
    float mn = (hi < q) ? hi : q;         /* std::min(q, hi) */
    float mx = (0.0f < mn) ? mn : 0.0f;   /* std::max(0, mn) */
    return (uint32_t)mx;
}

void orc_cube_pos(const orc_world *w, orc_f3 p, uint32_t c[3])
{
    c[0] = cell_axis(p.x, w->lo[0], w->cell[0], w->inv_cell[0], w->dims[0], w->ref_cell_math);
    c[1] = cell_axis(p.y, w->lo[1], w->cell[1], w->inv_cell[1], w->dims[1], w->ref_cell_math);
    c[2] = cell_axis(p.z, w->lo[2], w->cell[2], w->inv_cell[2], w->dims[2], w->ref_cell_math);
}

/* calcCubeHash, dSim3/src/spinSystem.cpp:756-759 (n*n*z + n*y + x) */
uint32_t orc_cube_hash(const orc_world *w, const uint32_t c[3])
{
    return (c[2] * w->dims[1] + c[1]) * w->dims[0] + c[0];
}

/* pointInCube, dSim3/src/spinSystem.cpp:1355-1357 */
int orc_point_in_cube(orc_f3 p, orc_f3 mn, orc_f3 mx)
{
    return ((p.x <= mx.x) && (p.x >= mn.x)) && ((p.y <= mx.y) && (p.y >= mn.y)) &&
           ((p.z <= mx.z) && (p.z >= mn.z));
}

static inline float dot3(orc_f3 a, orc_f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline orc_f3 sub3(orc_f3 a, orc_f3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline orc_f3 add3(orc_f3 a, orc_f3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline orc_f3 scl3(orc_f3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }

/* rayTriangleIntersect, dSim3/src/spinSystem.cpp:1360-1379 */
int orc_ray_tri_intersect(orc_f3 r1, orc_f3 r2, orc_f3 t1, orc_f3 t2, orc_f3 t3)
{
    orc_f3 u = sub3(t2, t1);
    orc_f3 v = sub3(t3, t1);
    orc_f3 n = f3(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
    float nl = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
    float s = 2.0f, t = 2.0f;
    n.x = n.x / nl; n.y = n.y / nl; n.z = n.z / nl;
    float r = dot3(n, sub3(t1, r1)) / dot3(n, sub3(r2, r1));
    if ((0 < r) & (r < 1)) {
        orc_f3 w = sub3(add3(r1, scl3(sub3(r2, r1), r)), t1);
        float uv = dot3(u, v), wv = dot3(w, v), vv = dot3(v, v), wu = dot3(w, u), uu = dot3(u, u);
        s = (uv * wv - vv * wu) / (uv * uv - uu * vv);
        t = (uv * wu - uu * wv) / (uv * uv - uu * vv);
    }
    return (s >= 0) && (t >= 0) && (s + t <= 1);
}

/* Cube corners by bit mask (bit0: x max, bit1: y max, bit2: z max). */
static inline orc_f3 corner(orc_f3 mn, orc_f3 mx, int m)
{
    return f3((m & 1) ? mx.x : mn.x, (m & 2) ? mx.y : mn.y, (m & 4) ? mx.z : mn.z);
}

/* The 12 cube edges (start, end) and 12 face half-triangles, in the argument
 * order of dSim3/src/spinSystem.cpp:1189-1244. */
static const int k_edges[12][2] = {{0, 1}, {0, 2}, {0, 4}, {6, 2}, {6, 4}, {6, 7},
                                   {3, 2}, {3, 1}, {3, 7}, {5, 7}, {5, 1}, {5, 4}};
static const int k_faces[12][3] = {{0, 1, 2}, {0, 4, 2}, {0, 4, 1}, {6, 2, 4}, {6, 2, 7}, {6, 4, 7},
                                   {3, 2, 1}, {3, 7, 1}, {3, 7, 2}, {5, 4, 7}, {5, 1, 7}, {5, 4, 1}};

static int tri_in_cube(orc_f3 p1, orc_f3 p2, orc_f3 p3, orc_f3 mn, orc_f3 mx)
{
    if (orc_point_in_cube(p1, mn, mx) || orc_point_in_cube(p2, mn, mx) || orc_point_in_cube(p3, mn, mx))
        return 1;
    for (int e = 0; e < 12; ++e)
        if (orc_ray_tri_intersect(corner(mn, mx, k_edges[e][0]), corner(mn, mx, k_edges[e][1]), p1, p2, p3))
            return 1;
    const orc_f3 a[3] = {p1, p1, p2}, b[3] = {p2, p3, p3};
    for (int s = 0; s < 3; ++s)
        for (int f = 0; f < 12; ++f)
            if (orc_ray_tri_intersect(a[s], b[s], corner(mn, mx, k_faces[f][0]),
                                      corner(mn, mx, k_faces[f][1]), corner(mn, mx, k_faces[f][2])))
                return 1;
    return 0;
}

/* populateCubes, dSim3/src/spinSystem.cpp:1094-1290 (grid fixed at n_cubes^3
 * over [-1,1]^3; two passes: size then fill, triangles appended in index
 * order). */
uint32_t orc_populate_cubes(const float *V, const uint32_t *T, uint32_t n_tri, uint32_t n_cubes,
                            uint32_t *count, uint32_t *list, uint32_t max_per_cell)
{
    orc_world g;
    memset(&g, 0, sizeof(g));
    float L = 2.0f / (float)n_cubes;  /* m_cubeLength = 2.0f / m_numCubes */
    for (int k = 0; k < 3; ++k) { g.dims[k] = n_cubes; g.lo[k] = -1.0f; g.cell[k] = L; }
    g.ref_cell_math = 1;
    size_t ncells = (size_t)n_cubes * n_cubes * n_cubes;
    memset(count, 0, ncells * sizeof(uint32_t));
    uint32_t maxc = 0;
    for (uint32_t t = 0; t < n_tri; ++t) {
        orc_f3 p1 = f3(V[T[t * 3] * 3], V[T[t * 3] * 3 + 1], V[T[t * 3] * 3 + 2]);
        orc_f3 p2 = f3(V[T[t * 3 + 1] * 3], V[T[t * 3 + 1] * 3 + 1], V[T[t * 3 + 1] * 3 + 2]);
        orc_f3 p3 = f3(V[T[t * 3 + 2] * 3], V[T[t * 3 + 2] * 3 + 1], V[T[t * 3 + 2] * 3 + 2]);
        orc_f3 mn = p1, mx = p1;
        if (p2.x < mn.x) mn.x = p2.x; if (p3.x < mn.x) mn.x = p3.x;
        if (p2.y < mn.y) mn.y = p2.y; if (p3.y < mn.y) mn.y = p3.y;
        if (p2.z < mn.z) mn.z = p2.z; if (p3.z < mn.z) mn.z = p3.z;
        if (p2.x > mx.x) mx.x = p2.x; if (p3.x > mx.x) mx.x = p3.x;
        if (p2.y > mx.y) mx.y = p2.y; if (p3.y > mx.y) mx.y = p3.y;
        if (p2.z > mx.z) mx.z = p2.z; if (p3.z > mx.z) mx.z = p3.z;
        uint32_t c0[3], c1[3];
        orc_cube_pos(&g, mn, c0);
        orc_cube_pos(&g, mx, c1);
        for (uint32_t xc = c0[0]; xc <= c1[0]; ++xc)
            for (uint32_t yc = c0[1]; yc <= c1[1]; ++yc)
                for (uint32_t zc = c0[2]; zc <= c1[2]; ++zc) {
                    uint32_t cc[3] = {xc, yc, zc};
                    uint32_t idx = orc_cube_hash(&g, cc);
                    orc_f3 bmin = f3(-1 + (float)xc * L, -1 + (float)yc * L, -1 + (float)zc * L);
                    orc_f3 bmax = f3(-1 + (float)(xc + 1) * L, -1 + (float)(yc + 1) * L, -1 + (float)(zc + 1) * L);
                    if (tri_in_cube(p1, p2, p3, bmin, bmax)) {
                        if (list) list[(size_t)idx * max_per_cell + count[idx]] = t;
                        count[idx] += 1;
                        if (count[idx] > maxc) maxc = count[idx];
                    }
                }
    }
    return maxc;
}

/* Build-defined conservative membership: triangle AABB grown by eps overlaps
 * the cell box [lo+i*cell, lo+(i+1)*cell] (inclusive); with `plane` set the
 * triangle's plane must also pass within eps of the box (a triangle lies in
 * its plane, so this drops only cells the triangle cannot touch). */
static uint32_t populate_boxes(const float *V, const uint32_t *T, uint32_t n_tri,
                               const uint32_t dims[3], const float lo[3], const float cell[3],
                               float eps, uint32_t *count, uint32_t *list, uint32_t max_per_cell,
                               int plane);

uint32_t orc_populate_aabb(const float *V, const uint32_t *T, uint32_t n_tri, const uint32_t dims[3],
                           const float lo[3], const float cell[3], float eps, uint32_t *count,
                           uint32_t *list, uint32_t max_per_cell)
{
    return populate_boxes(V, T, n_tri, dims, lo, cell, eps, count, list, max_per_cell, 0);
}

uint32_t orc_populate_aabb_plane(const float *V, const uint32_t *T, uint32_t n_tri,
                                 const uint32_t dims[3], const float lo[3], const float cell[3],
                                 float eps, uint32_t *count, uint32_t *list, uint32_t max_per_cell)
{
    return populate_boxes(V, T, n_tri, dims, lo, cell, eps, count, list, max_per_cell, 1);
}

static uint32_t populate_boxes(const float *V, const uint32_t *T, uint32_t n_tri,
                               const uint32_t dims[3], const float lo[3], const float cell[3],
                               float eps, uint32_t *count, uint32_t *list, uint32_t max_per_cell,
                               int plane)
{
    size_t ncells = (size_t)dims[0] * dims[1] * dims[2];
    memset(count, 0, ncells * sizeof(uint32_t));
    uint32_t maxc = 0;
    for (uint32_t t = 0; t < n_tri; ++t) {
        float mn[3], mx[3];
        for (int k = 0; k < 3; ++k) {
            float a = V[T[t * 3] * 3 + k], b = V[T[t * 3 + 1] * 3 + k], c = V[T[t * 3 + 2] * 3 + k];
            mn[k] = a; if (b < mn[k]) mn[k] = b; if (c < mn[k]) mn[k] = c;
            mx[k] = a; if (b > mx[k]) mx[k] = b; if (c > mx[k]) mx[k] = c;
            mn[k] = mn[k] - eps; mx[k] = mx[k] + eps;
        }
        int64_t i0[3], i1[3];
        int empty = 0;
        for (int k = 0; k < 3; ++k) {
            i0[k] = (int64_t)floorf((mn[k] - lo[k]) / cell[k]) - 1;
            i1[k] = (int64_t)floorf((mx[k] - lo[k]) / cell[k]) + 1;
            if (i0[k] < 0) i0[k] = 0;
            if (i1[k] > (int64_t)dims[k] - 1) i1[k] = (int64_t)dims[k] - 1;
            if (i0[k] > i1[k]) empty = 1;
        }
        if (empty) continue;
        for (int64_t z = i0[2]; z <= i1[2]; ++z) {
            float bz0 = lo[2] + (float)z * cell[2], bz1 = lo[2] + (float)(z + 1) * cell[2];
            if (bz0 > mx[2] || mn[2] > bz1) continue;
            for (int64_t y = i0[1]; y <= i1[1]; ++y) {
                float by0 = lo[1] + (float)y * cell[1], by1 = lo[1] + (float)(y + 1) * cell[1];
                if (by0 > mx[1] || mn[1] > by1) continue;
                for (int64_t x = i0[0]; x <= i1[0]; ++x) {
                    float bx0 = lo[0] + (float)x * cell[0], bx1 = lo[0] + (float)(x + 1) * cell[0];
                    if (bx0 > mx[0] || mn[0] > bx1) continue;
                    if (plane) {
                        const double b0[3] = {bx0, by0, bz0}, b1[3] = {bx1, by1, bz1};
                        double A[3], B[3], Cv[3], nrm[3], dist = 0.0, rad = 0.0, nn = 0.0;
                        for (int k = 0; k < 3; ++k) {
                            A[k] = V[T[t * 3] * 3 + k];
                            B[k] = V[T[t * 3 + 1] * 3 + k] - A[k];
                            Cv[k] = V[T[t * 3 + 2] * 3 + k] - A[k];
                        }
                        nrm[0] = B[1] * Cv[2] - B[2] * Cv[1];
                        nrm[1] = B[2] * Cv[0] - B[0] * Cv[2];
                        nrm[2] = B[0] * Cv[1] - B[1] * Cv[0];
                        for (int k = 0; k < 3; ++k) {
                            dist += nrm[k] * (0.5 * (b0[k] + b1[k]) - A[k]);
                            rad += 0.5 * (b1[k] - b0[k]) * fabs(nrm[k]);
                            nn += nrm[k] * nrm[k];
                        }
                        if (fabs(dist) > rad + (double)eps * sqrt(nn)) continue;
                    }
                    size_t idx = ((size_t)z * dims[1] + (size_t)y) * dims[0] + (size_t)x;
                    if (list) list[idx * max_per_cell + count[idx]] = t;
                    count[idx] += 1;
                    if (count[idx] > maxc) maxc = count[idx];
                }
            }
        }
    }
    return maxc;
}

/* ======================================================================== */
/* Collision                                                                 */
/* ======================================================================== */

/* triCollDetect, dSim3/src/spinKernelCpu.cpp:354-395 (softsurfer
 * plane-then-barycentric test; expression order kept). */
orc_coll orc_tri_coll_detect(const orc_world *w, orc_f3 o, orc_f3 p, uint32_t tri)
{
    orc_coll res;
    res.idx = ORC_NONE;
    res.type = 0;
    res.dist_sq = 400000000.0f;
    res.pt = f3(0, 0, 0);
    const float *h = w->helpers + (size_t)tri * 12;
    uint32_t P1 = w->tri_info[tri * 3 + 2];
    orc_f3 q = f3(w->vertices[P1 * 3 + 0], w->vertices[P1 * 3 + 1], w->vertices[P1 * 3 + 2]);
    float r = (h[0] * (q.x - o.x) + h[1] * (q.y - o.y) + h[2] * (q.z - o.z)) /
              (h[0] * (p.x - o.x) + h[1] * (p.y - o.y) + h[2] * (p.z - o.z));
    if ((0 < r) & (r < 1)) {
        float wu = (o.x + r * (p.x - o.x) - q.x) * h[3] + (o.y + r * (p.y - o.y) - q.y) * h[4] +
                   (o.z + r * (p.z - o.z) - q.z) * h[5];
        float wv = (o.x + r * (p.x - o.x) - q.x) * h[6] + (o.y + r * (p.y - o.y) - q.y) * h[7] +
                   (o.z + r * (p.z - o.z) - q.z) * h[8];
        float s = (h[9] * wv - h[11] * wu) / (h[9] * h[9] - h[10] * h[11]);
        float t = (h[9] * wu - h[10] * wv) / (h[9] * h[9] - h[10] * h[11]);
        if ((s >= 0) & (t >= 0) & (s + t <= 1)) {
            /* collPoint = oPos + r*(pos-oPos) (float3 operators of spinSystem.h) */
            res.pt = f3(o.x + (p.x - o.x) * r, o.y + (p.y - o.y) * r, o.z + (p.z - o.z) * r);
            res.dist_sq = (o.x - res.pt.x) * (o.x - res.pt.x) + (o.y - res.pt.y) * (o.y - res.pt.y) +
                          (o.z - res.pt.z) * (o.z - res.pt.z);
            res.idx = tri;
            res.type = 1;
            if ((s == 0) | (t == 0) | (s + t == 1)) {
                res.type = 2;
                if (((s == 0) & (t == 0)) | ((s == 0) & (t == 1)) | ((s == 1) & (t == 0)))
                    res.type = 3;
            }
        }
    }
    return res;
}

/* Build-defined.  Moller-Trumbore in FP64 for a near-grazing segment.  The
 * differences of float coordinates are exact in double; the inside tests
 * compare the numerators u*det, v*det, t*det with det after making det
 * positive (no division); t and the hit point are rounded back to float. */
static orc_coll mt_fp64(const float *r, orc_f3 o, orc_f3 p, uint32_t tri, orc_coll res)
{
    const double Dx = (double)p.x - (double)o.x, Dy = (double)p.y - (double)o.y,
                 Dz = (double)p.z - (double)o.z;
    const double Tx = (double)o.x - (double)r[4], Ty = (double)o.y - (double)r[5],
                 Tz = (double)o.z - (double)r[6];
    const double E1x = r[8], E1y = r[9], E1z = r[10], E2x = r[12], E2y = r[13], E2z = r[14];
    const double Qx = Dy * E2z - Dz * E2y, Qy = Dz * E2x - Dx * E2z, Qz = Dx * E2y - Dy * E2x;
    double det = E1x * Qx + E1y * Qy + E1z * Qz;
    if (det == 0.0) return res;
    const double s = det < 0.0 ? -1.0 : 1.0;
    det *= s;
    const double un = s * (Tx * Qx + Ty * Qy + Tz * Qz);
    if (!(un >= 0.0 && un <= det)) return res;
    const double Cx = Ty * E1z - Tz * E1y, Cy = Tz * E1x - Tx * E1z, Cz = Tx * E1y - Ty * E1x;
    const double vn = s * (Dx * Cx + Dy * Cy + Dz * Cz);
    if (!(vn >= 0.0 && un + vn <= det)) return res;
    const double tn = s * (E2x * Cx + E2y * Cy + E2z * Cz);
    if (!(tn > 0.0 && tn < det)) return res;
    const double t = tn / det;
    res.pt = f3((float)((double)o.x + t * Dx), (float)((double)o.y + t * Dy),
                (float)((double)o.z + t * Dz));
    res.dist_sq = (float)t;
    res.idx = tri;
    res.type = 1;
    if (un == 0.0 || vn == 0.0 || un + vn == det) {
        res.type = 2;
        if ((un == 0.0 && vn == 0.0) || (un == det && vn == 0.0) || (un == 0.0 && vn == det))
            res.type = 3;
    }
    return res;
}

/* Build-defined production test: plane-side prefilter (end points on different
 * sides as the float expression evaluates them), then Moller-Trumbore
 * ("Fast, minimum storage ray/triangle intersection", JGT 1997) restricted to
 * the open segment 0<t<1, in FP64 when the segment is within about 1e-3 rad
 * of the plane (det^2 < ORC_GRAZE_SQ |e1 x e2|^2 |d|^2).  dist_sq carries t. */
orc_coll orc_mt_coll_detect(const orc_world *w, orc_f3 o, orc_f3 p, uint32_t tri)
{
    orc_coll res;
    res.idx = ORC_NONE;
    res.type = 0;
    res.dist_sq = 2.0f;
    res.pt = f3(0, 0, 0);
    const float *r = w->tri_mt + (size_t)tri * 16;
    float s0 = fmaf(r[0], o.x, fmaf(r[1], o.y, fmaf(r[2], o.z, -r[3])));
    float s1 = fmaf(r[0], p.x, fmaf(r[1], p.y, fmaf(r[2], p.z, -r[3])));
    if ((s0 > 0.0f) == (s1 > 0.0f)) return res;
    float dx = p.x - o.x, dy = p.y - o.y, dz = p.z - o.z;
    float e1x = r[8], e1y = r[9], e1z = r[10];
    float e2x = r[12], e2y = r[13], e2z = r[14];
    float px = fmaf(dy, e2z, -(dz * e2y));
    float py = fmaf(dz, e2x, -(dx * e2z));
    float pz = fmaf(dx, e2y, -(dy * e2x));
    float det = fmaf(e1x, px, fmaf(e1y, py, e1z * pz));
    float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    if (det * det < r[11] * d2) return mt_fp64(r, o, p, tri, res);
    if (det == 0.0f) return res; /* unreachable unless r[11] * d2 == 0 */
    float inv = 1.0f / det;
    float tx = o.x - r[4], ty = o.y - r[5], tz = o.z - r[6];
    float u = fmaf(tx, px, fmaf(ty, py, tz * pz)) * inv;
    if (!(u >= 0.0f && u <= 1.0f)) return res;
    float qx = fmaf(ty, e1z, -(tz * e1y));
    float qy = fmaf(tz, e1x, -(tx * e1z));
    float qz = fmaf(tx, e1y, -(ty * e1x));
    float v = fmaf(dx, qx, fmaf(dy, qy, dz * qz)) * inv;
    if (!(v >= 0.0f && u + v <= 1.0f)) return res;
    float t = fmaf(e2x, qx, fmaf(e2y, qy, e2z * qz)) * inv;
    if (!(t > 0.0f && t < 1.0f)) return res;
    res.pt = f3(fmaf(t, dx, o.x), fmaf(t, dy, o.y), fmaf(t, dz, o.z));
    res.dist_sq = t;
    res.idx = tri;
    res.type = 1;
    if (u == 0.0f || v == 0.0f || u + v == 1.0f) {
        res.type = 2;
        if ((u == 0.0f && v == 0.0f) || (u == 1.0f && v == 0.0f) || (u == 0.0f && v == 1.0f)) res.type = 3;
    }
    return res;
}

static inline orc_coll test_tri(const orc_world *w, orc_f3 o, orc_f3 p, uint32_t tri)
{
    return w->test ? orc_mt_coll_detect(w, o, p, tri) : orc_tri_coll_detect(w, o, p, tri);
}

/* Strict "closer" rule.  Reference: keep the first minimal distance in list
 * order (spinKernelCpu.cpp:328).  Nearest mode: (t, index) lexicographic, so
 * the winner does not depend on candidate order. */
static inline int better(const orc_world *w, const orc_coll *c, const orc_coll *best)
{
    if (c->type == 0) return 0;
    if (w->traversal == 0) return c->dist_sq < best->dist_sq;
    if (best->type == 0) return 1;
    return (c->dist_sq < best->dist_sq) || (c->dist_sq == best->dist_sq && c->idx < best->idx);
}

/* cubeCollDetect, dSim3/src/spinKernelCpu.cpp:311-337 */
orc_coll orc_cube_coll_detect(const orc_world *w, orc_f3 o, orc_f3 p, uint32_t cube, uint32_t excluded)
{
    orc_coll best;
    best.type = 0;
    best.idx = ORC_NONE;
    best.dist_sq = w->test ? 2.0f : 400000000.0f;
    best.pt = f3(0, 0, 0);
    uint32_t kmax = w->cell_count[cube];
    for (uint32_t k = 0; k < kmax; ++k) {
        uint32_t i = w->cell_list[(size_t)cube * w->max_per_cell + k];
        if (i != excluded) {
            orc_coll c = test_tri(w, o, p, i);
            if (better(w, &c, &best)) best = c;
        }
    }
    return best;
}

/* reflectPos, dSim3/src/spinKernelCpu.cpp:406-422 (reference expression
 * order).  The production test uses the same Householder reflection written
 * with fmaf. */
orc_f3 orc_reflect_pos(const orc_world *w, orc_f3 sp, orc_f3 tp, orc_f3 cp, uint32_t tri, uint32_t type)
{
    orc_f3 r;
    if ((w->reflection_type == 0) | (type > 1)) return sp;
    if (w->test == 0) {
        orc_f3 s = sub3(tp, cp);
        const float *h = w->helpers + (size_t)tri * 12;
        float nx = h[0], ny = h[1], nz = h[2];
        r.x = (1 - 2 * nx * nx) * s.x - 2 * nx * ny * s.y - 2 * nx * nz * s.z + cp.x;
        r.y = -2 * nx * ny * s.x + (1 - 2 * ny * ny) * s.y - 2 * ny * nz * s.z + cp.y;
        r.z = -2 * nx * nz * s.x - 2 * ny * nz * s.y + (1 - 2 * nz * nz) * s.z + cp.z;
    } else {
        const float *h = w->tri_mt + (size_t)tri * 16;
        float sx = tp.x - cp.x, sy = tp.y - cp.y, sz = tp.z - cp.z;
        float dn = fmaf(h[0], sx, fmaf(h[1], sy, h[2] * sz));
        float k = -2.0f * dn;
        r.x = fmaf(k, h[0], sx) + cp.x;
        r.y = fmaf(k, h[1], sy) + cp.y;
        r.z = fmaf(k, h[2], sz) + cp.z;
    }
    return r;
}

/* Compartment change on transmission.  Reference table:
 * dSim3/src/spinKernelCpu.cpp:251-270.  Physical table (build-defined):
 * crossing the axon wall moves between intra-axonal and the fibre's myelin
 * (or extra-axonal space when the fibre has no myelin), myelin wall between
 * myelin and extra, glia wall between glia interior and extra. */
static void transit(const orc_world *w, uint32_t tri, uint8_t *comp, uint16_t *fiber)
{
    uint32_t f = w->tri_info[tri * 3 + 0], m = w->tri_info[tri * 3 + 1];
    uint8_t c = *comp;
    if (w->transitions == 0) {
        if (c == 2) {
            if (m == 0) { c = 1; } else { c = 0; *fiber = ORC_FIBER_NONE; }
        } else if (c == 1) {
            c = 2;
        } else if (c == 3) {
            c = 0; *fiber = ORC_FIBER_NONE;
        } else {
            *fiber = (uint16_t)f;
            c = (m == 1) ? 2 : 3;
        }
    } else {
        int myel = (w->fiber_has_myelin && f < w->n_fibers) ? w->fiber_has_myelin[f] : 0;
        if (c == 0) {
            *fiber = (uint16_t)f;
            c = (m == 0) ? 1 : (m == 1) ? 2 : 3;
        } else if (c == 1) {
            if (myel) { c = 2; } else { c = 0; *fiber = ORC_FIBER_NONE; }
        } else if (c == 2) {
            if (m == 0) { c = 1; } else { c = 0; *fiber = ORC_FIBER_NONE; }
        } else {
            c = 0; *fiber = ORC_FIBER_NONE;
        }
    }
    *comp = c;
}

/* Apply one hit: reflect (u <= threshold) or transmit,
 * dSim3/src/spinKernelCpu.cpp:238-272. */
static orc_f3 resolve_hit(const orc_world *w, orc_f3 o, orc_f3 p, float u, const orc_coll *hit,
                          float *umax, float *umin, uint8_t *comp, uint16_t *fiber, orc_f3 endp)
{
    float thr = *umax - (*umax - *umin) * w->permeability;
    if (u <= thr) {
        endp = orc_reflect_pos(w, o, p, hit->pt, hit->idx, hit->type);
        *umax = thr;
    } else {
        *umin = thr;
        transit(w, hit->idx, comp, fiber);
    }
    return endp;
}

static inline void note_hit(orc_hits *h, uint32_t idx)
{
    if (!h) return;
    if (h->ncoll == 0) h->first = idx;
    h->last = idx;
    h->ncoll += 1;
}

/* collDetectRectGrid, dSim3/src/spinKernelCpu.cpp:198-301.  Cells are walked
 * x-major from the start cell toward the end cell; traversal 0 stops at the
 * first cell holding any hit (reference), traversal 1 visits every cell of
 * the range and keeps the nearest hit. */
orc_f3 orc_coll_detect_grid(const orc_world *w, orc_f3 o, orc_f3 p, float u, uint8_t *comp,
                            uint16_t *fiber, orc_hits *hits)
{
    orc_f3 endp = p;
    uint32_t excluded = ORC_NONE;
    float umax = 1.0f, umin = 0.0f;
    if (hits) { hits->first = hits->last = ORC_NONE; hits->ncoll = 0; }
    for (int guard = 0; guard < 100000; ++guard) {
        uint32_t sc[3], ec[3];
        orc_cube_pos(w, o, sc);
        orc_cube_pos(w, p, ec);
        int inc[3];
        for (int k = 0; k < 3; ++k) inc[k] = (ec[k] > sc[k]) - (ec[k] < sc[k]);
        orc_coll best;
        best.type = 0; best.idx = ORC_NONE; best.dist_sq = w->test ? 2.0f : 400000000.0f; best.pt = f3(0, 0, 0);
        uint32_t cx = sc[0];
        do {
            uint32_t cy = sc[1];
            do {
                uint32_t cz = sc[2];
                do {
                    uint32_t cc[3] = {cx, cy, cz};
                    orc_coll c = orc_cube_coll_detect(w, o, p, orc_cube_hash(w, cc), excluded);
                    if (w->traversal == 0) best = c;
                    else if (better(w, &c, &best)) best = c;
                    cz += (uint32_t)inc[2];
                } while (cz != ec[2] + (uint32_t)inc[2] && (w->traversal == 1 || best.type == 0));
                cy += (uint32_t)inc[1];
            } while (cy != ec[1] + (uint32_t)inc[1] && (w->traversal == 1 || best.type == 0));
            cx += (uint32_t)inc[0];
        } while (cx != ec[0] + (uint32_t)inc[0] && (w->traversal == 1 || best.type == 0));
        if (best.type == 0) break;
        note_hit(hits, best.idx);
        endp = resolve_hit(w, o, p, u, &best, &umax, &umin, comp, fiber, endp);
        o = best.pt;
        p = endp;
        excluded = best.idx;
    }
    return endp;
}

/* Grid-free check of the nearest-hit mode: every triangle is a candidate. */
orc_f3 orc_coll_detect_brute(const orc_world *w, orc_f3 o, orc_f3 p, float u, uint8_t *comp,
                             uint16_t *fiber, orc_hits *hits)
{
    orc_f3 endp = p;
    uint32_t excluded = ORC_NONE;
    float umax = 1.0f, umin = 0.0f;
    if (hits) { hits->first = hits->last = ORC_NONE; hits->ncoll = 0; }
    for (int guard = 0; guard < 100000; ++guard) {
        orc_coll best;
        best.type = 0; best.idx = ORC_NONE; best.dist_sq = w->test ? 2.0f : 400000000.0f; best.pt = f3(0, 0, 0);
        for (uint32_t t = 0; t < w->n_tri; ++t) {
            if (t == excluded) continue;
            orc_coll c = test_tri(w, o, p, t);
            if (better(w, &c, &best)) best = c;
        }
        if (best.type == 0) break;
        note_hit(hits, best.idx);
        endp = resolve_hit(w, o, p, u, &best, &umax, &umin, comp, fiber, endp);
        o = best.pt;
        p = endp;
        excluded = best.idx;
    }
    return endp;
}

/* ======================================================================== */
/* Stepping                                                                  */
/* ======================================================================== */

/* Number of distinct triangles whose bounding box overlaps the box of the
 * segment o-p: the candidate count of an exact bounding-box search (what the
 * reference's R-tree query returns, spinKernelCpu.cpp:125), independent of the
 * grid.  Work-model counter only. */
static uint32_t aabb_candidates(const orc_world *w, orc_f3 o, orc_f3 p)
{
    float mn[3] = {o.x, o.y, o.z}, mx[3] = {o.x, o.y, o.z};
    if (p.x < mn[0]) mn[0] = p.x; if (p.y < mn[1]) mn[1] = p.y; if (p.z < mn[2]) mn[2] = p.z;
    if (p.x > mx[0]) mx[0] = p.x; if (p.y > mx[1]) mx[1] = p.y; if (p.z > mx[2]) mx[2] = p.z;
    uint32_t c0[3], c1[3];
    orc_cube_pos(w, f3(mn[0], mn[1], mn[2]), c0);
    orc_cube_pos(w, f3(mx[0], mx[1], mx[2]), c1);
    uint32_t buf[4096];
    uint32_t nb = 0;
    for (uint32_t z = c0[2]; z <= c1[2]; ++z)
        for (uint32_t y = c0[1]; y <= c1[1]; ++y)
            for (uint32_t x = c0[0]; x <= c1[0]; ++x) {
                uint32_t cc[3] = {x, y, z}, h = orc_cube_hash(w, cc);
                for (uint32_t k = 0; k < w->cell_count[h] && nb < 4096; ++k) {
                    uint32_t t = w->cell_list[(size_t)h * w->max_per_cell + k], dup = 0;
                    for (uint32_t q = 0; q < nb; ++q)
                        if (buf[q] == t) { dup = 1; break; }
                    if (!dup) buf[nb++] = t;
                }
            }
    uint32_t cnt = 0;
    for (uint32_t q = 0; q < nb; ++q) {
        const uint32_t *T = w->trgls + (size_t)buf[q] * 3;
        int ok = 1;
        for (int k = 0; k < 3 && ok; ++k) {
            float a = w->vertices[T[0] * 3 + k], b = w->vertices[T[1] * 3 + k], c = w->vertices[T[2] * 3 + k];
            float lo = fminf(a, fminf(b, c)), hi = fmaxf(a, fmaxf(b, c));
            if (lo > mx[k] || mn[k] > hi) ok = 0;
        }
        cnt += (uint32_t)ok;
    }
    return cnt;
}

/* cpuIntegrate, dSim3/src/spinKernelCpu.cpp:430-499, one spin one step. */
static void step_one(const orc_world *w, orc_state *s, uint32_t i, uint32_t it, const orc_step_args *a)
{
    orc_f3 pos = f3(s->pos[i * 3 + 0], s->pos[i * 3 + 1], s->pos[i * 3 + 2]);
    orc_f3 o = pos;
    float n[3], u = 0.0f;
    uint64_t gstep = a->step0 + it;
    if (a->rng == ORC_RNG_INJECT) {
        const float *q = a->inject + ((size_t)it * s->n + i) * 4;
        n[0] = q[0]; n[1] = q[1]; n[2] = q[2]; u = q[3];
    } else if (a->rng == ORC_RNG_MWC) {
        orc_randn(s->seed + (size_t)i * 2, &n[0], &n[1], &n[2], &u);
    } else {
        orc_philox_step_normals(a->philox_key, i, gstep, n);
    }
    uint8_t comp = s->comp[i];
    uint16_t fiber = s->fiber[i];
    float sd = w->std_dev[comp];
    if (w->test == 0) {
        pos.x += n[0] * sd; pos.y += n[1] * sd; pos.z += n[2] * sd;
    } else {
        pos.x = fmaf(n[0], sd, pos.x); pos.y = fmaf(n[1], sd, pos.y); pos.z = fmaf(n[2], sd, pos.z);
    }
    orc_hits hits;
    if (a->rng == ORC_RNG_PHILOX) {
        /* the permeation uniform is only needed when a hit happens; a dry run
         * would change nothing, so draw it up front for simplicity */
        u = orc_philox_step_uniform(a->philox_key, i, gstep);
    }
    if (a->stats) {
        uint32_t cand = aabb_candidates(w, o, pos);
#ifdef _OPENMP
#pragma omp atomic
#endif
        a->stats[1] += cand;
    }
    pos = orc_coll_detect_grid(w, o, pos, u, &comp, &fiber, &hits);
    if (a->stats) {
#ifdef _OPENMP
#pragma omp atomic
#endif
        a->stats[0] += 1;
#ifdef _OPENMP
#pragma omp atomic
#endif
        a->stats[2] += hits.ncoll;
    }
    if (w->boundary == 0) {
        /* spinKernelCpu.cpp:469-474 */
        float hi = 1.0f - w->spin_radius, lo = -1.0f + w->spin_radius;
        if (pos.x > hi) pos.x = hi; else if (pos.x < lo) pos.x = lo;
        if (pos.y > hi) pos.y = hi; else if (pos.y < lo) pos.y = lo;
        if (pos.z > hi) pos.z = hi; else if (pos.z < lo) pos.z = lo;
    }
    if (s->disp) {
        s->disp[i * 3 + 0] += pos.x - o.x;
        s->disp[i * 3 + 1] += pos.y - o.y;
        s->disp[i * 3 + 2] += pos.z - o.z;
    }
    int32_t kw[3] = {0, 0, 0};
    if (s->wrap) { kw[0] = s->wrap[i * 3 + 0]; kw[1] = s->wrap[i * 3 + 1]; kw[2] = s->wrap[i * 3 + 2]; }
    float wlo[3] = {-1.0f, -1.0f, -1.0f}, whi[3] = {1.0f, 1.0f, 1.0f};
    if (w->wrap_lo[0] != 0.0f || w->wrap_hi[0] != 0.0f || w->wrap_lo[1] != 0.0f ||
        w->wrap_hi[1] != 0.0f || w->wrap_lo[2] != 0.0f || w->wrap_hi[2] != 0.0f)
        for (int k = 0; k < 3; ++k) { wlo[k] = w->wrap_lo[k]; whi[k] = w->wrap_hi[k]; }
    float wlen[3] = {whi[0] - wlo[0], whi[1] - wlo[1], whi[2] - wlo[2]};
    if (w->boundary == 1) {
        /* build-defined periodic wrap into [lo, hi) per axis, counting crossings */
        float *c[3] = {&pos.x, &pos.y, &pos.z};
        for (int k = 0; k < 3; ++k) {
            if (*c[k] >= whi[k]) { *c[k] -= wlen[k]; kw[k] += 1; }
            else if (*c[k] < wlo[k]) { *c[k] += wlen[k]; kw[k] -= 1; }
        }
        if (s->wrap) { s->wrap[i * 3 + 0] = kw[0]; s->wrap[i * 3 + 1] = kw[1]; s->wrap[i * 3 + 2] = kw[2]; }
    }
    s->pos[i * 3 + 0] = pos.x; s->pos[i * 3 + 1] = pos.y; s->pos[i * 3 + 2] = pos.z;
    /* spinKernelCpu.cpp:486 forward-Euler T2; production uses dt/T2 */
    if (w->test == 0) s->mag[i] += -s->mag[i] / w->t2[comp] * w->dt;
    else s->mag[i] = fmaf(-s->mag[i], w->decay[comp], s->mag[i]);
    /* spinKernelCpu.cpp:490 phase, post-step position; in the periodic voxel
     * (build-defined) the unwrapped position x + k * period */
    if (s->phase && !(w->boundary == 1 && s->wrap))
        s->phase[i] += (a->grad.x * pos.x + a->grad.y * pos.y + a->grad.z * pos.z) * a->phase_constant;
    else if (s->phase) {
        float ux = pos.x + (float)kw[0] * wlen[0];
        float uy = pos.y + (float)kw[1] * wlen[1];
        float uz = pos.z + (float)kw[2] * wlen[2];
        s->phase[i] += (a->grad.x * ux + a->grad.y * uy + a->grad.z * uz) * a->phase_constant;
    }
    if (s->moment && a->weight != 0.0f) {
        for (int k = 0; k < 3; ++k) {
            float y = fmaf(a->weight, s->disp[i * 3 + k], -s->moment_c[i * 3 + k]);
            float t = s->moment[i * 3 + k] + y;
            s->moment_c[i * 3 + k] = (t - s->moment[i * 3 + k]) - y;
            s->moment[i * 3 + k] = t;
        }
    }
    s->comp[i] = comp;
    s->fiber[i] = fiber;
    if (a->trace) {
        uint32_t *tr = a->trace + ((size_t)it * s->n + i) * 5;
        uint32_t cc[3];
        orc_cube_pos(w, pos, cc);
        tr[0] = hits.first; tr[1] = hits.last; tr[2] = hits.ncoll;
        tr[3] = orc_cube_hash(w, cc);
        tr[4] = (uint32_t)comp | ((uint32_t)fiber << 8);
    }
}

/* cpuIntegrateSystem loop, dSim3/src/spinKernelCpu.cpp:6-16, iterated
 * `iters` times as SpinSystem::updateSpins does (spinSystem.cpp:1964-1968).
 * Spins are independent, so the spin loop is split over threads and each
 * spin runs its steps back to back (same results as step-major order). */
void orc_integrate(const orc_world *w, orc_state *s, uint32_t iters, const orc_step_args *a)
{
#ifdef _OPENMP
    int nt = a->n_threads > 0 ? a->n_threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
#endif
    for (int64_t i = 0; i < (int64_t)s->n; ++i)
        for (uint32_t it = 0; it < iters; ++it) step_one(w, s, (uint32_t)i, it, a);
}

/* ======================================================================== */
/* Labelling and readout                                                     */
/* ======================================================================== */

static const uint32_t k_label_key[2] = {0x6A09E667u, 0xBB67AE85u};

/* findSpinsInFibers, dSim3/src/spinSystem.cpp:1615-1852.  Candidates are the
 * triangles whose bounding box overlaps the ray's (the R-tree query of
 * spinSystem.cpp:1707, overlap test RTree.h:1550-1563), taken in index order.
 * The retry jitter replaces frand() by a Philox draw so that it is
 * reproducible, a ray that finds two odd fibres of one membrane type is
 * re-cast as well, and a spin that stays ambiguous after max_tries keeps the
 * last attempt's label (the reference instead moves it to a new random
 * position, spinSystem.cpp:1628-1633). */
uint32_t orc_find_spins(const orc_world *w, const float *pos, uint32_t n, float xmax, float ymax,
                        float zmax, uint32_t max_tries, uint32_t plus_x, uint8_t *comp_out,
                        uint16_t *fiber_out, uint32_t *tries_out)
{
    uint32_t nT = w->n_tri;
    const float *V = w->vertices;
    const uint32_t *T = w->trgls;
    float *bb = (float *)malloc((size_t)nT * 6 * sizeof(float));
    for (uint32_t t = 0; t < nT; ++t) {
        for (int k = 0; k < 3; ++k) {
            float a = V[T[t * 3] * 3 + k], b = V[T[t * 3 + 1] * 3 + k], c = V[T[t * 3 + 2] * 3 + k];
            float mn = a, mx = a;
            if (b < mn) mn = b; if (c < mn) mn = c;
            if (b > mx) mx = b; if (c > mx) mx = c;
            bb[t * 6 + k] = mn; bb[t * 6 + 3 + k] = mx;
        }
    }
    uint32_t ambiguous = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : ambiguous)
#endif
    for (int64_t si = 0; si < (int64_t)n; ++si) {
        uint32_t *hit = (uint32_t *)malloc((size_t)(nT ? nT : 1) * sizeof(uint32_t));
        uint32_t *grp = (uint32_t *)malloc((size_t)(nT ? nT : 1) * sizeof(uint32_t));
        orc_f3 sp = f3(pos[si * 3 + 0], pos[si * 3 + 1], pos[si * 3 + 2]);
        uint32_t tries = 0;
        uint8_t comp = 0;
        uint16_t fib = ORC_FIBER_NONE;
        for (;;) {
            orc_f3 le = sp;
            /* build-defined: ray along +x, +y or +z (plus_x = 1, 2, 3), so that
             * open prisms not parallel to that axis are counted */
            if (plus_x == 1)
                le.x = (float)(1.1 * (double)xmax);
            else if (plus_x == 2)
                le.y = (float)(1.1 * (double)ymax);
            else if (plus_x == 3)
                le.z = (float)(1.1 * (double)zmax);
            else if ((fabsf(sp.x) > fabsf(sp.y)) & (fabsf(sp.x) > fabsf(sp.z)))
                le.x = (float)(((sp.x >= 0) - (sp.x < 0)) * 1.1 * (double)xmax);
            else if (fabsf(sp.y) > fabsf(sp.z))
                le.y = (float)(((sp.y >= 0) - (sp.y < 0)) * 1.1 * (double)ymax);
            else
                le.z = (float)(((sp.z >= 0) - (sp.z < 0)) * 1.1 * (double)zmax);
            if (tries > 0) {
                uint32_t ctr[4] = {(uint32_t)si, tries, 0u, 0u}, o[4];
                orc_philox4x32_10(ctr, k_label_key, o);
                le.x = (float)((double)le.x + (double)orc_u32_to_unit(o[0]) * 0.05);
                le.y = (float)((double)le.y + (double)orc_u32_to_unit(o[1]) * 0.05);
                le.z = (float)((double)le.z + (double)orc_u32_to_unit(o[2]) * 0.05);
            }
            float rmin[3] = {sp.x, sp.y, sp.z}, rmax[3] = {sp.x, sp.y, sp.z};
            if (le.x < rmin[0]) rmin[0] = le.x; if (le.y < rmin[1]) rmin[1] = le.y; if (le.z < rmin[2]) rmin[2] = le.z;
            if (le.x > rmax[0]) rmax[0] = le.x; if (le.y > rmax[1]) rmax[1] = le.y; if (le.z > rmax[2]) rmax[2] = le.z;
            uint32_t nh = 0;
            for (uint32_t t = 0; t < nT; ++t) {
                const float *b = bb + (size_t)t * 6;
                if (rmin[0] > b[3] || b[0] > rmax[0] || rmin[1] > b[4] || b[1] > rmax[1] ||
                    rmin[2] > b[5] || b[2] > rmax[2])
                    continue;
                hit[nh++] = t;
            }
            uint32_t nedge = 0, nvert = 0, inside = ORC_FIBER_NONE, multi = 0;
            uint8_t sc = 0;
            for (uint32_t mt = 0; mt < 3 && inside == ORC_FIBER_NONE; ++mt) {
                uint32_t nodd = 0;
                for (uint32_t k = 0; k < nh; ++k) {
                    uint32_t ti = hit[k];
                    if (ti == ORC_NONE) continue;
                    uint32_t fi = w->tri_info[ti * 3], mi = w->tri_info[ti * 3 + 1];
                    if (mi != mt) continue;
                    hit[k] = ORC_NONE;
                    uint32_t ng = 0;
                    grp[ng++] = ti;
                    for (uint32_t q = k + 1; q < nh; ++q) {
                        uint32_t t2 = hit[q];
                        if (t2 == ORC_NONE) continue;
                        if (w->tri_info[t2 * 3] == fi && w->tri_info[t2 * 3 + 1] == mi) {
                            grp[ng++] = t2;
                            hit[q] = ORC_NONE;
                        }
                    }
                    uint32_t ncoll = 0;
                    for (uint32_t p = 0; p < ng; ++p) {
                        /* the physical mode (test 1) counts crossings with its own
                         * Moller-Trumbore test */
                        orc_coll r = test_tri(w, sp, le, grp[p]);
                        if (r.type > 0) {
                            ncoll++;
                            if (r.type == 2) nedge++;
                            if (r.type == 3) nvert++;
                        }
                    }
                    if (ncoll % 2 == 1) {
                        inside = fi;
                        sc = (mt == 0) ? 1 : (mt == 2) ? 3 : 2;
                        nodd++;
                    }
                }
                /* two odd fibres of one membrane type: a crossing was counted
                 * twice or missed at a shared edge; re-cast like an edge hit */
                if (nodd > 1) multi = 1;
            }
            if (inside == ORC_FIBER_NONE) sc = 0;
            comp = sc;
            fib = (uint16_t)inside;
            if (nedge == 0 && nvert == 0 && !multi) break;
            tries++;
            if (tries > max_tries) { ambiguous++; break; }
        }
        comp_out[si] = comp;
        fiber_out[si] = fib;
        if (tries_out) tries_out[si] = tries;
        free(hit);
        free(grp);
    }
    free(bb);
    return ambiguous;
}

/* getMrSignal(double[]), dSim3/src/spinSystem.cpp:1452-1500 */
void orc_mr_signal(uint32_t n, const float *mag, const float *phase, const uint8_t *comp, double out[5])
{
    double xm[5] = {0, 0, 0, 0, 0}, ym[5] = {0, 0, 0, 0, 0};
    uint32_t cnt[5] = {0, 0, 0, 0, 0};
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t c = comp[i];
        xm[0] += cos(phase[i]) * mag[i];
        ym[0] += sin(phase[i]) * mag[i];
        cnt[0]++;
        xm[c + 1] += cos(phase[i]) * mag[i];
        ym[c + 1] += sin(phase[i]) * mag[i];
        cnt[c + 1]++;
    }
    for (int i = 0; i < 5; ++i)
        out[i] = cnt[i] == 0 ? 0.0 : sqrt(xm[i] * xm[i] + ym[i] * ym[i]) / (double)cnt[i];
}

/* Build-defined: phase of direction j is phase_constant * g_j . (M + W x0)
 * evaluated in float with fmaf, then the reference's FP64 cos/sin sum. */
void orc_moment_signal(uint32_t n, const float *mag, const float *moment, const float *x0, float wsum,
                       const uint8_t *comp, uint32_t ndir, const float *g, float pc, double *out)
{
    for (uint32_t j = 0; j < ndir; ++j) {
        double xm[5] = {0, 0, 0, 0, 0}, ym[5] = {0, 0, 0, 0, 0};
        uint32_t cnt[5] = {0, 0, 0, 0, 0};
        float gx = g[j * 3 + 0], gy = g[j * 3 + 1], gz = g[j * 3 + 2];
        for (uint32_t i = 0; i < n; ++i) {
            float mx = moment[i * 3 + 0], my = moment[i * 3 + 1], mz = moment[i * 3 + 2];
            if (x0 && wsum != 0.0f) {
                mx = fmaf(wsum, x0[i * 3 + 0], mx);
                my = fmaf(wsum, x0[i * 3 + 1], my);
                mz = fmaf(wsum, x0[i * 3 + 2], mz);
            }
            float ph = fmaf(gx, mx, fmaf(gy, my, gz * mz)) * pc;
            double c = cos((double)ph) * (double)mag[i], s = sin((double)ph) * (double)mag[i];
            uint32_t k = comp[i];
            xm[0] += c; ym[0] += s; cnt[0]++;
            xm[k + 1] += c; ym[k + 1] += s; cnt[k + 1]++;
        }
        for (int i = 0; i < 5; ++i)
            out[j * 5 + i] = cnt[i] == 0 ? 0.0 : sqrt(xm[i] * xm[i] + ym[i] * ym[i]) / (double)cnt[i];
    }
}

/* ======================================================================== */
/* Formats and sequencing                                                    */
/* ======================================================================== */

/* initFibers, dSim3/src/spinSystem.cpp:290-673: two passes of a character
 * scanner over the mesh file (V, T, A n, M n, G n sections), counting on the
 * first pass and filling on the second.  Vertices are scaled by 1/spaceScale
 * (:422-424); triInfo = {fiber, membrane (A=0,M=1,G=2), first vertex}. */
int orc_parse_mesh(const char *path, float space_scale, orc_mesh *m)
{
    memset(m, 0, sizeof(*m));
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    float x, y, z, v1, v2, v3, nr, tri;
    int ns;
    uint32_t i, n_fibers = 0, *fibers = NULL, *tricount = NULL;
    float xmax = 1.0f, ymax = 1.0f, zmax = 1.0f;
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            n_fibers = m->n_axon + m->n_glia;
            m->vertices = (float *)calloc((size_t)m->n_vert * 3 + 1, sizeof(float));
            m->trgls = (uint32_t *)calloc((size_t)m->n_tri * 3 + 1, sizeof(uint32_t));
            m->tri_info = (uint32_t *)calloc((size_t)m->n_tri * 3 + 1, sizeof(uint32_t));
            tricount = (uint32_t *)calloc((size_t)3 * n_fibers + 1, sizeof(uint32_t));
            fibers = (uint32_t *)calloc((size_t)3 * n_fibers * m->max_tri_on_surface + 1, sizeof(uint32_t));
        }
        char dt = 'X';
        while (!feof(fp)) {
            ns = fscanf(fp, "%c", &dt);
            (void)ns;
            if (dt == 'V') {
                i = 0;
                ns = 3;
                while (ns == 3) {
                    ns = fscanf(fp, "%g %g %g", &x, &y, &z);
                    if (ns == 3) {
                        if (pass == 0) {
                            m->n_vert++;
                        } else {
                            m->vertices[i * 3 + 0] = x / space_scale;
                            m->vertices[i * 3 + 1] = y / space_scale;
                            m->vertices[i * 3 + 2] = z / space_scale;
                            xmax = fmaxf(fabsf(x) / space_scale, xmax);
                            ymax = fmaxf(fabsf(y) / space_scale, ymax);
                            zmax = fmaxf(fabsf(z) / space_scale, zmax);
                            i++;
                        }
                    }
                }
            } else if (dt == 'T') {
                i = 0;
                ns = 3;
                while (ns == 3) {
                    ns = fscanf(fp, "%g %g %g", &v1, &v2, &v3);
                    if (ns == 3) {
                        if (pass == 0) {
                            m->n_tri++;
                        } else {
                            m->trgls[i * 3 + 0] = (uint32_t)v1;
                            m->trgls[i * 3 + 1] = (uint32_t)v2;
                            m->trgls[i * 3 + 2] = (uint32_t)v3;
                            m->tri_info[i * 3 + 2] = (uint32_t)v1;
                            i++;
                        }
                    }
                }
            } else if (dt == 'A' || dt == 'M' || dt == 'G') {
                uint32_t mem = dt == 'A' ? 0 : dt == 'M' ? 1 : 2;
                if (pass == 0) {
                    if (mem == 0) m->n_axon++; else if (mem == 1) m->n_myelin++; else m->n_glia++;
                }
                nr = 0.0f;
                ns = fscanf(fp, "%g", &nr);
                ns = 1;
                i = 0;
                while (ns == 1) {
                    ns = fscanf(fp, "%g", &tri);
                    if (ns == 1) {
                        if (pass == 1) {
                            uint32_t f = (uint32_t)nr, t = (uint32_t)tri;
                            if (f < n_fibers && i < m->max_tri_on_surface)
                                fibers[(mem * n_fibers + f) * m->max_tri_on_surface + i] = t;
                            if (f < n_fibers) tricount[mem * n_fibers + f] += 1;
                            if (t < m->n_tri) {
                                m->tri_info[t * 3] = f;
                                m->tri_info[t * 3 + 1] = mem;
                            }
                        }
                        i++;
                    }
                }
                if (i > m->max_tri_on_surface) m->max_tri_on_surface = i;
            }
        }
        rewind(fp);
    }
    fclose(fp);
    free(fibers);
    free(tricount);
    m->n_fibers = n_fibers;
    m->xmax = xmax; m->ymax = ymax; m->zmax = zmax;
    if (m->n_axon != m->n_myelin) return -2;
    return 0;
}

void orc_free_mesh(orc_mesh *m)
{
    free(m->vertices); free(m->trgls); free(m->tri_info);
    memset(m, 0, sizeof(*m));
}

/* Gradient file reader, dSim3/src/diffusion.cpp:177-202 (six numbers per
 * line: delta Delta readout gx gy gz).  A token that is not a number would
 * stall the reference's loop forever; here it ends the read. */
int orc_parse_grads(const char *path, float *rows, int cap)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    int n = 0;
    float v[6];
    while (!feof(fp)) {
        int ns = fscanf(fp, "%g %g %g %g %g %g", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5]);
        if (ns == 6) {
            if (n < cap)
                for (int k = 0; k < 6; ++k) rows[n * 6 + k] = v[k];
            n++;
        } else if (ns == 0) {
            break;
        }
    }
    fclose(fp);
    return n;
}

/* Gradient on/off state machine of update(), dSim3/src/diffusion.cpp:564-608:
 * L=(int)(delta/(dt*S)+0.5), B, R likewise; the if/else-if chain is applied
 * at the start of each update k, then S steps run with the current gradient.
 * sign[k] is the lobe sign during update k; the readout happens at the start
 * of update L+B+R unless an earlier branch of the chain claims that update
 * (lbr[3] = 1 if the readout happens). */
int orc_pgse_schedule(float ld, float bd, float ro, float dt, int S, int *sign, int cap, int *lbr)
{
    int L = (int)(ld / (dt * (float)S) + 0.5);
    int B = (int)(bd / (dt * (float)S) + 0.5);
    int R = (int)(ro / (dt * (float)S) + 0.5);
    int n = B + L + R;
    int g = 0, readout = 0;
    for (int k = 0; k <= n; ++k) {
        if (k == 0) g = 1;
        else if (k == L && k < B) g = 0;
        else if (k == B) g = -1;
        else if (k == B + L) g = 0;
        else if (k == B + L + R) readout = 1;
        if (k < n && k < cap) sign[k] = g;
    }
    if (lbr) { lbr[0] = L; lbr[1] = B; lbr[2] = R; lbr[3] = readout; }
    return n;
}
