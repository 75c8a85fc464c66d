// Device build of a substrate (dsim_substrate_create_device, include/dsim.h):
// triangle records, the uniform grid as CSR with its cell-major copies and the
// per-cell clearance, computed on the GPU with the arithmetic of the host build
// (host.cpp) in the same order, so both builds give identical arrays.
//
//   k_tri_records   one thread per triangle: reference helpers
//                   (spinSystem.cpp:612-648), the 64-byte reference and
//                   Moller-Trumbore records, the bounding box
//   k_pairs<false>  per triangle, the number of cells it is listed in
//                   (reference grid: populateCubes' box test,
//                   spinSystem.cpp:1094-1352; AABB grid: box overlap)
//   exclusive scan  -> each triangle's first pair
//   k_pairs<true>   the (cell, triangle) pairs in triangle order, and the
//                   per-cell counts
//   radix sort      stable by cell: lists in ascending triangle order
//   exclusive scan  of the counts -> cell_start
//   k_cell_records  cell-major plane and triangle records
//   k_ring_*        Chebyshev ring distance to the nearest occupied cell
//   k_clearance     the clearance of each cell
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/dsim.h"
#include "host_common.h"
#include "tribox.h"

using dsim_host::DevBuf;
using dsim_host::fail;

namespace {

constexpr uint32_t kFar = 0xFFFFFFFFu;
constexpr uint32_t kBlock = 256;

__device__ __forceinline__ int64_t imin(int64_t a, int64_t b) { return a < b ? a : b; }
__device__ __forceinline__ int64_t imax(int64_t a, int64_t b) { return a > b ? a : b; }

struct Geom {
    uint32_t kind;
    uint32_t dims[3];
    float lo[3], cell[3];
    float eps;
    float L;  // reference grid: cube edge 2/n
    int64_t exact_rings;  // cells at ring distance <= this refine the clearance
};

struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dotv(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// m_triangleHelpers, spinSystem.cpp:612-648.
__device__ void helpers_of(const float *V, const uint32_t *T, float *h) {
    const uint32_t p1 = T[0], p2 = T[1], p3 = T[2];
    const float x1 = V[p1 * 3 + 0], y1 = V[p1 * 3 + 1], z1 = V[p1 * 3 + 2];
    const float x2 = V[p2 * 3 + 0], y2 = V[p2 * 3 + 1], z2 = V[p2 * 3 + 2];
    const float x3 = V[p3 * 3 + 0], y3 = V[p3 * 3 + 1], z3 = V[p3 * 3 + 2];
    const float nx = (y2 - y1) * (z3 - z1) - (z2 - z1) * (y3 - y1);
    const float ny = (z2 - z1) * (x3 - x1) - (x2 - x1) * (z3 - z1);
    const float nz = (x2 - x1) * (y3 - y1) - (y2 - y1) * (x3 - x1);
    const float len = sqrtf(nx * nx + ny * ny + nz * nz);
    h[0] = nx / len; h[1] = ny / len; h[2] = nz / len;
    h[3] = x3 - x1; h[4] = y3 - y1; h[5] = z3 - z1;
    h[6] = x2 - x1; h[7] = y2 - y1; h[8] = z2 - z1;
    h[9] = (x3 - x1) * (x2 - x1) + (y3 - y1) * (y2 - y1) + (z3 - z1) * (z2 - z1);
    h[10] = (x3 - x1) * (x3 - x1) + (y3 - y1) * (y3 - y1) + (z3 - z1) * (z3 - z1);
    h[11] = (x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1) + (z2 - z1) * (z2 - z1);
}

__global__ __launch_bounds__(kBlock) void k_tri_records(uint32_t n_tri, const float *V,
                                                        const uint32_t *T, const uint32_t *info3,
                                                        float *helpers, float4 *ref, float4 *mt,
                                                        float4 *box) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n_tri) return;
    float h[12];
    helpers_of(V, T + (size_t)t * 3, h);
    for (int k = 0; k < 12; ++k) helpers[(size_t)t * 12 + k] = h[k];
    const uint32_t info = (info3[t * 3] & 0xFFFFu) | (info3[t * 3 + 1] << 16);
    const uint32_t P1 = info3[t * 3 + 2];
    ref[t * 4 + 0] = make_float4(h[0], h[1], h[2], h[3]);
    ref[t * 4 + 1] = make_float4(h[4], h[5], h[6], h[7]);
    ref[t * 4 + 2] = make_float4(h[8], h[9], h[10], h[11]);
    ref[t * 4 + 3] = make_float4(V[P1 * 3], V[P1 * 3 + 1], V[P1 * 3 + 2], __uint_as_float(info));
    const uint32_t a = T[t * 3], b = T[t * 3 + 1], c = T[t * 3 + 2];
    const float v0x = V[a * 3], v0y = V[a * 3 + 1], v0z = V[a * 3 + 2];
    const float nd = fmaf(h[0], v0x, fmaf(h[1], v0y, h[2] * v0z));
    mt[t * 4 + 0] = make_float4(h[0], h[1], h[2], nd);
    mt[t * 4 + 1] = make_float4(v0x, v0y, v0z, __uint_as_float(info));
    const float e1x = V[b * 3] - v0x, e1y = V[b * 3 + 1] - v0y, e1z = V[b * 3 + 2] - v0z;
    const float e2x = V[c * 3] - v0x, e2y = V[c * 3 + 1] - v0y, e2z = V[c * 3 + 2] - v0z;
    const float cx = fmaf(e1y, e2z, -(e1z * e2y));
    const float cy = fmaf(e1z, e2x, -(e1x * e2z));
    const float cz = fmaf(e1x, e2y, -(e1y * e2x));
    const float graze = 1e-6f * fmaf(cx, cx, fmaf(cy, cy, cz * cz));
    mt[t * 4 + 2] = make_float4(e1x, e1y, e1z, graze);
    mt[t * 4 + 3] = make_float4(e2x, e2y, e2z, 0.f);
    float mn[3], mx[3];
    for (int k = 0; k < 3; ++k) {
        const float p = V[a * 3 + k], q = V[b * 3 + k], r = V[c * 3 + k];
        mn[k] = p; if (q < mn[k]) mn[k] = q; if (r < mn[k]) mn[k] = r;
        mx[k] = p; if (q > mx[k]) mx[k] = q; if (r > mx[k]) mx[k] = r;
    }
    box[t * 2 + 0] = make_float4(mn[0], mn[1], mn[2], __uint_as_float(info));
    box[t * 2 + 1] = make_float4(mx[0], mx[1], mx[2], 0.f);
}

// ---- reference grid membership (populateCubes, spinSystem.cpp:1094-1379) ----

__device__ __forceinline__ bool point_in_box(F3 p, F3 mn, F3 mx) {
    return ((p.x <= mx.x) && (p.x >= mn.x)) && ((p.y <= mx.y) && (p.y >= mn.y)) &&
           ((p.z <= mx.z) && (p.z >= mn.z));
}

// rayTriangleIntersect, spinSystem.cpp:1360-1379.
__device__ bool segment_hits_triangle(F3 r1, F3 r2, F3 t1, F3 t2, F3 t3) {
    const F3 u = t2 - t1, v = t3 - t1;
    F3 n = f3(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
    const float nl = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
    float s = 2.0f, t = 2.0f;
    n.x = n.x / nl; n.y = n.y / nl; n.z = n.z / nl;
    const float r = dotv(n, t1 - r1) / dotv(n, r2 - r1);
    if ((0 < r) & (r < 1)) {
        const F3 w = (r1 + (r2 - r1) * r) - t1;
        const float uv = dotv(u, v), wv = dotv(w, v), vv = dotv(v, v), wu = dotv(w, u),
                    uu = dotv(u, u);
        s = (uv * wv - vv * wu) / (uv * uv - uu * vv);
        t = (uv * wu - uu * wv) / (uv * uv - uu * vv);
    }
    return (s >= 0) && (t >= 0) && (s + t <= 1);
}

// Corner c of a box: bit 0 picks max x, bit 1 max y, bit 2 max z; edges and
// faces in the argument order of spinSystem.cpp:1189-1244.
__device__ __forceinline__ F3 corner(F3 mn, F3 mx, int c) {
    return f3((c & 1) ? mx.x : mn.x, (c & 2) ? mx.y : mn.y, (c & 4) ? mx.z : mn.z);
}
__constant__ int kBoxEdges[12][2] = {{0, 1}, {0, 2}, {0, 4}, {6, 2}, {6, 4}, {6, 7},
                                     {3, 2}, {3, 1}, {3, 7}, {5, 7}, {5, 1}, {5, 4}};
__constant__ int kBoxFaces[12][3] = {{0, 1, 2}, {0, 4, 2}, {0, 4, 1}, {6, 2, 4},
                                     {6, 2, 7}, {6, 4, 7}, {3, 2, 1}, {3, 7, 1},
                                     {3, 7, 2}, {5, 4, 7}, {5, 1, 7}, {5, 4, 1}};

__device__ bool triangle_in_box(F3 p1, F3 p2, F3 p3, F3 mn, F3 mx) {
    if (point_in_box(p1, mn, mx) || point_in_box(p2, mn, mx) || point_in_box(p3, mn, mx))
        return true;
    for (int e = 0; e < 12; ++e)
        if (segment_hits_triangle(corner(mn, mx, kBoxEdges[e][0]), corner(mn, mx, kBoxEdges[e][1]),
                                  p1, p2, p3))
            return true;
    const F3 a[3] = {p1, p1, p2}, b[3] = {p2, p3, p3};
    for (int s = 0; s < 3; ++s)
        for (int f = 0; f < 12; ++f)
            if (segment_hits_triangle(a[s], b[s], corner(mn, mx, kBoxFaces[f][0]),
                                      corner(mn, mx, kBoxFaces[f][1]),
                                      corner(mn, mx, kBoxFaces[f][2])))
                return true;
    return false;
}

// calcCubePos with the reference's division and min/max clamp.
__device__ __forceinline__ uint32_t ref_cell(float p, float lo, float cell, uint32_t n) {
    const float q = floorf((p - lo) / cell);
    const float hi = (float)(n - 1);
    const float mn = (hi < q) ? hi : q;
    const float mx = (0.0f < mn) ? mn : 0.0f;
    return (uint32_t)mx;
}

// The cells triangle t is listed in, in any order: fn(cell) for each.
template <class Fn>
__device__ void cells_of(const Geom &g, const float *V, const uint32_t *T, uint32_t t, Fn fn) {
    if (g.kind == DSIM_GRID_REFERENCE) {
        const uint32_t n = g.dims[0];
        const float L = g.L;
        F3 p[3];
        for (int j = 0; j < 3; ++j) {
            const uint32_t vi = T[t * 3 + j];
            p[j] = f3(V[vi * 3], V[vi * 3 + 1], V[vi * 3 + 2]);
        }
        F3 mn = p[0], mx = p[0];
        if (p[1].x < mn.x) mn.x = p[1].x; if (p[2].x < mn.x) mn.x = p[2].x;
        if (p[1].y < mn.y) mn.y = p[1].y; if (p[2].y < mn.y) mn.y = p[2].y;
        if (p[1].z < mn.z) mn.z = p[1].z; if (p[2].z < mn.z) mn.z = p[2].z;
        if (p[1].x > mx.x) mx.x = p[1].x; if (p[2].x > mx.x) mx.x = p[2].x;
        if (p[1].y > mx.y) mx.y = p[1].y; if (p[2].y > mx.y) mx.y = p[2].y;
        if (p[1].z > mx.z) mx.z = p[1].z; if (p[2].z > mx.z) mx.z = p[2].z;
        const uint32_t c0[3] = {ref_cell(mn.x, -1.0f, L, n), ref_cell(mn.y, -1.0f, L, n),
                                ref_cell(mn.z, -1.0f, L, n)};
        const uint32_t c1[3] = {ref_cell(mx.x, -1.0f, L, n), ref_cell(mx.y, -1.0f, L, n),
                                ref_cell(mx.z, -1.0f, L, n)};
        for (uint32_t xc = c0[0]; xc <= c1[0]; ++xc)
            for (uint32_t yc = c0[1]; yc <= c1[1]; ++yc)
                for (uint32_t zc = c0[2]; zc <= c1[2]; ++zc) {
                    const F3 bmin = f3(-1 + (float)xc * L, -1 + (float)yc * L, -1 + (float)zc * L);
                    const F3 bmax = f3(-1 + (float)(xc + 1) * L, -1 + (float)(yc + 1) * L,
                                       -1 + (float)(zc + 1) * L);
                    if (triangle_in_box(p[0], p[1], p[2], bmin, bmax)) fn((zc * n + yc) * n + xc);
                }
        return;
    }
    float mn[3], mx[3];
    int64_t i0[3], i1[3];
    for (int k = 0; k < 3; ++k) {
        const float a = V[T[t * 3] * 3 + k], b = V[T[t * 3 + 1] * 3 + k],
                    c = V[T[t * 3 + 2] * 3 + k];
        mn[k] = a; if (b < mn[k]) mn[k] = b; if (c < mn[k]) mn[k] = c;
        mx[k] = a; if (b > mx[k]) mx[k] = b; if (c > mx[k]) mx[k] = c;
        mn[k] = mn[k] - g.eps; mx[k] = mx[k] + g.eps;
        i0[k] = (int64_t)floorf((mn[k] - g.lo[k]) / g.cell[k]) - 1;
        i1[k] = (int64_t)floorf((mx[k] - g.lo[k]) / g.cell[k]) + 1;
        if (i0[k] < 0) i0[k] = 0;
        if (i1[k] > (int64_t)g.dims[k] - 1) i1[k] = (int64_t)g.dims[k] - 1;
        if (i0[k] > i1[k]) return;
    }
    for (int64_t z = i0[2]; z <= i1[2]; ++z) {
        const float bz0 = g.lo[2] + (float)z * g.cell[2], bz1 = g.lo[2] + (float)(z + 1) * g.cell[2];
        if (bz0 > mx[2] || mn[2] > bz1) continue;
        for (int64_t y = i0[1]; y <= i1[1]; ++y) {
            const float by0 = g.lo[1] + (float)y * g.cell[1],
                        by1 = g.lo[1] + (float)(y + 1) * g.cell[1];
            if (by0 > mx[1] || mn[1] > by1) continue;
            for (int64_t x = i0[0]; x <= i1[0]; ++x) {
                const float bx0 = g.lo[0] + (float)x * g.cell[0],
                            bx1 = g.lo[0] + (float)(x + 1) * g.cell[0];
                if (bx0 > mx[0] || mn[0] > bx1) continue;
                if (g.kind == DSIM_GRID_TRIBOX) {
                    // exact membership: host.cpp tri_cell_separated
                    double a[3], b[3], c[3], b0[3], b1[3];
                    const int64_t idx[3] = {x, y, z};
                    for (int k = 0; k < 3; ++k) {
                        a[k] = V[T[t * 3] * 3 + k];
                        b[k] = V[T[t * 3 + 1] * 3 + k];
                        c[k] = V[T[t * 3 + 2] * 3 + k];
                        b0[k] = (double)g.lo[k] + (double)idx[k] * (double)g.cell[k];
                        b1[k] = (double)g.lo[k] + (double)(idx[k] + 1) * (double)g.cell[k];
                    }
                    dsim_host::TriBoxAxes X;
                    dsim_host::tribox_setup(a, b, c, b0, b1, X);
                    if (dsim_host::tribox_separated(X, (double)g.eps)) continue;
                }
                fn((uint32_t)((z * g.dims[1] + y) * g.dims[0] + x));
            }
        }
    }
}

// EMIT false: count[t] = pairs of triangle t.  EMIT true: the pairs of t at
// offset[t] (cell keys, triangle values) and the per-cell counts.
template <bool EMIT>
__global__ __launch_bounds__(kBlock) void k_pairs(Geom g, uint32_t n_tri, const float *V,
                                                  const uint32_t *T, uint32_t *count_or_offset,
                                                  uint32_t *keys, uint32_t *vals,
                                                  uint32_t *cell_count) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n_tri) return;
    if (!EMIT) {
        uint32_t n = 0;
        cells_of(g, V, T, t, [&](uint32_t) { ++n; });
        count_or_offset[t] = n;
    } else {
        uint32_t o = count_or_offset[t];
        cells_of(g, V, T, t, [&](uint32_t c) {
            keys[o] = c;
            vals[o] = t;
            ++o;
            atomicAdd(&cell_count[c], 1u);
        });
    }
}

// Per grid entry in cell order: the plane of its triangle, and v0 (with the
// triangle index in .w), e1, e2 of its Moller-Trumbore record.
__global__ __launch_bounds__(kBlock) void k_cell_records(uint32_t n_entries, const uint32_t *tri,
                                                         const float4 *mt, float4 *plane,
                                                         float4 *cm) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n_entries) return;
    const uint32_t t = tri[k];
    const float4 *r = mt + (size_t)t * 4;
    plane[k] = r[0];
    const float4 v0 = r[1];
    cm[(size_t)k * 3 + 0] = make_float4(v0.x, v0.y, v0.z, __uint_as_float(t));
    cm[(size_t)k * 3 + 1] = r[2];
    cm[(size_t)k * 3 + 2] = r[3];
}

__global__ __launch_bounds__(kBlock) void k_ring_init(uint32_t n_cells, const uint32_t *start,
                                                      uint32_t *ring) {
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n_cells) return;
    ring[c] = start[c + 1] > start[c] ? 0u : kFar;
}

// Cells still unreached that touch (26-neighbourhood) a cell of ring r-1 get
// ring r: after step r every ring value up to r is final.
__global__ __launch_bounds__(kBlock) void k_ring_step(uint32_t dx, uint32_t dy, uint32_t dz,
                                                      uint32_t *ring, uint32_t r,
                                                      uint32_t *changed) {
    const uint64_t c = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= (uint64_t)dx * dy * dz) return;
    if (ring[c] != kFar) return;
    const int64_t ix = c % dx, iy = (c / dx) % dy, iz = c / ((uint64_t)dx * dy);
    for (int64_t z = imax(0, iz - 1); z <= imin((int64_t)dz - 1, iz + 1); ++z)
        for (int64_t y = imax(0, iy - 1); y <= imin((int64_t)dy - 1, iy + 1); ++y)
            for (int64_t x = imax(0, ix - 1); x <= imin((int64_t)dx - 1, ix + 1); ++x)
                if (ring[((uint64_t)z * dy + y) * dx + x] == r - 1) {
                    ring[c] = r;
                    *changed = 1u;
                    return;
                }
}

// The largest float whose square is at most x >= 0.
__device__ float float_sqrt_floor(double x) {
    if (!(x >= 0.0)) return 0.0f;
    float f = (float)sqrt(x);
    while (f > 0.0f && (double)f * (double)f > x) f = nextafterf(f, 0.0f);
    for (;;) {
        const float g = nextafterf(f, 3.0e38f);
        if ((double)g * (double)g > x) break;
        f = g;
    }
    return f;
}

// Clearance of each cell, as compute_clearance in host.cpp: from the ring
// distance alone beyond g.exact_rings, else the smallest box-to-box distance to
// the triangles of the rings around the cell.
__global__ __launch_bounds__(kBlock) void k_clearance(Geom g, const uint32_t *ring,
                                                      const uint32_t *start, const uint32_t *tri,
                                                      const float4 *box, const float *V,
                                                      const uint32_t *T, float *out) {
    const int64_t D[3] = {g.dims[0], g.dims[1], g.dims[2]};
    const uint64_t c = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= (uint64_t)D[0] * D[1] * D[2]) return;
    const uint32_t rc = ring[c];
    if (rc == kFar) { out[c] = 1e30f; return; }
    double hmin = (double)g.cell[0];
    if ((double)g.cell[1] < hmin) hmin = (double)g.cell[1];
    if ((double)g.cell[2] < hmin) hmin = (double)g.cell[2];
    if ((int64_t)rc > g.exact_rings) {
        // an occupied cell (ring 0) has no clearance
        const double best = rc > 0 ? (double)(rc - 1) * hmin : 0.0;
        float f = (float)best;
        if ((double)f > best) f = nextafterf(f, 0.0f);
        out[c] = f;
        return;
    }
    const int64_t ix = c % D[0], iy = (c / D[0]) % D[1], iz = c / ((uint64_t)D[0] * D[1]);
    const int64_t idx[3] = {ix, iy, iz};
    double b0[3], b1[3];
    for (int k = 0; k < 3; ++k) {
        b0[k] = (double)g.lo[k] + (double)idx[k] * g.cell[k];
        b1[k] = b0[k] + g.cell[k];
    }
    double best2 = 1e60;
    for (int64_t r = rc;; ++r) {
        if (r > 0) {
            const double q = (double)(r - 1) * hmin;
            if (q * q >= best2) break;
        }
        if (r > D[0] && r > D[1] && r > D[2]) break;
        for (int64_t z = imax(0, iz - r); z <= imin(D[2] - 1, iz + r); ++z)
            for (int64_t y = imax(0, iy - r); y <= imin(D[1] - 1, iy + r); ++y)
                for (int64_t x = imax(0, ix - r); x <= imin(D[0] - 1, ix + r); ++x) {
                    const int64_t ax = x > ix ? x - ix : ix - x, ay = y > iy ? y - iy : iy - y,
                                  az = z > iz ? z - iz : iz - z;
                    if (imax(imax(ax, ay), az) != r) continue;  // the ring at radius r only
                    const uint64_t cc = ((uint64_t)z * D[1] + y) * D[0] + x;
                    for (uint32_t k = start[cc]; k < start[cc + 1]; ++k) {
                        double d2 = 0;
                        if (g.kind == DSIM_GRID_TRIBOX) {
                            // host.cpp: the largest separating-axis gap
                            const uint32_t t = tri[k];
                            double a[3], b[3], c[3];
                            for (int q = 0; q < 3; ++q) {
                                a[q] = V[T[t * 3] * 3 + q];
                                b[q] = V[T[t * 3 + 1] * 3 + q];
                                c[q] = V[T[t * 3 + 2] * 3 + q];
                            }
                            dsim_host::TriBoxAxes X;
                            dsim_host::tribox_setup(a, b, c, b0, b1, X);
                            d2 = dsim_host::tribox_gap2(X);
                        } else {
                            const float4 a = box[(size_t)tri[k] * 2], b = box[(size_t)tri[k] * 2 + 1];
                            const double lo3[3] = {a.x, a.y, a.z}, hi3[3] = {b.x, b.y, b.z};
                            for (int q = 0; q < 3; ++q) {
                                double gap = 0.0;
                                if (lo3[q] - b1[q] > gap) gap = lo3[q] - b1[q];
                                if (b0[q] - hi3[q] > gap) gap = b0[q] - hi3[q];
                                d2 += gap * gap;
                            }
                        }
                        if (d2 < best2) best2 = d2;
                    }
                }
    }
    out[c] = float_sqrt_floor(best2);
}

uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

}  // namespace

extern "C" int dsim_substrate_create_device(const dsim_mesh *mesh, const dsim_grid_params *grid,
                                            dsim_substrate **out) {
    if (!mesh || !grid || !out) return fail(DSIM_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (int rc = dsim_host::check_mesh(mesh)) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(DSIM_ERR_STATE, "dsim_substrate_create_device needs a HIP device");
    const uint32_t nT = mesh->n_triangles, nV = mesh->n_vertices;
    Geom g{};
    g.kind = grid->kind;
    auto *S = new dsim_substrate();
    S->n_tri = nT;
    S->n_vert = nV;
    S->kind = grid->kind;
    if (grid->kind == DSIM_GRID_REFERENCE) {
        const uint32_t n = grid->dims[0];
        if (n == 0 || n > 1024) { delete S; return fail(DSIM_ERR_ARG, "reference grid n=%u", n); }
        const float L = 2.0f / (float)n;  // m_cubeLength, spinSystem.cpp:1115
        g.L = L;
        for (int k = 0; k < 3; ++k) {
            S->dims[k] = n; S->lo[k] = -1.0f; S->cell[k] = L; S->inv_cell[k] = 1.0f / L;
        }
    } else if (grid->kind == DSIM_GRID_AABB || grid->kind == DSIM_GRID_TRIBOX) {
        for (int k = 0; k < 3; ++k) {
            if (grid->dims[k] == 0 || !(grid->cell[k] > 0.0f)) {
                delete S;
                return fail(DSIM_ERR_ARG, "AABB grid dims/cell invalid on axis %d", k);
            }
            S->dims[k] = grid->dims[k]; S->lo[k] = grid->lo[k]; S->cell[k] = grid->cell[k];
            S->inv_cell[k] = 1.0f / grid->cell[k];
        }
        if ((uint64_t)S->dims[0] * S->dims[1] * S->dims[2] > (1ull << 28)) {
            delete S;
            return fail(DSIM_ERR_ARG, "grid too large");
        }
        g.eps = grid->eps;
        g.exact_rings = dsim_host::clearance_rings(grid);
    } else {
        delete S;
        return fail(DSIM_ERR_ARG, "unknown grid kind %u", grid->kind);
    }
    for (int k = 0; k < 3; ++k) { g.dims[k] = S->dims[k]; g.lo[k] = S->lo[k]; g.cell[k] = S->cell[k]; }
    S->n_cells = S->dims[0] * S->dims[1] * S->dims[2];
    uint32_t nf = 0;
    for (uint32_t t = 0; t < nT; ++t) nf = std::max(nf, mesh->tri_info[t * 3] + 1);
    S->n_fibers = nf;
    std::vector<uint8_t> myel(std::max(1u, nf), 0);
    for (uint32_t t = 0; t < nT; ++t)
        if (mesh->tri_info[t * 3 + 1] == 1) myel[mesh->tri_info[t * 3]] = 1;
    S->on_device = true;
    S->device_built = true;
    S->label_axis = dsim_host::label_axis_of(mesh);

    int rc = DSIM_OK;
    const hipStream_t st = nullptr;
    DevBuf<float> dV;
    DevBuf<uint32_t> dT, dI, cnt, offs, keys, vals, keys2, vals2, ccount, ring, flag, maxc;
    DevBuf<uint8_t> temp;
    auto cleanup = [&] {
        dV.release(); dT.release(); dI.release(); cnt.release(); offs.release(); keys.release();
        vals.release();
        keys2.release(); vals2.release(); ccount.release(); ring.release(); flag.release();
        maxc.release(); temp.release();
    };
    auto bail = [&](int code) {
        cleanup();
        delete S;
        return code;
    };
#define GRID_TRY(expr)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return bail(fail(DSIM_ERR_HIP, "device grid build: %s: %s", #expr,              \
                             hipGetErrorString(e_)));                                       \
    } while (0)

    // mesh upload and triangle records
    const size_t nt1 = std::max<uint32_t>(nT, 1);
    if ((rc = dV.alloc(std::max<size_t>((size_t)nV * 3, 1))) || (rc = dT.alloc(nt1 * 3)) ||
        (rc = dI.alloc(nt1 * 3)) || (rc = S->d_helpers.alloc(nt1 * 12)) ||
        (rc = S->tri_ref.alloc(nt1 * 4)) || (rc = S->tri_mt.alloc(nt1 * 4)) ||
        (rc = S->tri_aabb.alloc(nt1 * 2)) || (rc = cnt.alloc(nt1 + 1)) || (rc = offs.alloc(nt1 + 1)) ||
        (rc = S->d_cell_start.alloc((size_t)S->n_cells + 1)) ||
        (rc = ccount.alloc((size_t)S->n_cells + 1)) || (rc = maxc.alloc(1)) ||
        (rc = flag.alloc(1)))
        return bail(rc);
    if (nV) GRID_TRY(hipMemcpy(dV.p, mesh->vertices, (size_t)nV * 12, hipMemcpyHostToDevice));
    if (nT) {
        GRID_TRY(hipMemcpy(dT.p, mesh->triangles, (size_t)nT * 12, hipMemcpyHostToDevice));
        GRID_TRY(hipMemcpy(dI.p, mesh->tri_info, (size_t)nT * 12, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_tri_records, dim3(blocks_for(nT)), dim3(kBlock), 0, st, nT, dV.p,
                           dT.p, dI.p, S->d_helpers.p, S->tri_ref.p, S->tri_mt.p, S->tri_aabb.p);
        GRID_TRY(hipGetLastError());
    }
    if ((rc = dsim_host::upload(S->fiber_myelin, myel))) return bail(rc);

    // pairs per triangle, their offsets, the pairs, the per-cell counts
    GRID_TRY(hipMemsetAsync(cnt.p, 0, (nt1 + 1) * 4, st));
    GRID_TRY(hipMemsetAsync(ccount.p, 0, ((size_t)S->n_cells + 1) * 4, st));
    if (nT) {
        hipLaunchKernelGGL(k_pairs<false>, dim3(blocks_for(nT)), dim3(kBlock), 0, st, g, nT, dV.p,
                           dT.p, cnt.p, (uint32_t *)nullptr, (uint32_t *)nullptr,
                           (uint32_t *)nullptr);
        GRID_TRY(hipGetLastError());
    }
    // exclusive scan over nT + 1 counts: offsets, and the total in the last
    size_t tb_scan1 = 0, tb_scan2 = 0, tb_sort = 0, tb_red = 0;
    GRID_TRY(rocprim::exclusive_scan(nullptr, tb_scan1, cnt.p, offs.p, 0u, nT + 1,
                                     rocprim::plus<uint32_t>(), st));
    GRID_TRY(rocprim::exclusive_scan(nullptr, tb_scan2, ccount.p, S->d_cell_start.p, 0u,
                                     (size_t)S->n_cells + 1, rocprim::plus<uint32_t>(), st));
    GRID_TRY(rocprim::reduce(nullptr, tb_red, ccount.p, maxc.p, 0u, (size_t)S->n_cells,
                             rocprim::maximum<uint32_t>(), st));
    const size_t tb0 = std::max({tb_scan1, tb_scan2, tb_red, (size_t)1});
    if ((rc = temp.alloc(tb0))) return bail(rc);
    {
        size_t b = tb_scan1;
        GRID_TRY(rocprim::exclusive_scan(temp.p, b, cnt.p, offs.p, 0u, nT + 1,
                                         rocprim::plus<uint32_t>(), st));
    }
    uint32_t n_pairs = 0;
    GRID_TRY(hipMemcpyAsync(&n_pairs, offs.p + nT, 4, hipMemcpyDeviceToHost, st));
    GRID_TRY(hipStreamSynchronize(st));
    const size_t np1 = std::max<uint32_t>(n_pairs, 1);
    if ((rc = keys.alloc(np1)) || (rc = vals.alloc(np1)) || (rc = keys2.alloc(np1)) ||
        (rc = S->d_cell_tri.alloc(np1)))
        return bail(rc);
    if (nT) {
        hipLaunchKernelGGL(k_pairs<true>, dim3(blocks_for(nT)), dim3(kBlock), 0, st, g, nT, dV.p,
                           dT.p, offs.p, keys.p, vals.p, ccount.p);
        GRID_TRY(hipGetLastError());
    }
    // stable sort by cell: each cell's list in ascending triangle order
    unsigned end_bit = 1;
    while (end_bit < 32 && (1ull << end_bit) < S->n_cells) ++end_bit;
    if (n_pairs) {
        GRID_TRY(rocprim::radix_sort_pairs(nullptr, tb_sort, keys.p, keys2.p, vals.p,
                                           S->d_cell_tri.p, n_pairs, 0, end_bit, st));
        if (tb_sort > temp.n) {
            temp.release();
            if ((rc = temp.alloc(tb_sort))) return bail(rc);
        }
        size_t b = tb_sort;
        GRID_TRY(rocprim::radix_sort_pairs(temp.p, b, keys.p, keys2.p, vals.p, S->d_cell_tri.p,
                                           n_pairs, 0, end_bit, st));
    } else {
        GRID_TRY(hipMemsetAsync(S->d_cell_tri.p, 0, 4, st));
    }
    {
        size_t b = tb_scan2;
        GRID_TRY(rocprim::exclusive_scan(temp.p, b, ccount.p, S->d_cell_start.p, 0u,
                                         (size_t)S->n_cells + 1, rocprim::plus<uint32_t>(), st));
        b = tb_red;
        GRID_TRY(rocprim::reduce(temp.p, b, ccount.p, maxc.p, 0u, (size_t)S->n_cells,
                                 rocprim::maximum<uint32_t>(), st));
    }
    GRID_TRY(hipMemcpyAsync(&S->max_per_cell, maxc.p, 4, hipMemcpyDeviceToHost, st));
    S->n_entries = n_pairs;

    // cell-major records (not for exact-membership grids: their long
    // triangles are listed in many cells, so entries keep only the index)
    const bool cell_major = grid->kind != DSIM_GRID_TRIBOX;
    if (cell_major && ((rc = S->cell_plane.alloc(np1)) || (rc = S->cell_mt.alloc(np1 * 3))))
        return bail(rc);
    if (!cell_major) {
    } else if (n_pairs) {
        hipLaunchKernelGGL(k_cell_records, dim3(blocks_for(n_pairs)), dim3(kBlock), 0, st,
                           n_pairs, S->d_cell_tri.p, S->tri_mt.p, S->cell_plane.p, S->cell_mt.p);
        GRID_TRY(hipGetLastError());
    } else {
        GRID_TRY(hipMemsetAsync(S->cell_plane.p, 0, 16, st));
        GRID_TRY(hipMemsetAsync(S->cell_mt.p, 0, 48, st));
    }

    // clearance (AABB and exact-membership grids)
    if (grid->kind != DSIM_GRID_REFERENCE) {
        if ((rc = ring.alloc(S->n_cells)) || (rc = S->d_clearance.alloc(S->n_cells)))
            return bail(rc);
        const uint32_t nb = blocks_for(S->n_cells);
        hipLaunchKernelGGL(k_ring_init, dim3(nb), dim3(kBlock), 0, st, S->n_cells,
                           S->d_cell_start.p, ring.p);
        GRID_TRY(hipGetLastError());
        // rings grow by one per step; the flag is read every 8 steps (steps
        // past the last ring change nothing)
        const uint32_t max_r = std::max({S->dims[0], S->dims[1], S->dims[2]});
        for (uint32_t r = 1; n_pairs && r <= max_r; r += 8) {
            GRID_TRY(hipMemsetAsync(flag.p, 0, 4, st));
            for (uint32_t q = r; q < r + 8; ++q)
                hipLaunchKernelGGL(k_ring_step, dim3(nb), dim3(kBlock), 0, st, S->dims[0],
                                   S->dims[1], S->dims[2], ring.p, q, flag.p);
            GRID_TRY(hipGetLastError());
            uint32_t changed = 0;
            GRID_TRY(hipMemcpyAsync(&changed, flag.p, 4, hipMemcpyDeviceToHost, st));
            GRID_TRY(hipStreamSynchronize(st));
            if (!changed) break;
        }
        hipLaunchKernelGGL(k_clearance, dim3(nb), dim3(kBlock), 0, st, g, ring.p,
                           S->d_cell_start.p, S->d_cell_tri.p, S->tri_aabb.p, dV.p, dT.p,
                           S->d_clearance.p);
        GRID_TRY(hipGetLastError());
    }
    GRID_TRY(hipStreamSynchronize(st));
#undef GRID_TRY
    cleanup();
    *out = S;
    return DSIM_OK;
}
