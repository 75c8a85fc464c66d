// Host runtime of the dsim C ABI (include/dsim.h): substrate preparation,
// simulation handles, state transfer and kernel dispatch.
//
// Built with -ffp-contract=off: the triangle records and grid membership are
// computed with exactly the float operations the reference uses
// (spinSystem.cpp:612-648, 1094-1379) so the device sees the same numbers.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../../include/dsim.h"
#include "dsim_internal.h"
#include "host_common.h"
#include "tribox.h"

namespace dsim_host {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

}  // namespace dsim_host

using dsim_host::DevBuf;
using dsim_host::fail;
using dsim_host::upload;

namespace {

struct V3 {
    float x, y, z;
};
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
inline float dotv(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

inline float bits_to_float(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

unsigned n_workers() {
    unsigned n = std::thread::hardware_concurrency();
    if (n == 0) n = 1;
    return std::min(n, 16u);
}

// ---------------------------------------------------------------------------
// Triangle records
// ---------------------------------------------------------------------------

// m_triangleHelpers of spinSystem.cpp:612-648.
void helpers_of(const float *V, const uint32_t *T, float *h) {
    uint32_t p1 = T[0], p2 = T[1], p3 = T[2];
    float x1 = V[p1 * 3 + 0], y1 = V[p1 * 3 + 1], z1 = V[p1 * 3 + 2];
    float x2 = V[p2 * 3 + 0], y2 = V[p2 * 3 + 1], z2 = V[p2 * 3 + 2];
    float x3 = V[p3 * 3 + 0], y3 = V[p3 * 3 + 1], z3 = V[p3 * 3 + 2];
    float nx = (y2 - y1) * (z3 - z1) - (z2 - z1) * (y3 - y1);
    float ny = (z2 - z1) * (x3 - x1) - (x2 - x1) * (z3 - z1);
    float nz = (x2 - x1) * (y3 - y1) - (y2 - y1) * (x3 - x1);
    float len = std::sqrt(nx * nx + ny * ny + nz * nz);
    h[0] = nx / len; h[1] = ny / len; h[2] = nz / len;
    h[3] = x3 - x1; h[4] = y3 - y1; h[5] = z3 - z1;
    h[6] = x2 - x1; h[7] = y2 - y1; h[8] = z2 - z1;
    h[9] = (x3 - x1) * (x2 - x1) + (y3 - y1) * (y2 - y1) + (z3 - z1) * (z2 - z1);
    h[10] = (x3 - x1) * (x3 - x1) + (y3 - y1) * (y3 - y1) + (z3 - z1) * (z3 - z1);
    h[11] = (x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1) + (z2 - z1) * (z2 - z1);
}

// ---------------------------------------------------------------------------
// Reference grid membership (populateCubes, spinSystem.cpp:1094-1379)
// ---------------------------------------------------------------------------

bool point_in_box(V3 p, V3 mn, V3 mx) {
    return ((p.x <= mx.x) && (p.x >= mn.x)) && ((p.y <= mx.y) && (p.y >= mn.y)) &&
           ((p.z <= mx.z) && (p.z >= mn.z));
}

// rayTriangleIntersect, spinSystem.cpp:1360-1379.
bool segment_hits_triangle(V3 r1, V3 r2, V3 t1, V3 t2, V3 t3) {
    V3 u = t2 - t1, v = t3 - t1;
    V3 n = v3(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
    float nl = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
    float s = 2.0f, t = 2.0f;
    n.x = n.x / nl; n.y = n.y / nl; n.z = n.z / nl;
    float r = dotv(n, t1 - r1) / dotv(n, r2 - r1);
    if ((0 < r) & (r < 1)) {
        V3 w = (r1 + (r2 - r1) * r) - t1;
        float uv = dotv(u, v), wv = dotv(w, v), vv = dotv(v, v), wu = dotv(w, u), uu = dotv(u, u);
        s = (uv * wv - vv * wu) / (uv * uv - uu * vv);
        t = (uv * wu - uu * wv) / (uv * uv - uu * vv);
    }
    return (s >= 0) && (t >= 0) && (s + t <= 1);
}

// Corner c of a box: bit 0 picks max x, bit 1 max y, bit 2 max z.  The edge
// and face tables follow the argument order of spinSystem.cpp:1189-1244.
inline V3 corner(V3 mn, V3 mx, int c) {
    return v3((c & 1) ? mx.x : mn.x, (c & 2) ? mx.y : mn.y, (c & 4) ? mx.z : mn.z);
}
const int kBoxEdges[12][2] = {{0, 1}, {0, 2}, {0, 4}, {6, 2}, {6, 4}, {6, 7},
                              {3, 2}, {3, 1}, {3, 7}, {5, 7}, {5, 1}, {5, 4}};
const int kBoxFaces[12][3] = {{0, 1, 2}, {0, 4, 2}, {0, 4, 1}, {6, 2, 4}, {6, 2, 7}, {6, 4, 7},
                              {3, 2, 1}, {3, 7, 1}, {3, 7, 2}, {5, 4, 7}, {5, 1, 7}, {5, 4, 1}};

bool triangle_in_box(V3 p1, V3 p2, V3 p3, V3 mn, V3 mx) {
    if (point_in_box(p1, mn, mx) || point_in_box(p2, mn, mx) || point_in_box(p3, mn, mx))
        return true;
    for (const auto &e : kBoxEdges)
        if (segment_hits_triangle(corner(mn, mx, e[0]), corner(mn, mx, e[1]), p1, p2, p3))
            return true;
    const V3 a[3] = {p1, p1, p2}, b[3] = {p2, p3, p3};
    for (int s = 0; s < 3; ++s)
        for (const auto &f : kBoxFaces)
            if (segment_hits_triangle(a[s], b[s], corner(mn, mx, f[0]), corner(mn, mx, f[1]),
                                      corner(mn, mx, f[2])))
                return true;
    return false;
}

// calcCubePos with the reference's division and std::min/std::max clamp.
inline uint32_t ref_cell(float p, float lo, float cell, uint32_t n) {
    float q = std::floor((p - lo) / cell);
    float hi = (float)(n - 1);
    float mn = (hi < q) ? hi : q;
    float mx = (0.0f < mn) ? mn : 0.0f;
    return (uint32_t)mx;
}

struct Pair {
    uint32_t cell, tri;
};

// Runs fn(t, out) for every triangle over worker threads; the pair lists are
// concatenated in triangle order.
template <class Fn>
std::vector<Pair> collect_pairs(uint32_t n_tri, Fn fn) {
    unsigned nw = std::max(1u, std::min(n_workers(), (n_tri + 255) / 256));
    std::vector<std::vector<Pair>> parts(nw);
    std::vector<std::thread> th;
    for (unsigned w = 0; w < nw; ++w) {
        th.emplace_back([&, w] {
            uint32_t a = (uint32_t)((uint64_t)n_tri * w / nw);
            uint32_t b = (uint32_t)((uint64_t)n_tri * (w + 1) / nw);
            for (uint32_t t = a; t < b; ++t) fn(t, parts[w]);
        });
    }
    for (auto &t : th) t.join();
    size_t total = 0;
    for (auto &p : parts) total += p.size();
    std::vector<Pair> all;
    all.reserve(total);
    for (auto &p : parts) all.insert(all.end(), p.begin(), p.end());
    return all;
}

// Stable counting sort by cell: lists keep ascending triangle order.
void pairs_to_csr(const std::vector<Pair> &pairs, uint32_t n_cells, std::vector<uint32_t> &start,
                  std::vector<uint32_t> &tri, uint32_t &max_per_cell) {
    start.assign((size_t)n_cells + 1, 0);
    for (const Pair &p : pairs) start[p.cell + 1]++;
    max_per_cell = 0;
    for (uint32_t c = 0; c < n_cells; ++c) {
        max_per_cell = std::max(max_per_cell, start[c + 1]);
        start[c + 1] += start[c];
    }
    tri.assign(pairs.size(), 0);
    std::vector<uint32_t> fill(start.begin(), start.end() - 1);
    for (const Pair &p : pairs) tri[fill[p.cell]++] = p.tri;
}

// Vertices of triangle t in double.
inline void tri_vertices(const float *V, const uint32_t *T, uint32_t t, double a[3], double b[3],
                         double c[3]) {
    for (int k = 0; k < 3; ++k) {
        a[k] = V[T[t * 3] * 3 + k];
        b[k] = V[T[t * 3 + 1] * 3 + k];
        c[k] = V[T[t * 3 + 2] * 3 + k];
    }
}

// Exact-membership test of triangle t against cell idx (its box in double
// from the grid origin and cell edges).
inline bool tri_cell_separated(const float *V, const uint32_t *T, uint32_t t, const float lo[3],
                               const float ce[3], const int64_t idx[3], float eps) {
    double a[3], b[3], c[3], b0[3], b1[3];
    tri_vertices(V, T, t, a, b, c);
    for (int k = 0; k < 3; ++k) {
        b0[k] = (double)lo[k] + (double)idx[k] * (double)ce[k];
        b1[k] = (double)lo[k] + (double)(idx[k] + 1) * (double)ce[k];
    }
    dsim_host::TriBoxAxes X;
    dsim_host::tribox_setup(a, b, c, b0, b1, X);
    return dsim_host::tribox_separated(X, (double)eps);
}

// Per-cell clearance: a lower bound of the distance from any point of the cell
// to any triangle registered in the grid.  A breadth-first pass over the
// 26-neighbourhood gives each cell its Chebyshev ring distance D to the
// nearest cell holding a triangle; every triangle is then at least
// (D - 1) * (smallest cell edge) away.  Cells within `exact_rings` of a
// triangle refine that with the smallest box-to-box distance between the
// cell and the bounding boxes of the triangles listed in the rings around it
// (growing Chebyshev radius, stopping once a ring is farther than the best
// found).  Rounded down to float; 1e30 when the grid holds no triangle.  The
// device build (grid.hip) computes the same values.
// dist2(t, b0, b1) is a lower bound of the squared distance between triangle
// t and the box [b0, b1]: the distance between their bounding boxes (AABB
// grids) or the largest separating-axis gap (exact-membership grids).
template <class Dist2>
void compute_clearance(const uint32_t dims[3], const float lo[3], const float cell[3],
                       const std::vector<uint32_t> &start, const std::vector<uint32_t> &tri,
                       Dist2 dist2, int64_t exact_rings, std::vector<float> &out) {
    constexpr uint32_t kFar = 0xFFFFFFFFu;
    const uint64_t nc = (uint64_t)dims[0] * dims[1] * dims[2];
    out.assign(nc, 0.0f);
    const double hmin = std::min({(double)cell[0], (double)cell[1], (double)cell[2]});
    const int64_t D[3] = {dims[0], dims[1], dims[2]};
    // ring distance to the nearest occupied cell
    std::vector<uint32_t> ring(nc, kFar);
    std::vector<uint64_t> frontier, next;
    for (uint64_t c = 0; c < nc; ++c)
        if (start[c + 1] > start[c]) { ring[c] = 0; frontier.push_back(c); }
    for (uint32_t r = 1; !frontier.empty(); ++r) {
        next.clear();
        for (uint64_t c : frontier) {
            const int64_t ix = c % D[0], iy = (c / D[0]) % D[1], iz = c / ((uint64_t)D[0] * D[1]);
            for (int64_t z = std::max<int64_t>(0, iz - 1); z <= std::min(D[2] - 1, iz + 1); ++z)
                for (int64_t y = std::max<int64_t>(0, iy - 1); y <= std::min(D[1] - 1, iy + 1); ++y)
                    for (int64_t x = std::max<int64_t>(0, ix - 1); x <= std::min(D[0] - 1, ix + 1); ++x) {
                        const uint64_t cc = ((uint64_t)z * D[1] + y) * D[0] + x;
                        if (ring[cc] == kFar) { ring[cc] = r; next.push_back(cc); }
                    }
        }
        frontier.swap(next);
    }
    unsigned nw = std::max(1u, std::min(n_workers(), (unsigned)((nc + 4095) / 4096)));
    std::vector<std::thread> th;
    for (unsigned w = 0; w < nw; ++w) {
        th.emplace_back([&, w] {
            for (uint64_t c = nc * w / nw; c < nc * (w + 1) / nw; ++c) {
                if (ring[c] == kFar) { out[c] = 1e30f; continue; }
                // triangles lie only in rings >= ring[c], each at least
                // (ring - 1) * hmin from this cell
                double best = ring[c] > 0 ? (double)(ring[c] - 1) * hmin : 0.0;
                if ((int64_t)ring[c] <= exact_rings) {
                    int64_t ix = c % D[0], iy = (c / D[0]) % D[1], iz = c / ((uint64_t)D[0] * D[1]);
                    int64_t idx[3] = {ix, iy, iz};
                    double b0[3], b1[3];
                    for (int k = 0; k < 3; ++k) {
                        b0[k] = (double)lo[k] + (double)idx[k] * cell[k];
                        b1[k] = b0[k] + cell[k];
                    }
                    // squared distances; the float result is the largest float
                    // whose square is at most the minimum
                    double best2 = 1e60;
                    for (int64_t r = ring[c];; ++r) {
                        if (r > 0) {
                            const double q = (double)(r - 1) * hmin;
                            if (q * q >= best2) break;
                        }
                        if (r > D[0] && r > D[1] && r > D[2]) break;
                        for (int64_t z = std::max<int64_t>(0, iz - r); z <= std::min(D[2] - 1, iz + r); ++z)
                            for (int64_t y = std::max<int64_t>(0, iy - r); y <= std::min(D[1] - 1, iy + r); ++y)
                                for (int64_t x = std::max<int64_t>(0, ix - r); x <= std::min(D[0] - 1, ix + r); ++x) {
                                    if (std::max({std::llabs(x - ix), std::llabs(y - iy), std::llabs(z - iz)}) != r)
                                        continue;  // only the ring at Chebyshev radius r
                                    uint64_t cc = ((uint64_t)z * D[1] + y) * D[0] + x;
                                    for (uint32_t k = start[cc]; k < start[cc + 1]; ++k)
                                        best2 = std::min(best2, dist2(tri[k], b0, b1));
                                }
                    }
                    out[c] = dsim_host::float_sqrt_floor(best2);
                    continue;
                }
                float f = (float)best;
                if ((double)f > best) f = std::nextafter(f, 0.0f);
                out[c] = f;
            }
        });
    }
    for (auto &t : th) t.join();
}

}  // namespace

// ---------------------------------------------------------------------------
// Substrate
// ---------------------------------------------------------------------------

int dsim_host::check_mesh(const dsim_mesh *mesh) {
    if (mesh->n_triangles && (!mesh->vertices || !mesh->triangles || !mesh->tri_info))
        return fail(DSIM_ERR_ARG, "mesh arrays missing");
    const uint32_t nT = mesh->n_triangles, nV = mesh->n_vertices;
    for (uint64_t k = 0; k < (uint64_t)nT * 3; ++k)
        if (mesh->triangles[k] >= nV)
            return fail(DSIM_ERR_ARG, "triangle %llu references vertex %u of %u",
                        (unsigned long long)(k / 3), mesh->triangles[k], nV);
    for (uint32_t t = 0; t < nT; ++t) {
        if (mesh->tri_info[t * 3 + 1] > 2) return fail(DSIM_ERR_ARG, "triangle %u membrane > 2", t);
        if (mesh->tri_info[t * 3] >= DSIM_NO_FIBER)
            return fail(DSIM_ERR_ARG, "triangle %u fibre index too large", t);
        if (mesh->tri_info[t * 3 + 2] >= nV) return fail(DSIM_ERR_ARG, "triangle %u P1 bad", t);
    }
    return DSIM_OK;
}

// Rings of exact clearance refinement (cells at ring distance <= this):
// dsim_grid_params.clearance_rings - 2, or the grid kind's default.
int64_t dsim_host::clearance_rings(const dsim_grid_params *g) {
    if (g->clearance_rings == 0) return g->kind == DSIM_GRID_TRIBOX ? kTriboxRings : 3;
    return (int64_t)g->clearance_rings - 2;
}

// The axis of the physical mode's labelling rays.  A ray that runs along an
// open tube never crosses its wall, and one with a component along a cut
// tube can leave through its open end, so the ray must be perpendicular to
// the mesh's long walls: per axis, sum the triangles' extents along it, and
// take the smallest.  +x is kept unless its sum exceeds the smaller of the
// other two by more than 5% (z-aligned prisms, capped or not, keep +x).
uint32_t dsim_host::label_axis_of(const dsim_mesh *mesh) {
    const float *V = mesh->vertices;
    const uint32_t *T = mesh->triangles;
    double score[3] = {0, 0, 0};
    for (uint32_t t = 0; t < mesh->n_triangles; ++t)
        for (int k = 0; k < 3; ++k) {
            const double a = V[T[t * 3] * 3 + k], b = V[T[t * 3 + 1] * 3 + k],
                         c = V[T[t * 3 + 2] * 3 + k];
            score[k] += std::max({a, b, c}) - std::min({a, b, c});
        }
    const double other = std::min(score[1], score[2]);
    if (score[0] <= 1.05 * other) return 0;
    return score[2] < score[1] ? 2 : 1;
}

extern "C" const char *dsim_last_error(void) { return dsim_host::g_err.c_str(); }
extern "C" int dsim_abi_version(void) { return DSIM_ABI_VERSION; }

extern "C" int dsim_device_count(int *n) {
    if (!n) return fail(DSIM_ERR_ARG, "n is NULL");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    *n = (e == hipSuccess) ? c : 0;
    return DSIM_OK;
}

extern "C" int dsim_substrate_create(const dsim_mesh *mesh, const dsim_grid_params *grid,
                                     dsim_substrate **out) {
    if (!mesh || !grid || !out) return fail(DSIM_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (int rc = dsim_host::check_mesh(mesh)) return rc;
    const uint32_t nT = mesh->n_triangles, nV = mesh->n_vertices;
    int rc = DSIM_OK;
    int ndev = 0;
    const bool have_device = hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0;

    auto *S = new dsim_substrate();
    S->n_tri = nT;
    S->n_vert = nV;
    S->kind = grid->kind;
    const float *V = mesh->vertices;
    const uint32_t *T = mesh->triangles;
    uint32_t nf = 0;
    for (uint32_t t = 0; t < nT; ++t) nf = std::max(nf, mesh->tri_info[t * 3] + 1);
    S->n_fibers = nf;

    // Triangle helpers and the two 64-byte record layouts.
    S->helpers.resize((size_t)nT * 12);
    std::vector<float4> ref((size_t)nT * 4), mt((size_t)nT * 4), box((size_t)nT * 2);
    for (uint32_t t = 0; t < nT; ++t) {
        float *h = &S->helpers[(size_t)t * 12];
        helpers_of(V, T + (size_t)t * 3, h);
        uint32_t info = (mesh->tri_info[t * 3] & 0xFFFFu) | (mesh->tri_info[t * 3 + 1] << 16);
        uint32_t P1 = mesh->tri_info[t * 3 + 2];
        ref[t * 4 + 0] = make_float4(h[0], h[1], h[2], h[3]);
        ref[t * 4 + 1] = make_float4(h[4], h[5], h[6], h[7]);
        ref[t * 4 + 2] = make_float4(h[8], h[9], h[10], h[11]);
        ref[t * 4 + 3] = make_float4(V[P1 * 3], V[P1 * 3 + 1], V[P1 * 3 + 2], bits_to_float(info));
        uint32_t a = T[t * 3], b = T[t * 3 + 1], c = T[t * 3 + 2];
        float v0x = V[a * 3], v0y = V[a * 3 + 1], v0z = V[a * 3 + 2];
        float nd = std::fmaf(h[0], v0x, std::fmaf(h[1], v0y, h[2] * v0z));
        mt[t * 4 + 0] = make_float4(h[0], h[1], h[2], nd);
        mt[t * 4 + 1] = make_float4(v0x, v0y, v0z, bits_to_float(info));
        const float e1x = V[b * 3] - v0x, e1y = V[b * 3 + 1] - v0y, e1z = V[b * 3 + 2] - v0z;
        const float e2x = V[c * 3] - v0x, e2y = V[c * 3 + 1] - v0y, e2z = V[c * 3 + 2] - v0z;
        // 1e-6 |e1 x e2|^2: a segment d with det^2 < this * |d|^2 lies within
        // about 1e-3 rad of the plane and is re-tested in FP64
        const float cx = std::fmaf(e1y, e2z, -(e1z * e2y));
        const float cy = std::fmaf(e1z, e2x, -(e1x * e2z));
        const float cz = std::fmaf(e1x, e2y, -(e1y * e2x));
        const float graze = 1e-6f * std::fmaf(cx, cx, std::fmaf(cy, cy, cz * cz));
        mt[t * 4 + 2] = make_float4(e1x, e1y, e1z, graze);
        mt[t * 4 + 3] = make_float4(e2x, e2y, e2z, 0.f);
        float mn[3], mx[3];
        for (int k = 0; k < 3; ++k) {
            float p = V[a * 3 + k], q = V[b * 3 + k], r = V[c * 3 + k];
            mn[k] = p; if (q < mn[k]) mn[k] = q; if (r < mn[k]) mn[k] = r;
            mx[k] = p; if (q > mx[k]) mx[k] = q; if (r > mx[k]) mx[k] = r;
        }
        box[t * 2 + 0] = make_float4(mn[0], mn[1], mn[2], bits_to_float(info));
        box[t * 2 + 1] = make_float4(mx[0], mx[1], mx[2], 0.f);
    }
    std::vector<uint8_t> myel(std::max(1u, nf), 0);
    for (uint32_t t = 0; t < nT; ++t)
        if (mesh->tri_info[t * 3 + 1] == 1) myel[mesh->tri_info[t * 3]] = 1;

    // Uniform grid.
    std::vector<Pair> pairs;
    if (grid->kind == DSIM_GRID_REFERENCE) {
        uint32_t n = grid->dims[0];
        if (n == 0 || n > 1024) { delete S; return fail(DSIM_ERR_ARG, "reference grid n=%u", n); }
        float L = 2.0f / (float)n;  // m_cubeLength, spinSystem.cpp:1115
        for (int k = 0; k < 3; ++k) {
            S->dims[k] = n; S->lo[k] = -1.0f; S->cell[k] = L; S->inv_cell[k] = 1.0f / L;
        }
        pairs = collect_pairs(nT, [&](uint32_t t, std::vector<Pair> &o) {
            V3 p[3];
            for (int j = 0; j < 3; ++j) {
                uint32_t vi = T[t * 3 + j];
                p[j] = v3(V[vi * 3], V[vi * 3 + 1], V[vi * 3 + 2]);
            }
            V3 mn = p[0], mx = p[0];
            if (p[1].x < mn.x) mn.x = p[1].x; if (p[2].x < mn.x) mn.x = p[2].x;
            if (p[1].y < mn.y) mn.y = p[1].y; if (p[2].y < mn.y) mn.y = p[2].y;
            if (p[1].z < mn.z) mn.z = p[1].z; if (p[2].z < mn.z) mn.z = p[2].z;
            if (p[1].x > mx.x) mx.x = p[1].x; if (p[2].x > mx.x) mx.x = p[2].x;
            if (p[1].y > mx.y) mx.y = p[1].y; if (p[2].y > mx.y) mx.y = p[2].y;
            if (p[1].z > mx.z) mx.z = p[1].z; if (p[2].z > mx.z) mx.z = p[2].z;
            uint32_t c0[3] = {ref_cell(mn.x, -1.0f, L, n), ref_cell(mn.y, -1.0f, L, n),
                              ref_cell(mn.z, -1.0f, L, n)};
            uint32_t c1[3] = {ref_cell(mx.x, -1.0f, L, n), ref_cell(mx.y, -1.0f, L, n),
                              ref_cell(mx.z, -1.0f, L, n)};
            for (uint32_t xc = c0[0]; xc <= c1[0]; ++xc)
                for (uint32_t yc = c0[1]; yc <= c1[1]; ++yc)
                    for (uint32_t zc = c0[2]; zc <= c1[2]; ++zc) {
                        V3 bmin = v3(-1 + (float)xc * L, -1 + (float)yc * L, -1 + (float)zc * L);
                        V3 bmax = v3(-1 + (float)(xc + 1) * L, -1 + (float)(yc + 1) * L,
                                     -1 + (float)(zc + 1) * L);
                        if (triangle_in_box(p[0], p[1], p[2], bmin, bmax))
                            o.push_back(Pair{(zc * n + yc) * n + xc, t});
                    }
        });
    } else if (grid->kind == DSIM_GRID_AABB || grid->kind == DSIM_GRID_TRIBOX) {
        const bool exact = grid->kind == DSIM_GRID_TRIBOX;
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
        const float eps = grid->eps;
        const uint32_t *dm = S->dims;
        const float *lo = S->lo, *ce = S->cell;
        pairs = collect_pairs(nT, [&](uint32_t t, std::vector<Pair> &o) {
            float mn[3], mx[3];
            int64_t i0[3], i1[3];
            for (int k = 0; k < 3; ++k) {
                float a = V[T[t * 3] * 3 + k], b = V[T[t * 3 + 1] * 3 + k], c = V[T[t * 3 + 2] * 3 + k];
                mn[k] = a; if (b < mn[k]) mn[k] = b; if (c < mn[k]) mn[k] = c;
                mx[k] = a; if (b > mx[k]) mx[k] = b; if (c > mx[k]) mx[k] = c;
                mn[k] = mn[k] - eps; mx[k] = mx[k] + eps;
                i0[k] = (int64_t)std::floor((mn[k] - lo[k]) / ce[k]) - 1;
                i1[k] = (int64_t)std::floor((mx[k] - lo[k]) / ce[k]) + 1;
                if (i0[k] < 0) i0[k] = 0;
                if (i1[k] > (int64_t)dm[k] - 1) i1[k] = (int64_t)dm[k] - 1;
                if (i0[k] > i1[k]) return;
            }
            for (int64_t z = i0[2]; z <= i1[2]; ++z) {
                float bz0 = lo[2] + (float)z * ce[2], bz1 = lo[2] + (float)(z + 1) * ce[2];
                if (bz0 > mx[2] || mn[2] > bz1) continue;
                for (int64_t y = i0[1]; y <= i1[1]; ++y) {
                    float by0 = lo[1] + (float)y * ce[1], by1 = lo[1] + (float)(y + 1) * ce[1];
                    if (by0 > mx[1] || mn[1] > by1) continue;
                    for (int64_t x = i0[0]; x <= i1[0]; ++x) {
                        float bx0 = lo[0] + (float)x * ce[0], bx1 = lo[0] + (float)(x + 1) * ce[0];
                        if (bx0 > mx[0] || mn[0] > bx1) continue;
                        if (exact) {
                            const int64_t idx[3] = {x, y, z};
                            if (tri_cell_separated(V, T, t, lo, ce, idx, eps)) continue;
                        }
                        o.push_back(Pair{(uint32_t)((z * dm[1] + y) * dm[0] + x), t});
                    }
                }
            }
        });
    } else {
        delete S;
        return fail(DSIM_ERR_ARG, "unknown grid kind %u", grid->kind);
    }
    S->n_cells = S->dims[0] * S->dims[1] * S->dims[2];
    pairs_to_csr(pairs, S->n_cells, S->cell_start, S->cell_tri, S->max_per_cell);
    if (grid->kind == DSIM_GRID_AABB) {
        compute_clearance(S->dims, S->lo, S->cell, S->cell_start, S->cell_tri,
                          [&](uint32_t t, const double b0[3], const double b1[3]) {
                              const float4 a = box[(size_t)t * 2], b = box[(size_t)t * 2 + 1];
                              double lo3[3] = {a.x, a.y, a.z}, hi3[3] = {b.x, b.y, b.z}, d2 = 0;
                              for (int q = 0; q < 3; ++q) {
                                  double g = std::max({0.0, lo3[q] - b1[q], b0[q] - hi3[q]});
                                  d2 += g * g;
                              }
                              return d2;
                          },
                          dsim_host::clearance_rings(grid), S->clearance);
    } else if (grid->kind == DSIM_GRID_TRIBOX) {
        compute_clearance(S->dims, S->lo, S->cell, S->cell_start, S->cell_tri,
                          [&](uint32_t t, const double b0[3], const double b1[3]) {
                              double a[3], b[3], c[3];
                              tri_vertices(V, T, t, a, b, c);
                              dsim_host::TriBoxAxes X;
                              dsim_host::tribox_setup(a, b, c, b0, b1, X);
                              return dsim_host::tribox_gap2(X);
                          },
                          dsim_host::clearance_rings(grid), S->clearance);
    }
    S->label_axis = dsim_host::label_axis_of(mesh);

    if (!have_device) {  // host-only substrate: grid and helpers are inspectable, no sims
        *out = S;
        return DSIM_OK;
    }
    S->on_device = true;
    if ((rc = upload(S->tri_ref, ref)) || (rc = upload(S->tri_mt, mt)) ||
        (rc = upload(S->tri_aabb, box)) || (rc = upload(S->d_cell_start, S->cell_start)) ||
        (rc = upload(S->fiber_myelin, myel)) || (rc = upload(S->d_clearance, S->clearance))) {
        delete S;
        return rc;
    }
    std::vector<uint32_t> ct = S->cell_tri;
    if (ct.empty()) ct.push_back(0);  // keep a valid pointer for empty grids
    // Per grid entry, in cell order: the plane of its triangle (the
    // prefilter of a cell's candidates reads contiguous memory), and the
    // triangle's Moller-Trumbore vertices with its index in v0.w, so the test
    // needs no second lookup through cell_tri.
    // Exact-membership grids list long triangles in many cells: their
    // entries keep only the index and the step reads the triangle's record.
    const bool cell_major = grid->kind != DSIM_GRID_TRIBOX;
    std::vector<float4> cp(cell_major ? ct.size() : 0), cm(cell_major ? ct.size() * 3 : 0);
    for (size_t k = 0; k < cp.size(); ++k) {
        if (S->cell_tri.empty()) break;
        const float4 *r = &mt[(size_t)ct[k] * 4];
        cp[k] = r[0];
        cm[k * 3 + 0] = float4{r[1].x, r[1].y, r[1].z, bits_to_float(ct[k])};
        cm[k * 3 + 1] = r[2];
        cm[k * 3 + 2] = r[3];
    }
    if ((rc = upload(S->d_cell_tri, ct)) || (cell_major && (rc = upload(S->cell_plane, cp))) ||
        (cell_major && (rc = upload(S->cell_mt, cm)))) {
        delete S;
        return rc;
    }
    *out = S;
    return DSIM_OK;
}

extern "C" int dsim_substrate_destroy(dsim_substrate *s) {
    if (!s) return DSIM_OK;
    if (s->refs > 0) return fail(DSIM_ERR_STATE, "substrate still used by %d sims", s->refs);
    delete s;
    return DSIM_OK;
}

extern "C" int dsim_substrate_grid_info(const dsim_substrate *s, uint32_t *n_cells,
                                        uint32_t *n_entries, uint32_t *max_per_cell) {
    if (!s) return fail(DSIM_ERR_ARG, "NULL substrate");
    if (n_cells) *n_cells = s->n_cells;
    if (n_entries) *n_entries = s->device_built ? s->n_entries : (uint32_t)s->cell_tri.size();
    if (max_per_cell) *max_per_cell = s->max_per_cell;
    return DSIM_OK;
}

extern "C" int dsim_substrate_get_grid(const dsim_substrate *s, uint32_t *cell_start,
                                       uint32_t *cell_tri) {
    if (!s) return fail(DSIM_ERR_ARG, "NULL substrate");
    if (s->device_built) {
        if (cell_start)
            HIP_TRY(hipMemcpy(cell_start, s->d_cell_start.p, ((size_t)s->n_cells + 1) * 4,
                              hipMemcpyDeviceToHost));
        if (cell_tri && s->n_entries)
            HIP_TRY(hipMemcpy(cell_tri, s->d_cell_tri.p, (size_t)s->n_entries * 4,
                              hipMemcpyDeviceToHost));
        return DSIM_OK;
    }
    if (cell_start) memcpy(cell_start, s->cell_start.data(), s->cell_start.size() * 4);
    if (cell_tri && !s->cell_tri.empty()) memcpy(cell_tri, s->cell_tri.data(), s->cell_tri.size() * 4);
    return DSIM_OK;
}

extern "C" int dsim_substrate_get_clearance(const dsim_substrate *s, float *clearance) {
    if (!s || !clearance) return fail(DSIM_ERR_ARG, "NULL argument");
    if (s->device_built && s->d_clearance.p) {
        HIP_TRY(hipMemcpy(clearance, s->d_clearance.p, (size_t)s->n_cells * 4,
                          hipMemcpyDeviceToHost));
        return DSIM_OK;
    }
    if (s->clearance.empty()) return fail(DSIM_ERR_STATE, "the grid has no clearance table");
    memcpy(clearance, s->clearance.data(), s->clearance.size() * 4);
    return DSIM_OK;
}

extern "C" int dsim_substrate_get_helpers(const dsim_substrate *s, float *helpers) {
    if (!s || !helpers) return fail(DSIM_ERR_ARG, "NULL argument");
    if (s->device_built) {
        if (s->n_tri)
            HIP_TRY(hipMemcpy(helpers, s->d_helpers.p, (size_t)s->n_tri * 48, hipMemcpyDeviceToHost));
        return DSIM_OK;
    }
    if (!s->helpers.empty()) memcpy(helpers, s->helpers.data(), s->helpers.size() * 4);
    return DSIM_OK;
}

extern "C" int dsim_alloc_array(void **dev_ptr, size_t bytes) {
    if (!dev_ptr) return fail(DSIM_ERR_ARG, "NULL pointer");
    *dev_ptr = nullptr;
    if (bytes == 0) return DSIM_OK;
    hipError_t e = hipMalloc(dev_ptr, bytes);
    if (e != hipSuccess) {
        *dev_ptr = nullptr;
        return fail(DSIM_ERR_NOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    }
    return DSIM_OK;
}

extern "C" int dsim_free_array(void *dev_ptr) {
    if (dev_ptr) HIP_TRY(hipFree(dev_ptr));
    return DSIM_OK;
}

extern "C" int dsim_thread_sync(void) {
    HIP_TRY(hipDeviceSynchronize());
    return DSIM_OK;
}

extern "C" int dsim_copy_from_device(void *host, const void *device, size_t bytes) {
    if (bytes == 0) return DSIM_OK;
    if (!host || !device) return fail(DSIM_ERR_ARG, "NULL pointer");
    HIP_TRY(hipMemcpy(host, device, bytes, hipMemcpyDeviceToHost));
    return DSIM_OK;
}

extern "C" int dsim_copy_to_device(void *device, const void *host, size_t offset, size_t bytes) {
    if (bytes == 0) return DSIM_OK;
    if (!host || !device) return fail(DSIM_ERR_ARG, "NULL pointer");
    HIP_TRY(hipMemcpy(static_cast<char *>(device) + offset, host, bytes, hipMemcpyHostToDevice));
    return DSIM_OK;
}

extern "C" int dsim_substrate_label_axis(const dsim_substrate *s, uint32_t *axis) {
    if (!s || !axis) return fail(DSIM_ERR_ARG, "NULL argument");
    *axis = s->label_axis;
    return DSIM_OK;
}

extern "C" int dsim_substrate_bytes(const dsim_substrate *s, uint64_t *bytes) {
    if (!s || !bytes) return fail(DSIM_ERR_ARG, "NULL argument");
    *bytes = (uint64_t)s->tri_ref.n * 16 + s->tri_mt.n * 16 + s->tri_aabb.n * 16 + s->cell_plane.n * 16 +
             s->cell_mt.n * 16 +
             s->d_cell_start.n * 4 + s->d_cell_tri.n * 4 + s->fiber_myelin.n + s->d_clearance.n * 4 +
             s->d_helpers.n * 4;
    return DSIM_OK;
}

// ---------------------------------------------------------------------------
// Simulation handle
// ---------------------------------------------------------------------------
