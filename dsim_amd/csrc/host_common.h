// Shared pieces of the host runtime: error reporting, device buffers and the
// substrate object (used by host.cpp and sim.cpp).
#pragma once

#include <cstdint>
#include <cstring>
#include <vector>

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

#include "../../include/dsim.h"

namespace dsim_host {

// Records a message for dsim_last_error() and returns code.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

// Checks the mesh arrays of dsim_substrate_create (indices, membranes, fibres).
int check_mesh(const dsim_mesh *mesh);
// Axis (0 x, 1 y, 2 z) of the physical mode's labelling rays for a mesh.
uint32_t label_axis_of(const dsim_mesh *mesh);
// Exact clearance refinement of exact-membership grids by default: rings 0-1.
// Their exact distances cost a separating-axis gap per listed triangle of the
// surrounding cells; measured on crossing substrates, refining rings 0-1 keeps
// the walk within 1% of refining 0-3 at a third of the build time, and the
// ring distance alone costs 6-17% of the walk.
constexpr int64_t kTriboxRings = 1;
// Rings of exact clearance refinement of a grid (cells at ring distance at
// most this; -1 none).
int64_t clearance_rings(const dsim_grid_params *g);

// The largest float whose square is at most x >= 0 (the float floor of the
// square root, whatever the rounding of the double square root: the square of
// a float is exact in double).
inline float float_sqrt_floor(double x) {
    if (!(x >= 0.0)) return 0.0f;
    float f = (float)__builtin_sqrt(x);
    while (f > 0.0f && (double)f * (double)f > x) f = __builtin_nextafterf(f, 0.0f);
    for (;;) {
        const float g = __builtin_nextafterf(f, 3.0e38f);
        if ((double)g * (double)g > x) break;
        f = g;
    }
    return f;
}

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    int alloc(size_t count) {
        n = count;
        if (count == 0) return DSIM_OK;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            n = 0;
            return fail(DSIM_ERR_NOMEM, "hipMalloc(%zu bytes): %s", count * sizeof(T),
                        hipGetErrorString(e));
        }
        return DSIM_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

template <class T>
int upload(DevBuf<T> &d, const std::vector<T> &h) {
    int rc = d.alloc(h.size());
    if (rc) return rc;
    if (!h.empty()) {
        hipError_t e = hipMemcpy(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(DSIM_ERR_HIP, "upload: %s", hipGetErrorString(e));
    }
    return DSIM_OK;
}

}  // namespace dsim_host

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return dsim_host::fail(DSIM_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

struct dsim_substrate {
    uint32_t n_tri = 0, n_vert = 0, n_fibers = 0;
    std::vector<float> helpers;  // [n_tri*12]
    std::vector<uint32_t> cell_start, cell_tri;
    uint32_t n_cells = 0, max_per_cell = 0;
    uint32_t kind = 0;
    uint32_t label_axis = 0;  // ray axis of the physical-mode labeller (0 x, 1 y, 2 z)
    uint32_t dims[3] = {0, 0, 0};
    float lo[3] = {0, 0, 0}, cell[3] = {0, 0, 0}, inv_cell[3] = {0, 0, 0};
    std::vector<float> clearance;  // per cell: lower bound of the distance to any triangle
    dsim_host::DevBuf<float4> tri_ref, tri_mt, tri_aabb, cell_plane, cell_mt;
    dsim_host::DevBuf<uint32_t> d_cell_start, d_cell_tri;
    dsim_host::DevBuf<float> d_clearance, d_helpers;
    dsim_host::DevBuf<uint8_t> fiber_myelin;
    int refs = 0;
    bool on_device = false;
    // built by dsim_substrate_create_device: the grid, helpers and clearance
    // live on the device only (the host vectors stay empty)
    bool device_built = false;
    uint32_t n_entries = 0;
    ~dsim_substrate() {
        tri_ref.release(); tri_mt.release(); tri_aabb.release(); cell_plane.release(); cell_mt.release();
        d_clearance.release(); d_helpers.release();
        d_cell_start.release(); d_cell_tri.release(); fiber_myelin.release();
    }
};
