// Triangle / box separation by the separating-axis theorem (Akenine-Moller,
// "Fast 3D triangle-box overlap testing", JGT 2001): the box's three face
// normals, the triangle's normal and the nine cross products of box edges
// with triangle edges.  Evaluated in double, in one order, by the host and
// the device grid builds (host.cpp, grid.hip), so both make the same
// decisions.  Used for the exact-membership grid (DSIM_GRID_TRIBOX): a
// triangle is listed in a cell unless some axis separates them by more than
// the margin, and the largest separation over the axes is a lower bound of
// their distance (the cell clearance).
#pragma once

#include <hip/hip_runtime_api.h>

namespace dsim_host {

struct TriBoxAxes {
    double v[3][3];  // triangle vertices relative to the box centre
    double h[3];     // box half extents
};

__host__ __device__ inline void tribox_setup(const double a[3], const double b[3],
                                             const double c[3], const double lo[3],
                                             const double hi[3], TriBoxAxes &T) {
    for (int k = 0; k < 3; ++k) {
        const double ctr = 0.5 * (lo[k] + hi[k]);
        T.h[k] = 0.5 * (hi[k] - lo[k]);
        T.v[0][k] = a[k] - ctr;
        T.v[1][k] = b[k] - ctr;
        T.v[2][k] = c[k] - ctr;
    }
}

// Separation along `axis` (not normalised): max(min p - r, -r - max p) with p
// the triangle's projections and r the box's projected half extent; and
// |axis|^2.
__host__ __device__ inline double tribox_sep(const TriBoxAxes &T, const double ax[3], double &aa) {
    const double p0 = ax[0] * T.v[0][0] + ax[1] * T.v[0][1] + ax[2] * T.v[0][2];
    const double p1 = ax[0] * T.v[1][0] + ax[1] * T.v[1][1] + ax[2] * T.v[1][2];
    const double p2 = ax[0] * T.v[2][0] + ax[1] * T.v[2][1] + ax[2] * T.v[2][2];
    double mn = p0, mx = p0;
    if (p1 < mn) mn = p1; if (p2 < mn) mn = p2;
    if (p1 > mx) mx = p1; if (p2 > mx) mx = p2;
    const double r = T.h[0] * (ax[0] < 0 ? -ax[0] : ax[0]) + T.h[1] * (ax[1] < 0 ? -ax[1] : ax[1]) +
                     T.h[2] * (ax[2] < 0 ? -ax[2] : ax[2]);
    aa = ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2];
    const double g1 = mn - r, g2 = -r - mx;
    return g1 > g2 ? g1 : g2;
}

// The thirteen axes in a fixed order: x, y, z, the triangle normal, then
// unit_k x edge_j for edges v1-v0, v2-v1, v0-v2.
__host__ __device__ inline void tribox_axis(const TriBoxAxes &T, int q, double ax[3]) {
    ax[0] = ax[1] = ax[2] = 0.0;
    if (q < 3) { ax[q] = 1.0; return; }
    double e[3][3];
    for (int k = 0; k < 3; ++k) {
        e[0][k] = T.v[1][k] - T.v[0][k];
        e[1][k] = T.v[2][k] - T.v[1][k];
        e[2][k] = T.v[0][k] - T.v[2][k];
    }
    if (q == 3) {
        ax[0] = e[0][1] * e[1][2] - e[0][2] * e[1][1];
        ax[1] = e[0][2] * e[1][0] - e[0][0] * e[1][2];
        ax[2] = e[0][0] * e[1][1] - e[0][1] * e[1][0];
        return;
    }
    const int j = (q - 4) / 3, k = (q - 4) % 3;
    const double *E = e[j];
    if (k == 0) { ax[1] = -E[2]; ax[2] = E[1]; }
    else if (k == 1) { ax[0] = E[2]; ax[2] = -E[0]; }
    else { ax[0] = -E[1]; ax[1] = E[0]; }
}

// True when some axis separates the triangle from the box by more than eps
// (the triangle is then not listed in the cell).
__host__ __device__ inline bool tribox_separated(const TriBoxAxes &T, double eps) {
    for (int q = 0; q < 13; ++q) {
        double ax[3], aa;
        tribox_axis(T, q, ax);
        const double g = tribox_sep(T, ax, aa);
        if (g > 0.0 && aa > 0.0 && g * g > eps * eps * aa) return true;
    }
    return false;
}

// The square of the largest separation over the axes (0 when none
// separates): a lower bound of the squared distance between the two.
__host__ __device__ inline double tribox_gap2(const TriBoxAxes &T) {
    double best = 0.0;
    for (int q = 0; q < 13; ++q) {
        double ax[3], aa;
        tribox_axis(T, q, ax);
        const double g = tribox_sep(T, ax, aa);
        if (g > 0.0 && aa > 0.0) {
            const double d2 = g * g / aa;
            if (d2 > best) best = d2;
        }
    }
    return best;
}

}  // namespace dsim_host
