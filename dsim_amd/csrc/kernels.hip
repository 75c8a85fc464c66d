// HIP kernels of the dsim simulate path (gfx950).
//
// Built with -ffp-contract=off: every float expression rounds as written.
// The reference-mode code keeps the evaluation order of dSim3's CPU path
// (spinKernelCpu.cpp) so that, given identical normals, it reproduces that
// path's hit triangles, cells, labels, positions and phases bit for bit; the
// physical-mode code spells its fused multiply-adds out with fmaf().
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dsim_internal.h"

namespace dsim {

#define TWOPI_F 6.28318530717959f  // spinSystem.cpp:20

// ------------------------------------------------------------------------ //
// RNG
// ------------------------------------------------------------------------ //

// Marsaglia multiply-with-carry, spinKernelCpu.cpp:41-45
__device__ __forceinline__ uint32_t mwc_next(uint32_t &s0, uint32_t &s1) {
    s0 = 36969u * (s0 & 65535u) + (s0 >> 16);
    s1 = 18000u * (s1 & 65535u) + (s1 >> 16);
    return (s0 << 16) + s1;
}

// myRandn + boxMuller, spinKernelCpu.cpp:27-63 (ocml logf/cosf/sinf).
__device__ __forceinline__ void mwc_randn(uint32_t &s0, uint32_t &s1, float &n1, float &n2,
                                          float &n3, float &u) {
    n1 = ((float)mwc_next(s0, s1) + 1.0f) / 4294967296.0f;
    n2 = ((float)mwc_next(s0, s1) + 1.0f) / 4294967296.0f;
    n3 = ((float)mwc_next(s0, s1) + 1.0f) / 4294967296.0f;
    u = ((float)mwc_next(s0, s1) + 1.0f) / 4294967296.0f;
    float r1 = sqrtf(-2.0f * logf(n1));
    float p1 = TWOPI_F * n2;
    float r2 = sqrtf(-2.0f * logf(n3));
    float p2 = TWOPI_F * u;
    n1 = r1 * cosf(p1);
    n2 = r1 * sinf(p1);
    n3 = r2 * cosf(p2);
}

// Philox4x32-10 (Salmon et al., SC'11).  The 32x32->64 products compile to
// one v_mad_u64_u32 each.
__device__ __forceinline__ uint4 philox(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = make_uint4((uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1,
                       (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

__device__ __forceinline__ float u32_unit(uint32_t x) {
    return ((float)(x >> 9) + 0.5f) * 1.1920928955078125e-07f;
}

// One Philox block per spin and step (counter: spin, step lo, step hi, 0):
// Gaussian normals from the top 23 bits of each word with the hardware
// transcendental units (v_log_f32 is log2, v_cos_f32/v_sin_f32 take
// revolutions, so cos(2*pi*u) is one instruction for u in (0,1)), and the
// permeation uniform from 23 of the low bits.
__device__ __forceinline__ void philox_draw(uint32_t k0, uint32_t k1, uint32_t spin,
                                            uint64_t step, float &n0, float &n1, float &n2,
                                            float &u) {
    const uint4 o = philox(make_uint4(spin, (uint32_t)step, (uint32_t)(step >> 32), 0u), k0, k1);
    const float u1 = u32_unit(o.x), u2 = u32_unit(o.y), u3 = u32_unit(o.z), u4 = u32_unit(o.w);
    const float m2ln2 = -1.3862943611198906f;  // -2 ln 2
    const float r1 = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u1));
    const float r2 = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u3));
    n0 = r1 * __builtin_amdgcn_cosf(u2);
    n1 = r1 * __builtin_amdgcn_sinf(u2);
    n2 = r2 * __builtin_amdgcn_cosf(u4);
    const uint32_t m = (o.x & 511u) | ((o.y & 511u) << 9) | ((o.z & 31u) << 18);
    u = ((float)m + 0.5f) * 1.1920928955078125e-07f;
}

// ------------------------------------------------------------------------ //
// Grid
// ------------------------------------------------------------------------ //

// calcCubePos, spinSystem.cpp:691-717: std::min/std::max clamp in float.
__device__ __forceinline__ uint32_t clamp_cell(float q, uint32_t n) {
    float hi = (float)(n - 1);
    float mn = (hi < q) ? hi : q;
    float mx = (0.0f < mn) ? mn : 0.0f;
    return (uint32_t)mx;
}

template <int MODE>
__device__ __forceinline__ void cell_of(const StepParams &P, float x, float y, float z,
                                        uint32_t &cx, uint32_t &cy, uint32_t &cz) {
    if (MODE == 0) {
        cx = clamp_cell(floorf((x - P.lo[0]) / P.cell[0]), P.dims[0]);
        cy = clamp_cell(floorf((y - P.lo[1]) / P.cell[1]), P.dims[1]);
        cz = clamp_cell(floorf((z - P.lo[2]) / P.cell[2]), P.dims[2]);
    } else {
        cx = clamp_cell(floorf((x - P.lo[0]) * P.inv_cell[0]), P.dims[0]);
        cy = clamp_cell(floorf((y - P.lo[1]) * P.inv_cell[1]), P.dims[1]);
        cz = clamp_cell(floorf((z - P.lo[2]) * P.inv_cell[2]), P.dims[2]);
    }
}

__device__ __forceinline__ uint32_t cell_hash(const StepParams &P, uint32_t cx, uint32_t cy,
                                              uint32_t cz) {
    return (cz * P.dims[1] + cy) * P.dims[0] + cx;
}

// ------------------------------------------------------------------------ //
// Ray-triangle tests
// ------------------------------------------------------------------------ //

struct Hit {
    uint32_t type;
    float px, py, pz;
    uint32_t idx;
    float d;  // reference: squared distance; MT: parametric t
};

// triCollDetect, spinKernelCpu.cpp:354-395 (softsurfer test, reference order).
__device__ __forceinline__ void test_ref(const float4 *__restrict__ R, uint32_t t, float ox,
                                         float oy, float oz, float px, float py, float pz, Hit &h) {
    h.type = 0;
    const float4 a = R[t * 4 + 0];
    const float4 d = R[t * 4 + 3];
    float num = a.x * (d.x - ox) + a.y * (d.y - oy) + a.z * (d.z - oz);
    float den = a.x * (px - ox) + a.y * (py - oy) + a.z * (pz - oz);
    // r in (0,1) needs num and den of one sign with |num| <= |den|; the
    // division (and the rest of the record) is only needed then.
    if (!((num > 0.0f && den >= num) || (num < 0.0f && den <= num))) return;
    float r = num / den;
    if (!((0 < r) & (r < 1))) return;
    const float4 b = R[t * 4 + 1];
    const float4 c = R[t * 4 + 2];
    float wu = (ox + r * (px - ox) - d.x) * a.w + (oy + r * (py - oy) - d.y) * b.x +
               (oz + r * (pz - oz) - d.z) * b.y;
    float wv = (ox + r * (px - ox) - d.x) * b.z + (oy + r * (py - oy) - d.y) * b.w +
               (oz + r * (pz - oz) - d.z) * c.x;
    float h9 = c.y, h10 = c.z, h11 = c.w;
    float s = (h9 * wv - h11 * wu) / (h9 * h9 - h10 * h11);
    float tt = (h9 * wu - h10 * wv) / (h9 * h9 - h10 * h11);
    if (!((s >= 0) & (tt >= 0) & (s + tt <= 1))) return;
    h.px = ox + (px - ox) * r;
    h.py = oy + (py - oy) * r;
    h.pz = oz + (pz - oz) * r;
    h.d = (ox - h.px) * (ox - h.px) + (oy - h.py) * (oy - h.py) + (oz - h.pz) * (oz - h.pz);
    h.idx = t;
    h.type = 1;
    if ((s == 0) | (tt == 0) | (s + tt == 1)) {
        h.type = 2;
        if (((s == 0) & (tt == 0)) | ((s == 0) & (tt == 1)) | ((s == 1) & (tt == 0))) h.type = 3;
    }
}

// A segment within about 1e-3 rad of a triangle's plane, det^2 < 1e-6
// |e1 x e2|^2 |d|^2, is decided in FP64; the triangle record carries
// 1e-6 |e1 x e2|^2 in e1.w.

// Plane-side prefilter: do the end points lie on different sides of the plane
// (n, n.v0) as this expression evaluates it?  A point's side is a fixed
// function of its coordinates, so a spin changes side only across a segment
// that passes here.  (Admitting same-side segments with an end point near the
// plane lets a reflection point on the edge shared with a neighbouring
// triangle hit that neighbour at t ~ 0, and back, without end.)
__device__ __forceinline__ bool crosses_plane(const float4 pl, float ox, float oy, float oz,
                                              float px, float py, float pz) {
    float s0 = fmaf(pl.x, ox, fmaf(pl.y, oy, fmaf(pl.z, oz, -pl.w)));
    float s1 = fmaf(pl.x, px, fmaf(pl.y, py, fmaf(pl.z, pz, -pl.w)));
    return (s0 > 0.0f) != (s1 > 0.0f);
}

__device__ __forceinline__ float seg_len2(float ox, float oy, float oz, float px, float py,
                                          float pz) {
    const float dx = px - ox, dy = py - oy, dz = pz - oz;
    return fmaf(dx, dx, fmaf(dy, dy, dz * dz));
}

// Moller-Trumbore in FP64 for a near-grazing segment.  The differences of
// float coordinates are exact in double; the inside tests compare the
// numerators u*det, v*det, t*det with det after making det positive, so
// they involve no division; t and the hit point are rounded back to float.
__device__ __forceinline__ void mt_fp64(const float4 v0, const float4 e1, const float4 e2,
                                        uint32_t t, float ox, float oy, float oz, float px,
                                        float py, float pz, Hit &h) {
    const double Dx = (double)px - (double)ox, Dy = (double)py - (double)oy,
                 Dz = (double)pz - (double)oz;
    const double Tx = (double)ox - (double)v0.x, Ty = (double)oy - (double)v0.y,
                 Tz = (double)oz - (double)v0.z;
    const double Qx = Dy * (double)e2.z - Dz * (double)e2.y;
    const double Qy = Dz * (double)e2.x - Dx * (double)e2.z;
    const double Qz = Dx * (double)e2.y - Dy * (double)e2.x;
    double det = (double)e1.x * Qx + (double)e1.y * Qy + (double)e1.z * Qz;
    if (det == 0.0) return;
    const double s = det < 0.0 ? -1.0 : 1.0;
    det *= s;
    const double un = s * (Tx * Qx + Ty * Qy + Tz * Qz);
    if (!(un >= 0.0 && un <= det)) return;
    const double Cx = Ty * (double)e1.z - Tz * (double)e1.y;
    const double Cy = Tz * (double)e1.x - Tx * (double)e1.z;
    const double Cz = Tx * (double)e1.y - Ty * (double)e1.x;
    const double vn = s * (Dx * Cx + Dy * Cy + Dz * Cz);
    if (!(vn >= 0.0 && un + vn <= det)) return;
    const double tn = s * ((double)e2.x * Cx + (double)e2.y * Cy + (double)e2.z * Cz);
    if (!(tn > 0.0 && tn < det)) return;
    const double tt = tn / det;
    h.px = (float)((double)ox + tt * Dx);
    h.py = (float)((double)oy + tt * Dy);
    h.pz = (float)((double)oz + tt * Dz);
    h.d = (float)tt;
    h.idx = t;
    h.type = 1;
    if (un == 0.0 || vn == 0.0 || un + vn == det) {
        h.type = 2;
        if ((un == 0.0 && vn == 0.0) || (un == det && vn == 0.0) || (un == 0.0 && vn == det))
            h.type = 3;
    }
}

// Moller-Trumbore in FP32 on the open segment 0<t<1 for triangle t with first
// vertex v0 and edges e1, e2 (e1.w: the grazing scale), whose plane the
// segment may cross; d2 = |p - o|^2 (seg_len2).  Returns true, with no hit,
// for a near-grazing segment, which mt_fp64 decides.
__device__ __forceinline__ bool mt_fp32(const float4 v0, const float4 e1, const float4 e2,
                                        uint32_t t, float ox, float oy, float oz, float px,
                                        float py, float pz, float d2, Hit &h) {
    h.type = 0;
    float dx = px - ox, dy = py - oy, dz = pz - oz;
    float qx_ = fmaf(dy, e2.z, -(dz * e2.y));
    float qy_ = fmaf(dz, e2.x, -(dx * e2.z));
    float qz_ = fmaf(dx, e2.y, -(dy * e2.x));
    float det = fmaf(e1.x, qx_, fmaf(e1.y, qy_, e1.z * qz_));
    const bool graze = det * det < e1.w * d2;
    if (graze || det == 0.0f) return graze;
    float inv = 1.0f / det;
    float tx = ox - v0.x, ty = oy - v0.y, tz = oz - v0.z;
    float u = fmaf(tx, qx_, fmaf(ty, qy_, tz * qz_)) * inv;
    if (!(u >= 0.0f && u <= 1.0f)) return false;
    float cx = fmaf(ty, e1.z, -(tz * e1.y));
    float cy = fmaf(tz, e1.x, -(tx * e1.z));
    float cz = fmaf(tx, e1.y, -(ty * e1.x));
    float v = fmaf(dx, cx, fmaf(dy, cy, dz * cz)) * inv;
    if (!(v >= 0.0f && u + v <= 1.0f)) return false;
    float tt = fmaf(e2.x, cx, fmaf(e2.y, cy, e2.z * cz)) * inv;
    if (!(tt > 0.0f && tt < 1.0f)) return false;
    h.px = fmaf(tt, dx, ox);
    h.py = fmaf(tt, dy, oy);
    h.pz = fmaf(tt, dz, oz);
    h.d = tt;
    h.idx = t;
    h.type = 1;
    if (u == 0.0f || v == 0.0f || u + v == 1.0f) {
        h.type = 2;
        if ((u == 0.0f && v == 0.0f) || (u == 1.0f && v == 0.0f) || (u == 0.0f && v == 1.0f))
            h.type = 3;
    }
    return false;
}

// Moller-Trumbore with the FP64 re-test of near-grazing segments.
__device__ __forceinline__ void mt_core(const float4 v0, const float4 e1, const float4 e2,
                                        uint32_t t, float ox, float oy, float oz, float px,
                                        float py, float pz, Hit &h) {
    if (mt_fp32(v0, e1, e2, t, ox, oy, oz, px, py, pz, seg_len2(ox, oy, oz, px, py, pz), h))
        mt_fp64(v0, e1, e2, t, ox, oy, oz, px, py, pz, h);
}

// The same test on the triangle's record in the MT table.
__device__ __forceinline__ void test_mt_crossing(const float4 *__restrict__ M, uint32_t t,
                                                 float ox, float oy, float oz, float px, float py,
                                                 float pz, Hit &h) {
    mt_core(M[t * 4 + 1], M[t * 4 + 2], M[t * 4 + 3], t, ox, oy, oz, px, py, pz, h);
}

// Plane-side prefilter, then Moller-Trumbore on the open segment 0<t<1.
__device__ __forceinline__ void test_mt(const float4 *__restrict__ M, uint32_t t, float ox,
                                        float oy, float oz, float px, float py, float pz, Hit &h) {
    h.type = 0;
    if (!crosses_plane(M[t * 4 + 0], ox, oy, oz, px, py, pz)) return;
    test_mt_crossing(M, t, ox, oy, oz, px, py, pz, h);
}

template <int MODE>
__device__ __forceinline__ bool better(const Hit &c, const Hit &b) {
    if (c.type == 0) return false;
    if (MODE == 0) return c.d < b.d;
    if (b.type == 0) return true;
    return (c.d < b.d) || (c.d == b.d && c.idx < b.idx);
}

// Where the walk reads a grid entry's records: kRecsCellMajor from the
// cell-major copies (AABB and reference grids), kRecsIndirect from the
// triangle's own record through cell_tri (exact-membership grids keep no
// copy), kRecsEither deciding per launch parameter (a uniform branch; the
// batched kernel, whose jobs may use either kind of grid).
// Moller-Trumbore vertices of grid entry k and its triangle index.
template <int RECS>
__device__ __forceinline__ uint32_t entry_mt(const StepParams &P, uint32_t k, float4 &v0,
                                             float4 &e1, float4 &e2) {
    if (RECS == kRecsCellMajor || (RECS == kRecsEither && P.cell_mt)) {
        v0 = P.cell_mt[(size_t)k * 3 + 0];
        e1 = P.cell_mt[(size_t)k * 3 + 1];
        e2 = P.cell_mt[(size_t)k * 3 + 2];
        return __float_as_uint(v0.w);
    }
    const uint32_t t = P.cell_tri[k];
    v0 = P.tri_mt[(size_t)t * 4 + 1];
    e1 = P.tri_mt[(size_t)t * 4 + 2];
    e2 = P.tri_mt[(size_t)t * 4 + 3];
    return t;
}
// Plane (n, n.v0) of grid entry k's triangle.
template <int RECS>
__device__ __forceinline__ float4 entry_plane(const StepParams &P, uint32_t k) {
    if (RECS == kRecsCellMajor || (RECS == kRecsEither && P.cell_plane)) return P.cell_plane[k];
    return P.tri_mt[(size_t)P.cell_tri[k] * 4];
}

// cubeCollDetect, spinKernelCpu.cpp:311-337, over one CSR cell.
template <int MODE>
__device__ __forceinline__ void cell_test(const StepParams &P, uint32_t cell, uint32_t excluded,
                                          float ox, float oy, float oz, float px, float py,
                                          float pz, Hit &best) {
    uint32_t k0 = P.cell_start[cell], k1 = P.cell_start[cell + 1];
    for (uint32_t k = k0; k < k1; ++k) {
        if (MODE == 0) {
            uint32_t t = P.cell_tri[k];
            if (t == excluded) continue;
            Hit c;
            test_ref(P.tri_ref, t, ox, oy, oz, px, py, pz, c);
            if (better<MODE>(c, best)) best = c;
        } else {
            // the cell-major plane record rejects most candidates without
            // touching the triangle record
            if (!crosses_plane(entry_plane<kRecsEither>(P, k), ox, oy, oz, px, py, pz)) continue;
            uint32_t t = P.cell_tri[k];
            if (t == excluded) continue;
            Hit c;
            test_mt_crossing(P.tri_mt, t, ox, oy, oz, px, py, pz, c);
            if (better<MODE>(c, best)) best = c;
        }
    }
}

// reflectPos, spinKernelCpu.cpp:406-422.
template <int MODE>
__device__ __forceinline__ void reflect(const StepParams &P, float sx, float sy, float sz,
                                        float tx, float ty, float tz, const Hit &h, float &rx,
                                        float &ry, float &rz) {
    if ((P.reflection_type == 0) | (h.type > 1)) {
        rx = sx; ry = sy; rz = sz;
        return;
    }
    if (MODE == 0) {
        const float4 a = P.tri_ref[h.idx * 4];
        float nx = a.x, ny = a.y, nz = a.z;
        float qx = tx - h.px, qy = ty - h.py, qz = tz - h.pz;
        rx = (1 - 2 * nx * nx) * qx - 2 * nx * ny * qy - 2 * nx * nz * qz + h.px;
        ry = -2 * nx * ny * qx + (1 - 2 * ny * ny) * qy - 2 * ny * nz * qz + h.py;
        rz = -2 * nx * nz * qx - 2 * ny * nz * qy + (1 - 2 * nz * nz) * qz + h.pz;
    } else {
        const float4 pl = P.tri_mt[h.idx * 4];
        float qx = tx - h.px, qy = ty - h.py, qz = tz - h.pz;
        float dn = fmaf(pl.x, qx, fmaf(pl.y, qy, pl.z * qz));
        float k = -2.0f * dn;
        rx = fmaf(k, pl.x, qx) + h.px;
        ry = fmaf(k, pl.y, qy) + h.py;
        rz = fmaf(k, pl.z, qz) + h.pz;
    }
}

// Compartment change on transmission: reference table spinKernelCpu.cpp:251-270
// (MODE 0); physical table (MODE 1) moves across each wall to the compartment
// on its other side.
template <int MODE>
__device__ __forceinline__ void transit(const StepParams &P, uint32_t info, uint32_t &comp,
                                        uint32_t &fiber) {
    uint32_t f = info & 0xFFFFu, m = info >> 16;
    if (MODE == 0) {
        if (comp == 2) {
            if (m == 0) comp = 1; else { comp = 0; fiber = kFiberNone; }
        } else if (comp == 1) {
            comp = 2;
        } else if (comp == 3) {
            comp = 0; fiber = kFiberNone;
        } else {
            fiber = f;
            comp = (m == 1) ? 2 : 3;
        }
    } else {
        uint32_t myel = (f < P.n_fibers) ? P.fiber_myelin[f] : 0u;
        if (comp == 0) {
            fiber = f;
            comp = (m == 0) ? 1 : (m == 1) ? 2 : 3;
        } else if (comp == 1) {
            if (myel) comp = 2; else { comp = 0; fiber = kFiberNone; }
        } else if (comp == 2) {
            if (m == 0) comp = 1; else { comp = 0; fiber = kFiberNone; }
        } else {
            comp = 0; fiber = kFiberNone;
        }
    }
}

__device__ __forceinline__ uint32_t tri_info_of(const StepParams &P, uint32_t t, int mode) {
    return mode == 0 ? __float_as_uint(P.tri_ref[t * 4 + 3].w)
                     : __float_as_uint(P.tri_mt[t * 4 + 1].w);
}

// collDetectRectGrid, spinKernelCpu.cpp:198-301.  MODE 0 stops at the first
// cell (x-major walk) holding a hit; MODE 1 visits the whole cell range and
// keeps the nearest hit (t, then triangle index).
template <int MODE>
__device__ __forceinline__ void collide(const StepParams &P, float ox, float oy, float oz,
                                        float &px, float &py, float &pz, float u, uint32_t &comp,
                                        uint32_t &fiber, uint32_t &first, uint32_t &last,
                                        uint32_t &ncoll) {
    float ex = px, ey = py, ez = pz;
    uint32_t excluded = kNone;
    float umax = 1.0f, umin = 0.0f;
    for (int guard = 0; guard < 4096; ++guard) {
        uint32_t sx, sy, sz, tx, ty, tz;
        cell_of<MODE>(P, ox, oy, oz, sx, sy, sz);
        cell_of<MODE>(P, px, py, pz, tx, ty, tz);
        uint32_t ix = (uint32_t)((tx > sx) - (tx < sx));
        uint32_t iy = (uint32_t)((ty > sy) - (ty < sy));
        uint32_t iz = (uint32_t)((tz > sz) - (tz < sz));
        Hit best;
        best.type = 0;
        best.idx = kNone;
        best.d = (MODE == 0) ? 400000000.0f : 2.0f;
        best.px = best.py = best.pz = 0.0f;
        uint32_t cx = sx;
        do {
            uint32_t cy = sy;
            do {
                uint32_t cz = sz;
                do {
                    uint32_t cell = cell_hash(P, cx, cy, cz);
                    if (MODE == 0) {
                        Hit c;
                        c.type = 0;
                        c.idx = kNone;
                        c.d = 400000000.0f;
                        cell_test<0>(P, cell, excluded, ox, oy, oz, px, py, pz, c);
                        best = c;
                    } else {
                        cell_test<1>(P, cell, excluded, ox, oy, oz, px, py, pz, best);
                    }
                    cz += iz;
                } while (cz != tz + iz && (MODE == 1 || best.type == 0));
                cy += iy;
            } while (cy != ty + iy && (MODE == 1 || best.type == 0));
            cx += ix;
        } while (cx != tx + ix && (MODE == 1 || best.type == 0));
        if (best.type == 0) break;
        if (ncoll == 0) first = best.idx;
        last = best.idx;
        ncoll += 1;
        float thr = umax - (umax - umin) * P.perm;
        if (u <= thr) {
            reflect<MODE>(P, ox, oy, oz, px, py, pz, best, ex, ey, ez);
            umax = thr;
        } else {
            umin = thr;
            transit<MODE>(P, tri_info_of(P, best.idx, MODE), comp, fiber);
        }
        ox = best.px; oy = best.py; oz = best.pz;
        px = ex; py = ey; pz = ez;
        excluded = best.idx;
        if (guard == 4095) atomicOr(P.flags, 1u);
    }
    px = ex; py = ey; pz = ez;
}

// Clearance shortcut.  The clearance of a cell bounds from below the distance
// from any point of the cell to any triangle registered in the grid, so a
// spin inside the cell can walk a path shorter than that without a search
// finding a hit.  Two allowances keep the float arithmetic on the safe side:
// kClearMargin for a position that rounds into the cell from just outside it,
// and kStepSlack per step for the rounding of the new position (half an ulp
// per coordinate at |x| < 2, 1.1e-7 as a vector) and of the step length.
constexpr float kClearMargin = 1e-5f;
constexpr float kStepSlack = 4e-7f;

// Clearance budget of a spin at (x, y, z): the cell's clearance less the
// margin, or -1 (no shortcut) outside the grid, where lookups clamp to an
// edge cell that does not contain the position.
__device__ __forceinline__ float clearance_at(const StepParams &P, float x, float y, float z) {
    const float qx = floorf((x - P.lo[0]) * P.inv_cell[0]);
    const float qy = floorf((y - P.lo[1]) * P.inv_cell[1]);
    const float qz = floorf((z - P.lo[2]) * P.inv_cell[2]);
    if (!(qx >= 0.0f && qx < (float)P.dims[0] && qy >= 0.0f && qy < (float)P.dims[1] &&
          qz >= 0.0f && qz < (float)P.dims[2]))
        return -1.0f;
    return P.cell_clear[cell_hash(P, (uint32_t)qx, (uint32_t)qy, (uint32_t)qz)] - kClearMargin;
}

// ------------------------------------------------------------------------ //
// Reference-mode step kernel: cpuIntegrate (spinKernelCpu.cpp:430-499) for
// `iters` steps per launch with the spin's state held in registers.
// ------------------------------------------------------------------------ //
template <int MODE, bool PERIODIC, int RNG, bool TRACE, bool MOMENT>
__global__ __launch_bounds__(64) void k_step(StepParams P, uint32_t iters) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    // the spin's identity keys its random stream, injected normals and trace
    const uint32_t sid = P.spin_id ? P.spin_id[i] : i;
    float x = P.x[i], y = P.y[i], z = P.z[i];
    float mag = P.mag[i];
    float ph = MOMENT ? 0.0f : P.phase[i];
    uint32_t lab = P.label[i];
    uint32_t comp = lab & 0xFFu, fiber = (lab >> 8) & 0xFFFFu;
    uint32_t s0 = 0, s1 = 0;
    if (RNG == kRngMwc) { s0 = P.seed[2 * i]; s1 = P.seed[2 * i + 1]; }
    float dx = 0.f, dy = 0.f, dz = 0.f, mx = 0.f, my = 0.f, mz = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
    if (MOMENT) {
        dx = P.disp[i]; dy = P.disp[P.n + i]; dz = P.disp[2 * P.n + i];
        mx = P.mom[i]; my = P.mom[P.n + i]; mz = P.mom[2 * P.n + i];
        cx = P.momc[i]; cy = P.momc[P.n + i]; cz = P.momc[2 * P.n + i];
    }
    for (uint32_t it = 0; it < iters; ++it) {
        const uint64_t gstep = P.step0 + it;
        float n0, n1, n2, u = 0.0f;
        if (RNG == kRngInject) {
            float4 q = P.inject[(size_t)it * P.n + sid];
            n0 = q.x; n1 = q.y; n2 = q.z; u = q.w;
        } else if (RNG == kRngMwc) {
            mwc_randn(s0, s1, n0, n1, n2, u);
        } else {
            philox_draw(P.key0, P.key1, sid, gstep, n0, n1, n2, u);
        }
        const float sd = comp == 0 ? P.std_dev[0] : comp == 1 ? P.std_dev[1]
                       : comp == 2 ? P.std_dev[2] : P.std_dev[3];
        const float ox = x, oy = y, oz = z;
        float px, py, pz;
        if (MODE == 0) {
            px = x + n0 * sd; py = y + n1 * sd; pz = z + n2 * sd;
        } else {
            px = fmaf(n0, sd, x); py = fmaf(n1, sd, y); pz = fmaf(n2, sd, z);
        }
        uint32_t first = kNone, last = kNone, ncoll = 0;
        collide<MODE>(P, ox, oy, oz, px, py, pz, u, comp, fiber, first, last, ncoll);
        if (!PERIODIC) {
            // spinKernelCpu.cpp:469-474
            if (px > P.clamp_hi) px = P.clamp_hi; else if (px < P.clamp_lo) px = P.clamp_lo;
            if (py > P.clamp_hi) py = P.clamp_hi; else if (py < P.clamp_lo) py = P.clamp_lo;
            if (pz > P.clamp_hi) pz = P.clamp_hi; else if (pz < P.clamp_lo) pz = P.clamp_lo;
        }
        if (MOMENT) {
            dx += px - ox; dy += py - oy; dz += pz - oz;
        }
        if (PERIODIC) {
            if (px >= 1.0f) px -= 2.0f; else if (px < -1.0f) px += 2.0f;
            if (py >= 1.0f) py -= 2.0f; else if (py < -1.0f) py += 2.0f;
            if (pz >= 1.0f) pz -= 2.0f; else if (pz < -1.0f) pz += 2.0f;
        }
        x = px; y = py; z = pz;
        const float t2 = comp == 0 ? P.t2[0] : comp == 1 ? P.t2[1] : comp == 2 ? P.t2[2] : P.t2[3];
        const float dc = comp == 0 ? P.decay[0] : comp == 1 ? P.decay[1]
                       : comp == 2 ? P.decay[2] : P.decay[3];
        if (MODE == 0) mag += -mag / t2 * P.dt;          // spinKernelCpu.cpp:486
        else mag = fmaf(-mag, dc, mag);
        if (!MOMENT) ph += (P.gx * x + P.gy * y + P.gz * z) * P.pc;  // spinKernelCpu.cpp:490
        if (MOMENT && P.weight != 0.0f) {
            float yv, tv;
            yv = fmaf(P.weight, dx, -cx); tv = mx + yv; cx = (tv - mx) - yv; mx = tv;
            yv = fmaf(P.weight, dy, -cy); tv = my + yv; cy = (tv - my) - yv; my = tv;
            yv = fmaf(P.weight, dz, -cz); tv = mz + yv; cz = (tv - mz) - yv; mz = tv;
        }
        if (TRACE) {
            uint32_t qx, qy, qz;
            cell_of<MODE>(P, x, y, z, qx, qy, qz);
            uint32_t *tr = P.trace + ((size_t)it * P.n + sid) * 5;
            tr[0] = first; tr[1] = last; tr[2] = ncoll;
            tr[3] = cell_hash(P, qx, qy, qz);
            tr[4] = comp | (fiber << 8);
        }
    }
    P.x[i] = x; P.y[i] = y; P.z[i] = z;
    P.mag[i] = mag;
    if (!MOMENT) P.phase[i] = ph;
    P.label[i] = comp | (fiber << 8);
    if (RNG == kRngMwc) { P.seed[2 * i] = s0; P.seed[2 * i + 1] = s1; }
    if (MOMENT) {
        P.disp[i] = dx; P.disp[P.n + i] = dy; P.disp[2 * P.n + i] = dz;
        P.mom[i] = mx; P.mom[P.n + i] = my; P.mom[2 * P.n + i] = mz;
        P.momc[i] = cx; P.momc[P.n + i] = cy; P.momc[2 * P.n + i] = cz;
    }
}

// ------------------------------------------------------------------------ //
// Physical-mode walk, laid out for 64-wide wavefronts.
//
// Each lane runs a small state machine: it either starts its next step (draw
// normals, propose the end point, take the clearance shortcut if it applies)
// or has a segment of the current step pending a collision search.  Every
// trip round the wave loop does one search round for all pending lanes
// together, so a lane whose step ended without a hit starts its next step
// while another lane is still reflecting.
//
// The search is compacted over the wave in two passes.  Each pending lane
// lists the CSR ranges of the cells its segment's bounding box covers, and a
// wave prefix sum numbers the (lane, grid entry) pairs.  The plane pass tests
// 64 pairs per trip, whatever lane they belong to, against the entry's plane
// record and queues the pairs whose plane the segment crosses; the triangle
// pass runs Moller-Trumbore on the queue, 64 pairs per trip.  The nearest hit
// per lane is an LDS atomic minimum over (t bits, triangle index): for t > 0
// the float bits order like the values, so this is the same (t, then index)
// order as a scan of the lane's cells in any order; the pair holding the
// minimum then writes its hit point for the lane.
// ------------------------------------------------------------------------ //
constexpr uint32_t kWaveSegs = 256;   // cell ranges listed per wave and search round
constexpr uint32_t kPass = 128;       // pairs per pass: two per lane, loads in flight together
// plane crossings awaiting the triangle test; the scratch of 16 waves (four
// per SIMD) then takes 155.6 KB of a CU's 160 KB LDS
constexpr uint32_t kWaveQueue = 384;

struct WaveScratch {
    uint2 seg[kWaveSegs];  // (first entry - lane-local pair number, lane-local end)
    float4 o[64];          // segment start; .w = excluded triangle
    float4 p[64];          // segment end
    float4 hit[64];        // nearest hit point; .w = hit type
    uint2 meta[64];        // (first pair number, first slot in seg)
    unsigned long long best[64];
    uint32_t map[kPass];   // lane + 1 at the pass position where its pairs start
    uint2 queue[kWaveQueue];  // (grid entry, lane)
};

// Orders this wave's LDS accesses: a wave's LDS instructions complete in
// issue order, so only the compiler has to be kept from moving them.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inclusive scan over the 64 lanes for an operation whose identity is 0
// (sum, unsigned maximum); every lane must be active.  Four DPP row shifts
// (1, 2, 4, 8) scan each row of 16 lanes; row_bcast:15 then carries row 0
// into row 1 and row 2 into row 3 (row mask 0b1010), and row_bcast:31
// carries lane 31 into rows 2 and 3 (row mask 0b1100).  A lane whose source
// is outside its row, or whose row is masked off, reads the identity, so each
// step is one VALU operation with a DPP operand.
template <bool MAX>
__device__ __forceinline__ uint32_t wave_scan(uint32_t v) {
#define DSIM_SCAN_STEP(ctrl, rows)                                                      \
    {                                                                                  \
        const uint32_t t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, ctrl, rows, 0xf, \
                                                                 false);               \
        v = MAX ? max(v, t) : v + t;                                                   \
    }
    DSIM_SCAN_STEP(0x111, 0xf)  // row_shr:1
    DSIM_SCAN_STEP(0x112, 0xf)  // row_shr:2
    DSIM_SCAN_STEP(0x114, 0xf)  // row_shr:4
    DSIM_SCAN_STEP(0x118, 0xf)  // row_shr:8
    DSIM_SCAN_STEP(0x142, 0xa)  // row_bcast:15
    DSIM_SCAN_STEP(0x143, 0xc)  // row_bcast:31
#undef DSIM_SCAN_STEP
    return v;
}
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v) { return wave_scan<false>(v); }
__device__ __forceinline__ uint32_t wave_incl_max(uint32_t v) { return wave_scan<true>(v); }
__device__ __forceinline__ uint32_t lane_rank(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ unsigned long long hit_key(float t, uint32_t tri) {
    return ((unsigned long long)__float_as_uint(t) << 32) | tri;
}

// Work counters of one wave in the counting build of the walk (indices
// kStat*), added to P.stats when the wave ends.  Every count is wave-uniform
// where it is taken.
struct WalkCounters {
    uint32_t v[kWalkStats];
};
__device__ __forceinline__ uint32_t popc(unsigned long long m) { return (uint32_t)__popcll(m); }

// FP32 triangle test of queue entry q (cell-major record) for its lane; the
// lane's minimum key in best[] is lowered to it on a hit.  Returns true for a
// near-grazing entry, which wave_graze_pass decides.
template <int RECS>
__device__ __forceinline__ bool queue_test(const StepParams &P, WaveScratch &L, uint32_t q,
                                           uint32_t qlen, Hit &h, uint32_t &owner,
                                           unsigned long long &key) {
    h.type = 0;
    owner = 0;
    key = ~0ull;
    if (q >= qlen) return false;
    const uint2 e = L.queue[q];
    owner = e.y;
    float4 v0, e1, e2;
    const uint32_t t = entry_mt<RECS>(P, e.x, v0, e1, e2);
    const float4 so = L.o[owner], sp = L.p[owner];
    if (t == __float_as_uint(so.w)) return false;
    if (mt_fp32(v0, e1, e2, t, so.x, so.y, so.z, sp.x, sp.y, sp.z, sp.w, h)) return true;
    if (h.type) {
        key = hit_key(h.d, t);
        atomicMin(&L.best[owner], key);
    }
    return false;
}

// FP64 re-test of the near-grazing queue entries in [0, qlen) (found again
// by the FP32 test), one entry per lane per trip, with the nearest-hit
// bookkeeping of the FP32 pass.  Wave-uniform qlen.
template <int RECS, bool STATS>
__device__ __forceinline__ void wave_graze_pass(const StepParams &P, WaveScratch &L,
                                                uint32_t lane, uint32_t qlen, WalkCounters &C) {
    for (uint32_t qb = 0; qb < qlen; qb += 64) {
        wave_lds_sync();
        Hit h;
        h.type = 0;
        uint32_t owner = 0;
        unsigned long long key = ~0ull;
        const uint32_t q = qb + lane;
        bool graze = false;
        if (q < qlen) {
            const uint2 e = L.queue[q];
            float4 v0, e1, e2;
            const uint32_t t = entry_mt<RECS>(P, e.x, v0, e1, e2);
            const float4 so = L.o[e.y], sp = L.p[e.y];
            graze = t != __float_as_uint(so.w) &&
                    mt_fp32(v0, e1, e2, t, so.x, so.y, so.z, sp.x, sp.y, sp.z, sp.w, h);
            if (graze) {
                mt_fp64(v0, e1, e2, t, so.x, so.y, so.z, sp.x, sp.y, sp.z, h);
                if (h.type) {
                    owner = e.y;
                    key = hit_key(h.d, t);
                    atomicMin(&L.best[owner], key);
                }
            }
        }
        if (STATS) C.v[kStatGraze] += popc(__ballot(graze));
        wave_lds_sync();
        if (h.type && L.best[owner] == key)
            L.hit[owner] = make_float4(h.px, h.py, h.pz, __uint_as_float(h.type));
    }
}

// Triangle pass over queue[0, qlen): nearest hit per lane into best[] and its
// point into hit[].  Wave-uniform qlen.
template <int RECS, bool STATS>
__device__ __forceinline__ void wave_triangle_pass(const StepParams &P, WaveScratch &L,
                                                   uint32_t lane, uint32_t qlen,
                                                   WalkCounters &C) {
    if (STATS) {
        C.v[kStatTriangle] += qlen;
        C.v[kStatTriPasses] += (qlen + kPass - 1) / kPass;
    }
    bool graze = false;
    for (uint32_t qb = 0; qb < qlen; qb += kPass) {
        wave_lds_sync();
        Hit ha, hb;
        uint32_t oa, ob;
        unsigned long long ka, kb;
        graze |= queue_test<RECS>(P, L, qb + lane, qlen, ha, oa, ka);
        graze |= queue_test<RECS>(P, L, qb + 64 + lane, qlen, hb, ob, kb);
        wave_lds_sync();
        // the pair holding the lane's minimum so far records its hit point; a
        // later pass that lowers the minimum overwrites it
        if (ha.type && L.best[oa] == ka)
            L.hit[oa] = make_float4(ha.px, ha.py, ha.pz, __uint_as_float(ha.type));
        if (hb.type && L.best[ob] == kb)
            L.hit[ob] = make_float4(hb.px, hb.py, hb.pz, __uint_as_float(hb.type));
    }
    if (__builtin_expect(__ballot(graze) != 0, 0)) wave_graze_pass<RECS, STATS>(P, L, lane, qlen, C);
    wave_lds_sync();
}

// Plane test of pair j for its owner lane: sets k (grid entry) and whether
// the owner's segment crosses the entry's plane.
template <int RECS>
__device__ __forceinline__ bool pair_crosses(const StepParams &P, const WaveScratch &L,
                                             uint32_t j, uint32_t total, uint32_t owner,
                                             uint32_t &k) {
    k = 0;
    if (j >= total) return false;
    const uint2 m = L.meta[owner];
    const uint32_t r = j - m.x;
    uint32_t s = m.y;
    uint2 e = L.seg[s];
    while (r >= e.y && s + 1 < kWaveSegs) e = L.seg[++s];
    k = e.x + r;
    if (k >= P.n_entries) {  // cannot happen when the numbering is right
        atomicOr(P.flags, 4u);
        k = 0;
    }
    const float4 so = L.o[owner], sp = L.p[owner];
    return crosses_plane(entry_plane<RECS>(P, k), so.x, so.y, so.z, sp.x, sp.y, sp.z);
}

// One search round, called by all 64 lanes of the wave together.  For each
// lane with `search` set, either finds the nearest hit of the segment o-p
// (excluding `excluded`): returns true with key = hit_key(t, triangle) (or ~0
// when there is none) and the hit point and type in `hit`; or returns false
// when the wave's cell list is full in this search round, or when the lane's
// cell box is longer than the list and cells remain (the lane searches again).
// A box of more than kWaveSegs cells is searched kWaveSegs cells per round, in
// x-fastest order from cell `slice` (0 for a new segment; advanced here), its
// nearest hit carried in best[]/hit[] from round to round.
template <int RECS, bool STATS>
__device__ __forceinline__ bool wave_search(const StepParams &P, WaveScratch &L, uint32_t lane,
                                            bool search, float ox, float oy, float oz, float px,
                                            float py, float pz, uint32_t excluded,
                                            uint32_t &slice, unsigned long long &key,
                                            float4 &hit, WalkCounters &C) {
    uint32_t x0 = 0, x1 = 0, y0 = 0, y1 = 0, z0 = 0, z1 = 0, ncell = 0;
    if (search) {
        uint32_t sx, sy, sz, tx, ty, tz;
        cell_of<1>(P, ox, oy, oz, sx, sy, sz);
        cell_of<1>(P, px, py, pz, tx, ty, tz);
        x0 = min(sx, tx); x1 = max(sx, tx);
        y0 = min(sy, ty); y1 = max(sy, ty);
        z0 = min(sz, tz); z1 = max(sz, tz);
        ncell = (x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1);
    }
    // a lane lists at most kWaveSegs cells of its box per round; one that does
    // not fit behind the lanes before it waits for the next round (the first
    // lane always fits)
    const uint32_t nreq = search ? min(ncell - slice, kWaveSegs) : 0u;
    const uint32_t send = wave_incl_sum(nreq);
    const uint32_t soff = send - nreq;
    const bool coop = nreq > 0 && send <= kWaveSegs;
    uint32_t cnt = 0;
    if (coop) {
        uint32_t cx = x0, cy = y0, cz = z0;
        if (slice) {  // a later slice of a long box
            const uint32_t nx = x1 - x0 + 1, ny = y1 - y0 + 1;
            cx += slice % nx;
            cy += (slice / nx) % ny;
            cz += slice / (nx * ny);
        }
        for (uint32_t s = 0; s < nreq; s += 4) {
            uint32_t k0[4], k1[4];
#pragma unroll
            for (uint32_t q = 0; q < 4; ++q) {
                k0[q] = k1[q] = 0;
                if (s + q < nreq) {
                    const uint32_t c = cell_hash(P, cx, cy, cz);
                    k0[q] = P.cell_start[c];
                    k1[q] = P.cell_start[c + 1];
                    if (++cx > x1) { cx = x0; if (++cy > y1) { cy = y0; ++cz; } }
                }
            }
#pragma unroll
            for (uint32_t q = 0; q < 4; ++q) {
                if (s + q < nreq) {
                    L.seg[soff + s + q] = make_uint2(k0[q] - cnt, cnt + (k1[q] - k0[q]));
                    cnt += k1[q] - k0[q];
                }
            }
        }
    }
    const uint32_t cend = wave_incl_sum(cnt);
    const uint32_t off = cend - cnt;
    const uint32_t total = __builtin_amdgcn_readlane(cend, 63);
    if (STATS) {
        C.v[kStatSearchLanes] += popc(__ballot(coop));
        C.v[kStatDeferred] += popc(__ballot(nreq > 0 && !coop));
        C.v[kStatCells] += __builtin_amdgcn_readlane(wave_incl_sum(coop ? nreq : 0u), 63);
        C.v[kStatPlane] += total;
        C.v[kStatPasses] += (total + kPass - 1) / kPass;
    }
    if (coop) {
        L.o[lane] = make_float4(ox, oy, oz, __uint_as_float(excluded));
        L.p[lane] = make_float4(px, py, pz, seg_len2(ox, oy, oz, px, py, pz));
        L.meta[lane] = make_uint2(off, soff);
        if (slice == 0) L.best[lane] = ~0ull;  // later slices keep the minimum so far
    }
    uint32_t qlen = 0;  // wave-uniform
    for (uint32_t base = 0; base < total; base += kPass) {
        L.map[lane] = 0;
        L.map[lane + 64] = 0;
        wave_lds_sync();
        if (cnt && off < base + kPass && off + cnt > base) L.map[off > base ? off - base : 0] = lane + 1;
        wave_lds_sync();
        // owner of pass position q: the last lane whose pairs start at or
        // before it (position 0 always holds the lane covering `base`)
        const uint32_t own_a = wave_incl_max(L.map[lane]);
        const uint32_t own_b = max(wave_incl_max(L.map[lane + 64]),
                                   __builtin_amdgcn_readlane(own_a, 63));
        uint32_t ka, kb;
        const bool ca = pair_crosses<RECS>(P, L, base + lane, total, own_a - 1, ka);
        const bool cb = pair_crosses<RECS>(P, L, base + 64 + lane, total, own_b - 1, kb);
        const unsigned long long ba = __ballot(ca), bb = __ballot(cb);
        if (ca) L.queue[qlen + lane_rank(ba)] = make_uint2(ka, own_a - 1);
        qlen += (uint32_t)__popcll(ba);
        if (cb) L.queue[qlen + lane_rank(bb)] = make_uint2(kb, own_b - 1);
        qlen += (uint32_t)__popcll(bb);
        if (qlen > kWaveQueue - kPass) {
            wave_triangle_pass<RECS, STATS>(P, L, lane, qlen, C);
            qlen = 0;
        }
    }
    if (qlen) wave_triangle_pass<RECS, STATS>(P, L, lane, qlen, C);
    wave_lds_sync();
    key = ~0ull;
    const bool done = coop && slice + nreq >= ncell;
    if (coop && !done) slice += nreq;
    if (done) {
        key = L.best[lane];
        if (key != ~0ull) hit = L.hit[lane];
    }
    return done;
}

// Workgroups are handed to the 8 XCDs in turn (workgroup b to XCD b mod 8),
// and each XCD has its own L2.  The spins are renumbered so that each XCD
// takes runs of `run` consecutive spin blocks: once the spins are sorted
// (rate band, then Morton order) the blocks an XCD holds at a time cover a
// compact stretch of that order.  Blocks past the last whole round of runs
// keep their number.
constexpr uint32_t kXcds = 8;
constexpr uint32_t kWalkXcdRun = 128;  // k_walk: 8192 spins a run
constexpr uint32_t kBatchXcdRun = 32;  // k_walk_batch: 2048 spins a run (jobs are small)
__device__ __forceinline__ uint32_t xcd_block(uint32_t run) {
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t full = nb / (kXcds * run) * (kXcds * run);
    if (b >= full) return b;
    const uint32_t x = b % kXcds, k = b / kXcds;  // XCD and its k-th workgroup
    return ((k / run) * kXcds + x) * run + k % run;
}

// Periodic voxel: one face crossing per step at most (a step is far shorter
// than the voxel); k counts the crossings, +1 for each wrap back from hi.
__device__ __forceinline__ void wrap_axis(float &p, float lo, float hi, float len, int &k) {
    if (p >= hi) { p -= len; k += 1; } else if (p < lo) { p += len; k -= 1; }
}
// Three signed 10-bit crossing counts in one word (the count wraps at +-512,
// far beyond any walk of a voxel wider than a few steps).
__device__ __forceinline__ void unpack_wrap(uint32_t w, int &kx, int &ky, int &kz) {
    kx = (int)(w << 22) >> 22;
    ky = (int)(w << 12) >> 22;
    kz = (int)(w << 2) >> 22;
}
__device__ __forceinline__ uint32_t pack_wrap(int kx, int ky, int kz) {
    return ((uint32_t)kx & 1023u) | (((uint32_t)ky & 1023u) << 10) | (((uint32_t)kz & 1023u) << 20);
}

// cpuIntegrate (spinKernelCpu.cpp:430-499) in physical mode: `iters` steps
// for the blockDim.x spins of block `block` of the simulation described by P.
// Boundary of the walk: clamp box, the periodic cube [-1, 1)^3 (constant
// bounds), or a periodic box with the bounds of the launch parameters.
enum { kBoundClamp = 0, kBoundCube = 1, kBoundBox = 2 };

// Spin slots of the persistent walk: queue q (one per XCD) hands out the
// slots of runs q, q + 8, q + 16, ... of `run` consecutive slots, so that
// every XCD walks a like mix of the sorted order and a compact stretch of it
// at a time.  Slot s = (8 (g / run) + q) run + g % run grows with the queue
// count g, so a queue past the end stays there.  Returns n (no slot) past
// the end.
__device__ __forceinline__ uint32_t queue_slot(uint32_t q, uint32_t g, uint32_t n, uint32_t run) {
    const uint64_t s = ((uint64_t)(g / run) * kXcds + q) * run + g % run;
    return s < n ? (uint32_t)s : n;
}
// Run length for n spins: as near kWalkQueueRun as gives every XCD the same
// number of runs (the last one short).  The queues do not share work, so
// with a fixed length the XCDs holding one run fewer than the others sit
// idle at the end of every launch: 1e6 spins in runs of 8192 are 122 runs,
// 16 for two XCDs and 15 for six.  On 1e6 spins at a 2048-step reorder
// interval: fixed runs of 8192 2.03e10 spin-steps/s, of 4096 2.10e10;
// balanced runs near 4096 2.10e10, near 2048 2.12e10, near 1024 2.09e10.
constexpr uint32_t kWalkQueueRun = 2048;
inline uint32_t queue_run(uint32_t n) {
    const uint64_t per_xcd =
        std::max<uint64_t>(1, ((uint64_t)n + kXcds * kWalkQueueRun - 1) / (kXcds * kWalkQueueRun));
    return (uint32_t)(((uint64_t)n + kXcds * per_xcd - 1) / (kXcds * per_xcd));
}
// Shortest launch (steps) that takes the persistent walk (see launch_one).
constexpr uint32_t kQueueMinSteps = 768;

// TRAJBUF (trajectory runs): each lane keeps its last four steps'
// displacements in registers and stores them as three 16-byte chunks every
// fourth step (launches and buffers are multiples of four steps).
template <int BOUND, int RNG, int INSTR, bool MOMENT, int RECS, bool QUEUE = false,
          bool TRAJBUF = false>
__device__ __forceinline__ void walk_body(const StepParams &P, uint32_t block, uint32_t iters,
                                          WaveScratch &L) {
    constexpr bool PERIODIC = BOUND != kBoundClamp;
    constexpr bool TRACE = INSTR == kInstrTrace, STATS = INSTR == kInstrStats;
    WalkCounters C;
    if (STATS)
        for (uint32_t k = 0; k < kWalkStats; ++k) C.v[k] = 0;
    uint32_t shortcuts = 0;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t sid = 0, comp = 0, fiber = kFiberNone, s0 = 0, s1 = 0;
    float x = 0.f, y = 0.f, z = 0.f, mag = 0.f, ph = 0.f;
    float dx = 0.f, dy = 0.f, dz = 0.f, mx = 0.f, my = 0.f, mz = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
    float safe = -1.0f;
    float tb[TRAJBUF ? 12 : 1];  // TRAJBUF: displacements of the last four steps
    // voxel crossings (periodic phase): the unwrapped position is x + kx * period
    int kx = 0, ky = 0, kz = 0;
    uint32_t nhit = 0, hit_total = 0;
    auto load = [&](uint32_t i) {
        sid = P.spin_id ? P.spin_id[i] : i;
        x = P.x[i]; y = P.y[i]; z = P.z[i];
        mag = P.mag[i];
        if (!MOMENT) ph = P.phase[i];
        if (PERIODIC && !MOMENT) unpack_wrap(P.wrap[i], kx, ky, kz);
        const uint32_t lab = P.label[i];
        comp = lab & 0xFFu; fiber = (lab >> 8) & 0xFFFFu;
        if (RNG == kRngMwc) { s0 = P.seed[2 * i]; s1 = P.seed[2 * i + 1]; }
        if (MOMENT) {
            dx = P.disp[i]; dy = P.disp[P.n + i]; dz = P.disp[2 * P.n + i];
            mx = P.mom[i]; my = P.mom[P.n + i]; mz = P.mom[2 * P.n + i];
            cx = P.momc[i]; cy = P.momc[P.n + i]; cz = P.momc[2 * P.n + i];
        }
        if (P.cell_clear) safe = clearance_at(P, x, y, z);
        nhit = 0;
    };
    auto store = [&](uint32_t i) {
        P.x[i] = x; P.y[i] = y; P.z[i] = z;
        P.mag[i] = mag;
        if (!MOMENT) P.phase[i] = ph;
        if (PERIODIC && !MOMENT) P.wrap[i] = pack_wrap(kx, ky, kz);
        P.label[i] = comp | (fiber << 8);
        if (RNG == kRngMwc) { P.seed[2 * i] = s0; P.seed[2 * i + 1] = s1; }
        if (MOMENT) {
            P.disp[i] = dx; P.disp[P.n + i] = dy; P.disp[2 * P.n + i] = dz;
            P.mom[i] = mx; P.mom[P.n + i] = my; P.mom[2 * P.n + i] = mz;
            P.momc[i] = cx; P.momc[P.n + i] = cy; P.momc[2 * P.n + i] = cz;
        }
        if (P.hits) P.hits[i] += nhit;
        if (STATS) hit_total += nhit;
    };
    // The persistent walk (QUEUE): lanes take slots from their XCD's queue,
    // the finishing lanes of a trip together (one atomic per wave).
    const uint32_t q = block % kXcds;
    auto take = [&](bool want) -> uint32_t {
        const unsigned long long m = __ballot(want);
        uint32_t base = 0;
        const uint32_t lead = (uint32_t)__builtin_ctzll(m);
        if (lane == lead) base = atomicAdd(P.queue + q, popc(m));
        base = __builtin_amdgcn_readlane(base, lead);
        return want ? queue_slot(q, base + lane_rank(m), P.n, P.queue_run) : P.n;
    };
    // lanes past the last spin stay in the loop (done) for the wave operations
    uint32_t i = QUEUE ? take(true) : block * blockDim.x + threadIdx.x;
    bool live = i < P.n;
    if (live) load(i);
    uint32_t it = live ? 0u : iters;
    // the pending segment of the current step and its permeation interval
    bool pending = false, refresh = false;
    float ox = 0.f, oy = 0.f, oz = 0.f, px = 0.f, py = 0.f, pz = 0.f;
    float u = 0.f, umax = 1.f, umin = 0.f;
    uint32_t excluded = kNone, first = kNone, last = kNone, ncoll = 0;
    uint32_t slice = 0;  // cells of the segment's box searched so far
    while (__ballot(it < iters) != 0ull) {
        bool complete = false;
        if (STATS) {
            C.v[kStatTrips] += 1;
            C.v[kStatLiveLaneTrips] += popc(__ballot(it < iters));
        }
        if (it < iters && !pending) {
            float n0, n1, n2;
            if (RNG == kRngInject) {
                const float4 q = P.inject[(size_t)it * P.n + sid];
                n0 = q.x; n1 = q.y; n2 = q.z; u = q.w;
            } else if (RNG == kRngMwc) {
                mwc_randn(s0, s1, n0, n1, n2, u);
            } else {
                philox_draw(P.key0, P.key1, sid, P.step0 + it, n0, n1, n2, u);
            }
            const float sd = comp == 0 ? P.std_dev[0] : comp == 1 ? P.std_dev[1]
                           : comp == 2 ? P.std_dev[2] : P.std_dev[3];
            ox = x; oy = y; oz = z;
            px = fmaf(n0, sd, x); py = fmaf(n1, sd, y); pz = fmaf(n2, sd, z);
            first = last = excluded = kNone;
            ncoll = 0;
            umax = 1.0f; umin = 0.0f;
            pending = true;
            refresh = true;
            slice = 0;
            if (P.cell_clear) {
                // no triangle lies within `safe` of the path walked since the
                // last check, so a shorter step cannot hit one
                const float len =
                    sd * __builtin_amdgcn_sqrtf(fmaf(n0, n0, fmaf(n1, n1, n2 * n2))) + kStepSlack;
                if (len < safe) {
                    safe -= len;
                    pending = false;
                    refresh = false;
                    complete = true;
                    if (STATS) ++shortcuts;
                }
            }
        }
        unsigned long long key;
        float4 hp;
        const bool searched =
            wave_search<RECS, STATS>(P, L, lane, pending, ox, oy, oz, px, py, pz, excluded, slice,
                                     key, hp, C);
        if (pending && searched) {
            if (key == ~0ull) {
                pending = false;
                complete = true;
            } else {
                const uint32_t idx = (uint32_t)key;
                Hit h;
                h.px = hp.x; h.py = hp.y; h.pz = hp.z;
                h.type = __float_as_uint(hp.w);
                h.idx = idx;
                if (ncoll == 0) first = idx;
                last = idx;
                ncoll += 1;
                nhit += 1;
                const float thr = umax - (umax - umin) * P.perm;
                float ex = px, ey = py, ez = pz;
                if (u <= thr) {
                    reflect<1>(P, ox, oy, oz, px, py, pz, h, ex, ey, ez);
                    umax = thr;
                } else {
                    umin = thr;
                    transit<1>(P, tri_info_of(P, idx, 1), comp, fiber);
                }
                ox = h.px; oy = h.py; oz = h.pz;
                px = ex; py = ey; pz = ez;
                excluded = idx;
                slice = 0;
                if (ncoll == 4096) {
                    atomicOr(P.flags, 1u);
                    pending = false;
                    complete = true;
                }
            }
        }
        if (complete) {
            if (!PERIODIC) {
                if (px > P.clamp_hi) px = P.clamp_hi; else if (px < P.clamp_lo) px = P.clamp_lo;
                if (py > P.clamp_hi) py = P.clamp_hi; else if (py < P.clamp_lo) py = P.clamp_lo;
                if (pz > P.clamp_hi) pz = P.clamp_hi; else if (pz < P.clamp_lo) pz = P.clamp_lo;
            }
            if (MOMENT) {
                dx += px - x; dy += py - y; dz += pz - z;
            }
            if (PERIODIC) {
                const float wx = px, wy = py, wz = pz;
                if (BOUND == kBoundCube) {
                    wrap_axis(px, -1.0f, 1.0f, 2.0f, kx);
                    wrap_axis(py, -1.0f, 1.0f, 2.0f, ky);
                    wrap_axis(pz, -1.0f, 1.0f, 2.0f, kz);
                } else {
                    wrap_axis(px, P.wrap_lo[0], P.wrap_hi[0], P.wrap_len[0], kx);
                    wrap_axis(py, P.wrap_lo[1], P.wrap_hi[1], P.wrap_len[1], ky);
                    wrap_axis(pz, P.wrap_lo[2], P.wrap_hi[2], P.wrap_len[2], kz);
                }
                refresh |= (px != wx) | (py != wy) | (pz != wz);
            }
            x = px; y = py; z = pz;
            if (P.cell_clear && refresh) safe = clearance_at(P, x, y, z);
            const float dc = comp == 0 ? P.decay[0] : comp == 1 ? P.decay[1]
                           : comp == 2 ? P.decay[2] : P.decay[3];
            mag = fmaf(-mag, dc, mag);
            // spinKernelCpu.cpp:490 with the post-step position; in a periodic
            // voxel that is the unwrapped position (x + k * period, exactly x
            // for a spin that has not crossed a face)
            if (!MOMENT && !PERIODIC) ph += (P.gx * x + P.gy * y + P.gz * z) * P.pc;
            if (!MOMENT && PERIODIC) {
                const float lx = BOUND == kBoundCube ? 2.0f : P.wrap_len[0];
                const float ly = BOUND == kBoundCube ? 2.0f : P.wrap_len[1];
                const float lz = BOUND == kBoundCube ? 2.0f : P.wrap_len[2];
                const float ux = x + (float)kx * lx;
                const float uy = y + (float)ky * ly;
                const float uz = z + (float)kz * lz;
                ph += (P.gx * ux + P.gy * uy + P.gz * uz) * P.pc;
            }
            if (MOMENT && P.weight != 0.0f) {
                float yv, tv;
                yv = fmaf(P.weight, dx, -cx); tv = mx + yv; cx = (tv - mx) - yv; mx = tv;
                yv = fmaf(P.weight, dy, -cy); tv = my + yv; cy = (tv - my) - yv; my = tv;
                yv = fmaf(P.weight, dz, -cz); tv = mz + yv; cz = (tv - mz) - yv; mz = tv;
            }
            if (TRAJBUF) {
                // unwrapped displacement after this step, k = 3 * it + axis,
                // shifted into the four-step window; stored when it is full
#pragma unroll
                for (int q = 0; q < 9; ++q) tb[q] = tb[q + 3];
                tb[9] = dx; tb[10] = dy; tb[11] = dz;
                if ((it & 3u) == 3u) {
                    float4 *tr = reinterpret_cast<float4 *>(
                        P.traj + (size_t)(sid / kTrajTile) * P.traj_stride +
                        (size_t)(3 * (it - 3) / kTrajChunk) * (kTrajTile * kTrajChunk)) +
                        sid % kTrajTile;
                    tr[0] = make_float4(tb[0], tb[1], tb[2], tb[3]);
                    tr[kTrajTile] = make_float4(tb[4], tb[5], tb[6], tb[7]);
                    tr[2 * kTrajTile] = make_float4(tb[8], tb[9], tb[10], tb[11]);
                }
            } else if (MOMENT && P.traj) {
                // unwrapped displacement after this step, k = 3 * it + axis
                // of the spin's tile
                float *tr = P.traj + (size_t)(sid / kTrajTile) * P.traj_stride +
                            (sid % kTrajTile) * kTrajChunk;
#pragma unroll
                for (uint32_t a = 0; a < 3; ++a) {
                    const uint32_t k = 3 * it + a;
                    tr[(size_t)(k / kTrajChunk) * (kTrajTile * kTrajChunk) + k % kTrajChunk] =
                        a == 0 ? dx : a == 1 ? dy : dz;
                }
            }
            if (TRACE) {
                uint32_t qx, qy, qz;
                cell_of<1>(P, x, y, z, qx, qy, qz);
                uint32_t *tr = P.trace + ((size_t)it * P.n + sid) * 5;
                tr[0] = first; tr[1] = last; tr[2] = ncoll;
                tr[3] = cell_hash(P, qx, qy, qz);
                tr[4] = comp | (fiber << 8);
            }
            ++it;
        }
        if (QUEUE) {
            // a lane whose spin has walked its steps stores it and takes the
            // next slot of its queue
            const bool fin = live && it == iters;
            if (__ballot(fin) != 0ull) {
                if (fin) store(i);
                if (STATS) C.v[kStatSteps] += popc(__ballot(fin)) * iters;
                const uint32_t ni = take(fin);
                if (fin) {
                    i = ni;
                    live = i < P.n;
                    if (live) {
                        load(i);
                        it = 0;
                    }
                }
            }
        }
    }
    if (!QUEUE && live) store(i);
    if (STATS) {
        if (!QUEUE) C.v[kStatSteps] = popc(__ballot(live)) * iters;
        C.v[kStatHits] = __builtin_amdgcn_readlane(wave_incl_sum(hit_total), 63);
        C.v[kStatShortcut] = __builtin_amdgcn_readlane(wave_incl_sum(shortcuts), 63);
        C.v[kStatWaves] = 1;
        if (lane == 0)
            for (uint32_t k = 0; k < kWalkStats; ++k)
                if (C.v[k]) atomicAdd(P.stats + k, (unsigned long long)C.v[k]);
    }
}

template <int BOUND, int RNG, int INSTR, bool MOMENT, int RECS>
// One wave per workgroup: every wave runs until its own slowest lane is done,
// and a finished wave frees its slot for the next block at once (with four
// waves per workgroup the slot waited for the slowest of the four).
__global__ __launch_bounds__(kWalkBlockSpins) __attribute__((amdgpu_waves_per_eu(4))) void k_walk(
    StepParams P, uint32_t iters) {
    __shared__ WaveScratch scratch[1];
    walk_body<BOUND, RNG, INSTR, MOMENT, RECS>(P, xcd_block(kWalkXcdRun), iters, scratch[0]);
}

// The walk of trajectory runs (moment mode, Philox): k_walk with the
// trajectory stores gathered four steps at a time (walk_body's TRAJBUF).
template <int BOUND, int RECS>
__global__ __launch_bounds__(kWalkBlockSpins) __attribute__((amdgpu_waves_per_eu(4))) void k_walk_traj(
    StepParams P, uint32_t iters) {
    __shared__ WaveScratch scratch[1];
    walk_body<BOUND, kRngPhilox, kInstrNone, true, RECS, false, true>(P, xcd_block(kWalkXcdRun),
                                                                      iters, scratch[0]);
}

// Persistent walk: one one-wave workgroup per resident slot (block b on XCD
// b mod 8); each lane walks a spin's steps, stores it and takes the next slot
// from its XCD's queue, so no lane idles while its wave waits for the slowest
// spin of a fixed block.  Spins are independent (Philox counters keyed by the
// spin id), so the results equal k_walk's.
template <int BOUND, int INSTR, bool MOMENT, int RECS>
__global__ __launch_bounds__(kWalkBlockSpins) __attribute__((amdgpu_waves_per_eu(4))) void k_walk_queue(
    StepParams P, uint32_t iters) {
    __shared__ WaveScratch scratch[1];
    walk_body<BOUND, kRngPhilox, INSTR, MOMENT, RECS, true>(P, blockIdx.x, iters, scratch[0]);
}

// Batched walk for the server: the blocks of one launch are shared out over
// the jobs of a table in device memory (job j owns blocks [block_start[j],
// block_start[j+1])).  Each job is a whole simulation (its own substrate,
// spins and parameters) advanced by its own number of steps; its spins see
// exactly the arithmetic of a k_walk launch of their own.  Blocks past the
// table's total exit at once, so one launch shape (and one captured graph)
// serves any mix of jobs.
template <int BOUND, int RECS>
__global__ __launch_bounds__(kWalkBlockSpins) __attribute__((amdgpu_waves_per_eu(4))) void k_walk_batch(const WalkTable *__restrict__ T) {
    __shared__ WaveScratch scratch[1];
    const uint32_t b = xcd_block(kBatchXcdRun);
    if (b >= T->total_blocks) return;
    uint32_t j = 0;
    for (uint32_t st = kMaxJobs / 2; st; st >>= 1)
        if (j + st < T->n_jobs && T->block_start[j + st] <= b) j += st;
    const WalkJob &J = T->job[j];
    walk_body<BOUND, kRngPhilox, kInstrNone, true, RECS>(J.p, b - T->block_start[j], J.iters,
                                                    scratch[0]);
}

hipError_t launch_walk_batch(const WalkTable *table, uint32_t grid_blocks, bool periodic,
                             int recs, hipStream_t s) {
    if (grid_blocks == 0) return hipSuccess;
    const dim3 g(grid_blocks), b(kWalkBlockSpins);
    // every job's voxel comes from its parameters (kBoundBox): jobs differ
    if (periodic) {
        if (recs == kRecsCellMajor) hipLaunchKernelGGL((k_walk_batch<kBoundBox, kRecsCellMajor>), g, b, 0, s, table);
        else if (recs == kRecsIndirect) hipLaunchKernelGGL((k_walk_batch<kBoundBox, kRecsIndirect>), g, b, 0, s, table);
        else hipLaunchKernelGGL((k_walk_batch<kBoundBox, kRecsEither>), g, b, 0, s, table);
    } else {
        if (recs == kRecsCellMajor) hipLaunchKernelGGL((k_walk_batch<kBoundClamp, kRecsCellMajor>), g, b, 0, s, table);
        else if (recs == kRecsIndirect) hipLaunchKernelGGL((k_walk_batch<kBoundClamp, kRecsIndirect>), g, b, 0, s, table);
        else hipLaunchKernelGGL((k_walk_batch<kBoundClamp, kRecsEither>), g, b, 0, s, table);
    }
    return hipGetLastError();
}

// Resident one-wave workgroups of the walk on this device (four per SIMD).
static uint32_t resident_walk_blocks() {
    static uint32_t cached = 0;
    if (!cached) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
        cached = (uint32_t)cus * 16u;
    }
    return cached;
}

template <int MODE, int BOUND, int RNG, int INSTR, bool MOMENT>
static void launch_one(const StepParams &p, uint32_t iters, hipStream_t s) {
    if constexpr (MODE == 1 && RNG == kRngPhilox && INSTR != kInstrTrace) {
        // the persistent walk once every resident slot has two spins or more
        // and a launch walks at least kQueueMinSteps steps; not while a
        // trajectory is recorded (its lanes, at different steps of different
        // spins, would scatter every trajectory store over 64 rows).  Lanes
        // take a queue slot per spin walked, so short launches turn the eight
        // queue counters into a stream of atomics: 4e6 spins in 128-step
        // launches walked 2.3x slower than fixed blocks, at 512 steps 3 %
        // slower, at 768 3 % faster.
        const uint32_t slots = resident_walk_blocks();
        if (p.queue && !p.traj && p.n >= 2 * slots * kWalkBlockSpins && iters >= kQueueMinSteps) {
            if (hipMemsetAsync(p.queue, 0, kXcds * sizeof(uint32_t), s) != hipSuccess) return;
            StepParams pq = p;
            pq.queue_run = queue_run(p.n);
            if (p.cell_mt)
                hipLaunchKernelGGL((k_walk_queue<BOUND, INSTR, MOMENT, kRecsCellMajor>), dim3(slots),
                                   dim3(kWalkBlockSpins), 0, s, pq, iters);
            else
                hipLaunchKernelGGL((k_walk_queue<BOUND, INSTR, MOMENT, kRecsIndirect>), dim3(slots),
                                   dim3(kWalkBlockSpins), 0, s, pq, iters);
            return;
        }
    }
    if constexpr (MODE == 1 && MOMENT && RNG == kRngPhilox && INSTR == kInstrNone) {
        // four-step windows need whole windows per launch
        if (p.traj && iters % 4 == 0) {
            if (p.cell_mt)
                hipLaunchKernelGGL((k_walk_traj<BOUND, kRecsCellMajor>),
                                   dim3((p.n + kWalkBlockSpins - 1) / kWalkBlockSpins),
                                   dim3(kWalkBlockSpins), 0, s, p, iters);
            else
                hipLaunchKernelGGL((k_walk_traj<BOUND, kRecsIndirect>),
                                   dim3((p.n + kWalkBlockSpins - 1) / kWalkBlockSpins),
                                   dim3(kWalkBlockSpins), 0, s, p, iters);
            return;
        }
    }
    if constexpr (MODE == 0) {
        hipLaunchKernelGGL((k_step<0, BOUND != kBoundClamp, RNG, INSTR == kInstrTrace, MOMENT>),
                           dim3((p.n + 63) / 64), dim3(64), 0, s, p, iters);
    } else if (p.cell_mt) {
        hipLaunchKernelGGL((k_walk<BOUND, RNG, INSTR, MOMENT, kRecsCellMajor>),
                           dim3((p.n + kWalkBlockSpins - 1) / kWalkBlockSpins),
                           dim3(kWalkBlockSpins), 0, s, p, iters);
    } else {
        hipLaunchKernelGGL((k_walk<BOUND, RNG, INSTR, MOMENT, kRecsIndirect>),
                           dim3((p.n + kWalkBlockSpins - 1) / kWalkBlockSpins),
                           dim3(kWalkBlockSpins), 0, s, p, iters);
    }
}

template <int MODE, int BOUND, int RNG, int INSTR>
static void launch_m(const StepParams &p, uint32_t iters, bool moment, hipStream_t s) {
    if (moment) launch_one<MODE, BOUND, RNG, INSTR, true>(p, iters, s);
    else launch_one<MODE, BOUND, RNG, INSTR, false>(p, iters, s);
}

// The counting build of the walk exists for the physical mode with Philox
// or injected draws (the runs whose work it measures, and the injected runs
// that pin its counts to the oracle's).
template <int MODE, int BOUND, int RNG>
static hipError_t launch3(const StepParams &p, uint32_t iters, int instr, bool moment,
                          hipStream_t s) {
    if (instr == kInstrTrace) {
        launch_m<MODE, BOUND, RNG, kInstrTrace>(p, iters, moment, s);
    } else if (instr == kInstrStats) {
        if constexpr (MODE == 1 && RNG != kRngMwc)
            launch_m<MODE, BOUND, RNG, kInstrStats>(p, iters, moment, s);
        else
            return hipErrorInvalidValue;
    } else {
        launch_m<MODE, BOUND, RNG, kInstrNone>(p, iters, moment, s);
    }
    return hipGetLastError();
}

template <int MODE, int BOUND>
static hipError_t launch2(const StepParams &p, uint32_t iters, int rng, int instr, bool moment,
                          hipStream_t s) {
    switch (rng) {
        case kRngInject: return launch3<MODE, BOUND, kRngInject>(p, iters, instr, moment, s);
        case kRngMwc: return launch3<MODE, BOUND, kRngMwc>(p, iters, instr, moment, s);
        default: return launch3<MODE, BOUND, kRngPhilox>(p, iters, instr, moment, s);
    }
}

hipError_t launch_step(const StepParams &p, uint32_t iters, int mode, int periodic, int rng,
                       int instr, bool moment, hipStream_t s) {
    if (p.n == 0 || iters == 0) return hipSuccess;
    if (mode == 0) return launch2<0, kBoundClamp>(p, iters, rng, instr, moment, s);
    if (!periodic) return launch2<1, kBoundClamp>(p, iters, rng, instr, moment, s);
    bool cube = true;
    for (int k = 0; k < 3; ++k) cube &= p.wrap_lo[k] == -1.0f && p.wrap_hi[k] == 1.0f;
    if (cube) return launch2<1, kBoundCube>(p, iters, rng, instr, moment, s);
    return launch2<1, kBoundBox>(p, iters, rng, instr, moment, s);
}

// ------------------------------------------------------------------------ //
// Phase of arbitrary gradient waveforms on the matrix cores.
//
// Over a chunk of T steps the phase increment of spin s under waveform w is
// sum_t pc G_w(t) . u_s(t) with u the unwrapped displacement: a GEMM of the
// trajectory (spins x K, K = 3T) by the waveforms (K x n_wave).  Both factors
// are split into bf16 hi + lo parts (x = hi + lo to about 16 bits) and the
// product is hi*hi + hi*lo + lo*hi, accumulated in FP32 by
// v_mfma_f32_32x32x16_bf16; the chunk's result is folded into the phase with
// Kahan compensation.
//
// Workgroup: 256 threads, tile of 128 spins x 128 waveforms, four waves of
// 64 x 64 (2 x 2 MFMA tiles of 32 x 32), whose (phase, compensation) pairs
// stay in registers for the whole launch (one wave per SIMD).  K advances in
// slices of 32 through two LDS stages: the trajectory slice (fp32, spins
// contiguous) is split and transposed into LDS as [spin][k], the waveform
// slice is copied as [waveform][k], so every MFMA operand (8 consecutive k of
// one row) is one 16-byte LDS read.
// v_mfma_f32_32x32x16_bf16 lane layout (checked on the device): A row l%32,
// B column l%32, k = 8*(l/32) + v for v < 8; result register r holds row
// 8*(r/4) + 4*(l/32) + r%4, column l%32.
// ------------------------------------------------------------------------ //
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr uint32_t kLdsPitch = 40;  // bf16 per LDS row: 32 k + 8 padding (80 bytes)

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Two floats to bf16 hi + lo parts, each packed two to a word (element 0 in
// the low half): v_cvt_pk_bf16_f32 rounds to nearest even, and lo is the
// rounded remainder x - hi, which is exact in FP32.
__device__ __forceinline__ void bf16_split2(float a, float b, uint32_t &hi, uint32_t &lo) {
    const f32x2 v = {a, b};
    hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
    const f32x2 r = {a - __uint_as_float(hi << 16), b - __uint_as_float(hi & 0xffff0000u)};
    lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bf16x2));
}

// Raw buffer descriptor over a wave-uniform base (gfx9 word 3; no stride, byte
// offsets below 2^31).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void *base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, 0x7fffffff, 0x00020000);
}

// Register stage of one K slice of the phase GEMM (see k_phase_gemm) for
// kPhaseThreads threads and a workgroup tile of 128 spins x 128 * TJ
// waveforms: 8 consecutive k of one spin of the trajectory, and TJ 16-byte
// pieces each of the hi and lo waveform rows.
constexpr uint32_t kPhaseThreads = 512;
template <int TJ> struct PhaseStage {
    float tv[8];
    u32x4 wh[TJ], wl[TJ];
};
// Loads of a slice through buffer descriptors with wave-uniform bases: the
// spin tile (a contiguous block of k_len rows of kTrajTile spins), the wave's
// k row akg (t / 128 is uniform in a wave) and the waveform tile n0 go into
// the scalar base, the row stride into the scalar offset, so the loads take
// no vector address arithmetic.  The waveform offsets stay below 2^31 bytes:
// 128 * TJ * k_total * 2 (checked at launch).
template <int TJ>
__device__ __forceinline__ void phase_gload(PhaseStage<TJ> &S, const float *__restrict__ traj,
                                            uint32_t k_len, uint32_t tm, uint32_t am,
                                            uint32_t akg, uint32_t ks,
                                            const uint16_t *__restrict__ g_hi,
                                            const uint16_t *__restrict__ g_lo, uint32_t n0,
                                            uint32_t k_total, uint32_t k0, uint32_t t) {
    // k = ks + akg .. + 7 of spin am: two 16-byte chunks of kTrajChunk k
    const __amdgpu_buffer_rsrc_t ra =
        raw_rsrc(traj + ((size_t)tm * k_len + ks + akg) * kTrajTile);
#pragma unroll
    for (uint32_t c = 0; c < 2; ++c) {
        const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(
            ra, am * 16u, c * kTrajTile * kTrajChunk * 4u, 0);
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) S.tv[4 * c + q] = __uint_as_float(v[q]);
    }
    const size_t gb = (size_t)n0 * k_total + k0 + ks;
    const __amdgpu_buffer_rsrc_t rh = raw_rsrc(g_hi + gb), rl = raw_rsrc(g_lo + gb);
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
        // 16-byte piece c: waveform row c / 4, quarter c % 4
        const uint32_t c = t + j * kPhaseThreads;
        const uint32_t go = ((c >> 2) * k_total + (c & 3u) * 8u) * 2u;
        S.wh[j] = __builtin_amdgcn_raw_buffer_load_b128(rh, go, 0, 0);
        S.wl[j] = __builtin_amdgcn_raw_buffer_load_b128(rl, go, 0, 0);
    }
}
// Trajectory split into bf16 hi + lo as [spin][k] rows (one 16-byte row piece
// each); waveform rows as loaded.
template <int TJ>
__device__ __forceinline__ void phase_lstore(const PhaseStage<TJ> &S, uint16_t *a_hi,
                                             uint16_t *a_lo, uint16_t *b_hi, uint16_t *b_lo,
                                             uint32_t am, uint32_t akg, uint32_t t) {
    uint32_t hq[4], lq[4];
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) bf16_split2(S.tv[2 * q], S.tv[2 * q + 1], hq[q], lq[q]);
    const u32x4 h = {hq[0], hq[1], hq[2], hq[3]}, l = {lq[0], lq[1], lq[2], lq[3]};
    *reinterpret_cast<u32x4 *>(&a_hi[am * kLdsPitch + akg]) = h;
    *reinterpret_cast<u32x4 *>(&a_lo[am * kLdsPitch + akg]) = l;
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
        const uint32_t c = t + j * kPhaseThreads, r0 = c >> 2, part = (c & 3u) * 8u;
        *reinterpret_cast<u32x4 *>(&b_hi[r0 * kLdsPitch + part]) = S.wh[j];
        *reinterpret_cast<u32x4 *>(&b_lo[r0 * kLdsPitch + part]) = S.wl[j];
    }
}

// One launch covers k_len trajectory rows (the steps of one trajectory
// buffer) in folds of k_fold rows: each fold's products are summed in the
// MFMA accumulators from zero and then added, with Kahan compensation, to the
// (phase, compensation) pairs, which stay in registers from the first fold
// to the last and touch memory once per launch.  The compensation is held
// rounded to bf16 (it is below half an ulp of the phase, so rounding it
// perturbs the sum by under 2^-10 ulp per fold), two to a word: the spins
// 2m and 2m + 1 of one waveform, which are elements q and q + 1 (q even) of
// one MFMA result tile, so each word loads and stores as it is held.
// Workgroup tile 128 spins x 128 * TJ waveforms in eight waves of 64 x 32 * TJ,
// two waves per SIMD.  The K loop stages slices of 32 through two LDS
// stages: the next slice's global loads are in flight while the current
// slice's MFMAs run, and its split and LDS stores follow, one barrier per
// slice.
template <int TJ>
__global__ __launch_bounds__(kPhaseThreads) void k_phase_gemm(
    const float *__restrict__ traj, uint32_t n_pad, uint32_t k_len, uint32_t k_fold,
    const uint16_t *__restrict__ g_hi, const uint16_t *__restrict__ g_lo, uint32_t k_total,
    uint32_t k0, uint32_t n_wave, float *__restrict__ phase, uint32_t *__restrict__ comp) {
    constexpr uint32_t kTileN = 128u * TJ;
    constexpr uint32_t kAHalves = kPhaseTileM * kLdsPitch, kBHalves = kTileN * kLdsPitch;
    constexpr uint32_t kStage = 2 * kAHalves + 2 * kBHalves;
    __shared__ __align__(16) uint16_t lds[2 * kStage];
    const uint32_t t = threadIdx.x, lane = t & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(t >> 6);
    // Workgroups go to the 8 XCDs in turn; the numbering below hands all
    // waveform tiles of one spin tile to one XCD back to back, so the
    // trajectory slice is read from HBM once and then from that XCD's L2.
    const uint32_t n_tiles = n_wave / kTileN, m_tiles = n_pad / kPhaseTileM;
    const uint32_t xcd = blockIdx.x % 8u, r = blockIdx.x / 8u;
    const uint32_t tm = (r / n_tiles) * 8u + xcd;
    if (tm >= m_tiles) return;
    const uint32_t n0 = (r % n_tiles) * kTileN;  // waveform tile
    const uint32_t m0 = tm * kPhaseTileM;        // spin tile
    const uint32_t wm = (wave & 1u) * 64u, wn = (wave >> 1) * 32u * TJ;  // wave tile in the WG tile
    // v_mfma_f32_32x32x16_bf16 result register q of lane l: row
    // 8*(q/4) + 4*(l/32) + q%4, column l%32 of the 32 x 32 tile
    const uint32_t am = t & 127u, akg = (wave >> 1) * 8u;
    auto stage = [&](uint32_t sl) -> uint16_t * { return lds + (sl & 1u) * kStage; };
    auto store = [&](const PhaseStage<TJ> &S, uint16_t *st) {
        phase_lstore<TJ>(S, st, st + kAHalves, st + 2 * kAHalves, st + 2 * kAHalves + kBHalves, am,
                         akg, t);
    };
    PhaseStage<TJ> stg;
    const uint32_t n_slices = k_len / kPhaseSliceK, fold_slices = k_fold / kPhaseSliceK;
    auto slice_k = [&](uint32_t sl) -> uint32_t { return min(sl, n_slices - 1) * kPhaseSliceK; };
    // slice 0 into stage 0 and slice 1 into registers, then the pairs (issued
    // after the slices' loads, so that waiting for those does not wait for
    // the pairs)
    phase_gload<TJ>(stg, traj, k_len, tm, am, akg, 0, g_hi, g_lo, n0, k_total, k0, t);
    store(stg, stage(0));
    phase_gload<TJ>(stg, traj, k_len, tm, am, akg, slice_k(1), g_hi, g_lo, n0, k_total, k0, t);
    float ps[2][TJ][16];   // phase
    uint32_t pk[2][TJ][8];  // compensation, bf16 of elements q (low half) and q + 1, q even
    // Pairs through buffer descriptors based at the wave tile's first spin
    // and waveform (wave-uniform), so no per-element addresses live across
    // the K loop: element (i, j, q) is row 32 i + 8 (q / 4) + 4 (l / 32) +
    // q % 4 of the wave tile, column 32 j + l % 32; the lane's row and column
    // and q % 4 go into four lane offsets, 32 i + 8 (q / 4) into the scalar
    // offset and 32 j into the immediate one.  phase has n_pad rows.
    const __amdgpu_buffer_rsrc_t rp = raw_rsrc(phase + (size_t)(m0 + wm) * n_wave + n0 + wn);
    const __amdgpu_buffer_rsrc_t rc = raw_rsrc(comp + (size_t)((m0 + wm) / 2) * n_wave + n0 + wn);
    uint32_t lp[4], lc[2];
#pragma unroll
    for (int r = 0; r < 4; ++r) lp[r] = ((4u * (lane >> 5) + r) * n_wave + (lane & 31u)) * 4u;
#pragma unroll
    for (int r = 0; r < 2; ++r) lc[r] = ((2u * (lane >> 5) + r) * n_wave + (lane & 31u)) * 4u;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const uint32_t sp = (32u * i + 8u * (q / 4)) * n_wave * 4u;
            const uint32_t sc = (16u * i + 4u * (q / 4)) * n_wave * 4u;
#pragma unroll
            for (int j = 0; j < TJ; ++j) {
                ps[i][j][q] = __uint_as_float(
                    __builtin_amdgcn_raw_buffer_load_b32(rp, lp[q % 4] + 128u * j, sp, 0));
                if (q % 2 == 0)
                    pk[i][j][q / 2] =
                        __builtin_amdgcn_raw_buffer_load_b32(rc, lc[(q % 4) / 2] + 128u * j, sc, 0);
            }
        }
    __syncthreads();
    f32x16 acc[2][TJ];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[i][j][q] = 0.0f;
    // K in folds of fold_slices slices: the slices of a fold, then its
    // Kahan fold (one loop shape, so the pairs stay in registers throughout)
    for (uint32_t f0 = 0; f0 < n_slices; f0 += fold_slices) {
        const uint32_t f1 = min(f0 + fold_slices, n_slices);
        for (uint32_t sl = f0; sl < f1; ++sl) {
            const uint16_t *cur = stage(sl);
            const uint16_t *a_hi = cur, *a_lo = cur + kAHalves;
            const uint16_t *b_hi = cur + 2 * kAHalves, *b_lo = b_hi + kBHalves;
#pragma unroll
            for (uint32_t kk = 0; kk < kPhaseSliceK; kk += 16) {
                const uint32_t kofs = kk + 8u * (lane >> 5);
                bf16x8 ah[2], al[2], bh[TJ], bl[TJ];
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const uint32_t row = wm + 32u * i + (lane & 31u);
                    ah[i] = *reinterpret_cast<const bf16x8 *>(&a_hi[row * kLdsPitch + kofs]);
                    al[i] = *reinterpret_cast<const bf16x8 *>(&a_lo[row * kLdsPitch + kofs]);
                }
#pragma unroll
                for (int j = 0; j < TJ; ++j) {
                    const uint32_t row = wn + 32u * j + (lane & 31u);
                    bh[j] = *reinterpret_cast<const bf16x8 *>(&b_hi[row * kLdsPitch + kofs]);
                    bl[j] = *reinterpret_cast<const bf16x8 *>(&b_lo[row * kLdsPitch + kofs]);
                }
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < TJ; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < TJ; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < TJ; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
            }
            // the next slice into the other stage while these MFMAs run, and the
            // one after it into registers (unconditionally, so that the compiler
            // can interleave the split with the MFMAs: after the last slice the
            // other stage is written and never read, and the loads stay on the
            // last slice)
            store(stg, stage(sl + 1));
            phase_gload<TJ>(stg, traj, k_len, tm, am, akg, slice_k(sl + 2), g_hi, g_lo, n0, k_total,
                            k0, t);
            __syncthreads();
        }
        // Kahan fold of this chunk's sums
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < TJ; ++j)
#pragma unroll
                for (int q = 0; q < 16; q += 2) {
                    float c[2] = {__uint_as_float(pk[i][j][q / 2] << 16),
                                  __uint_as_float(pk[i][j][q / 2] & 0xffff0000u)};
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const float y = acc[i][j][q + e] - c[e];
                        const float s1 = ps[i][j][q + e] + y;
                        c[e] = (s1 - ps[i][j][q + e]) - y;
                        ps[i][j][q + e] = s1;
                        acc[i][j][q + e] = 0.0f;
                    }
                    const f32x2 cv = {c[0], c[1]};
                    pk[i][j][q / 2] = __builtin_bit_cast(uint32_t, __builtin_convertvector(cv, bf16x2));
                }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const uint32_t sp = (32u * i + 8u * (q / 4)) * n_wave * 4u;
            const uint32_t sc = (16u * i + 4u * (q / 4)) * n_wave * 4u;
#pragma unroll
            for (int j = 0; j < TJ; ++j) {
                __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(ps[i][j][q]), rp,
                                                      lp[q % 4] + 128u * j, sp, 0);
                if (q % 2 == 0)
                    __builtin_amdgcn_raw_buffer_store_b32(pk[i][j][q / 2], rc,
                                                          lc[(q % 4) / 2] + 128u * j, sc, 0);
            }
        }
}

hipError_t launch_phase_gemm(const float *traj, uint32_t n, uint32_t n_pad, uint32_t k_len,
                             uint32_t k_fold, const uint16_t *g_hi, const uint16_t *g_lo,
                             uint32_t k_total, uint32_t k0, uint32_t n_wave, float *phase,
                             uint32_t *comp, hipStream_t s) {
    if (n == 0 || k_len == 0) return hipSuccess;
    if (k_fold == 0 || k_fold % kPhaseSliceK || k_len % kPhaseSliceK || n_wave % kPhaseTileN ||
        n_pad % kPhaseTileM)
        return hipErrorInvalidValue;
    // 256-waveform tiles where the waveforms divide into them
    const bool wide = n_wave % (2 * kPhaseTileN) == 0;
    const uint32_t tile_n = wide ? 2 * kPhaseTileN : kPhaseTileN;
    // byte offsets of the waveform loads (phase_gload) and of the pairs below 2^31
    if ((uint64_t)tile_n * k_total * 2 >= (1ull << 31) || (uint64_t)64 * n_wave * 4 >= (1ull << 31))
        return hipErrorInvalidValue;
    const uint32_t m_groups = (n_pad / kPhaseTileM + 7u) / 8u;
    const uint32_t grid = m_groups * 8u * (n_wave / tile_n);
    if (wide)
        hipLaunchKernelGGL(k_phase_gemm<2>, dim3(grid), dim3(kPhaseThreads), 0, s, traj, n_pad,
                           k_len, k_fold, g_hi, g_lo, k_total, k0, n_wave, phase, comp);
    else
        hipLaunchKernelGGL(k_phase_gemm<1>, dim3(grid), dim3(kPhaseThreads), 0, s, traj, n_pad,
                           k_len, k_fold, g_hi, g_lo, k_total, k0, n_wave, phase, comp);
    return hipGetLastError();
}

// Readout of the waveform phases: one thread per waveform, one block row per
// range of spins; FP64 phase (x0 term added in FP64), cos/sin and sums.
__global__ __launch_bounds__(256) void k_phase_partial(uint32_t n, const float *__restrict__ phase,
                                                       uint32_t n_wave,
                                                       const double *__restrict__ gsum,
                                                       const float *__restrict__ mag,
                                                       const uint32_t *__restrict__ label,
                                                       const float *__restrict__ x0,
                                                       const uint32_t *__restrict__ id,
                                                       uint32_t per_block, double *partial) {
    const uint32_t w = blockIdx.x * 256u + threadIdx.x;
    const uint32_t sb = blockIdx.y;
    if (w >= n_wave) return;
    const double gx = gsum[3 * w], gy = gsum[3 * w + 1], gz = gsum[3 * w + 2];
    double v[15];
    for (int k = 0; k < 15; ++k) v[k] = 0.0;
    const uint32_t s1 = min(n, (sb + 1) * per_block);
    for (uint32_t s = sb * per_block; s < s1; ++s) {
        const double ph = (double)phase[(size_t)id[s] * n_wave + w] + gx * (double)x0[s] +
                          gy * (double)x0[n + s] + gz * (double)x0[2 * n + s];
        double sn, cs;
        sincos(ph, &sn, &cs);
        const double m = (double)mag[s];
        const uint32_t c = (label[s] & 0xFFu) + 1u;
        v[0] += m * cs; v[5] += m * sn; v[10] += 1.0;
        if (c < 5) { v[c] += m * cs; v[5 + c] += m * sn; v[10 + c] += 1.0; }
    }
    double *o = partial + ((size_t)sb * n_wave + w) * 15;
    for (int k = 0; k < 15; ++k) o[k] = v[k];
}

__global__ __launch_bounds__(256) void k_phase_final(const double *partial, uint32_t n_blocks,
                                                     uint32_t n_wave, double *out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;  // waveform*15 + quantity
    if (i >= n_wave * 15u) return;
    double acc = 0.0;
    for (uint32_t b = 0; b < n_blocks; ++b) acc += partial[(size_t)b * n_wave * 15 + i];
    out[i] = acc;
}

hipError_t launch_phase_signal(uint32_t n, const float *phase, uint32_t n_wave,
                               const double *gsum, const float *mag, const uint32_t *label,
                               const float *x0, const uint32_t *id, double *partial,
                               uint32_t n_blocks, double *out, hipStream_t s) {
    if (n_wave == 0 || n_blocks == 0) return hipSuccess;
    const uint32_t per = (n + n_blocks - 1) / n_blocks;
    hipLaunchKernelGGL(k_phase_partial, dim3((n_wave + 255) / 256, n_blocks), dim3(256), 0, s, n,
                       phase, n_wave, gsum, mag, label, x0, id, per, partial);
    hipLaunchKernelGGL(k_phase_final, dim3((n_wave * 15 + 255) / 256), dim3(256), 0, s, partial,
                       n_blocks, n_wave, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------------ //
// Signal readout: getMrSignal(double[]) (spinSystem.cpp:1452-1500), FP64.
// Per-block partials then one fixed-order final pass (deterministic).
// ------------------------------------------------------------------------ //
__device__ __forceinline__ void block_reduce15(double (&v)[15], double *out) {
    __shared__ double sh[15][256];
    const int t = threadIdx.x;
    for (int k = 0; k < 15; ++k) sh[k][t] = v[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s)
            for (int k = 0; k < 15; ++k) sh[k][t] += sh[k][t + s];
        __syncthreads();
    }
    if (t < 15) out[t] = sh[t][0];
}

__global__ __launch_bounds__(256) void k_signal_partial(uint32_t n, const float *mag,
                                                        const float *phase, const uint32_t *label,
                                                        double *partial) {
    double v[15];
    for (int k = 0; k < 15; ++k) v[k] = 0.0;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double m = (double)mag[i];
        double c = cos((double)phase[i]) * m, s = sin((double)phase[i]) * m;
        uint32_t k = (label[i] & 0xFFu) + 1u;
        v[0] += c; v[5] += s; v[10] += 1.0;
        for (uint32_t q = 1; q < 5; ++q) {
            bool on = (q == k);
            v[q] += on ? c : 0.0;
            v[5 + q] += on ? s : 0.0;
            v[10 + q] += on ? 1.0 : 0.0;
        }
    }
    block_reduce15(v, partial + (size_t)blockIdx.x * 15);
}

__global__ __launch_bounds__(256) void k_partial_final(const double *partial, uint32_t n_blocks,
                                                       double *out) {
    // out[group*15 + k] = sum over blocks of partial[(group*n_blocks + b)*15 + k]
    const uint32_t g = blockIdx.x;
    double v[15];
    for (int k = 0; k < 15; ++k) v[k] = 0.0;
    for (uint32_t b = threadIdx.x; b < n_blocks; b += 256)
        for (int k = 0; k < 15; ++k) v[k] += partial[((size_t)g * n_blocks + b) * 15 + k];
    block_reduce15(v, out + (size_t)g * 15);
}

hipError_t launch_signal(uint32_t n, const float *mag, const float *phase, const uint32_t *label,
                         double *partial, uint32_t n_blocks, double *out15, hipStream_t s) {
    hipLaunchKernelGGL(k_signal_partial, dim3(n_blocks), dim3(256), 0, s, n, mag, phase, label,
                       partial);
    hipLaunchKernelGGL(k_partial_final, dim3(1), dim3(256), 0, s, partial, n_blocks, out15);
    return hipGetLastError();
}

// Per-direction readout from gradient moments: phase_j = pc * g_j.(M + W x0).
__global__ __launch_bounds__(256) void k_moment_partial(uint32_t n, const float *mag,
                                                        const float *mom, const float *x0,
                                                        float wsum, const uint32_t *label,
                                                        const float *grads, float pc,
                                                        double *partial) {
    const uint32_t j = blockIdx.y;
    const float gx = grads[3 * j], gy = grads[3 * j + 1], gz = grads[3 * j + 2];
    double v[15];
    for (int k = 0; k < 15; ++k) v[k] = 0.0;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        float mx = mom[i], my = mom[n + i], mz = mom[2 * n + i];
        if (x0 && wsum != 0.0f) {
            mx = fmaf(wsum, x0[i], mx);
            my = fmaf(wsum, x0[n + i], my);
            mz = fmaf(wsum, x0[2 * n + i], mz);
        }
        float ph = fmaf(gx, mx, fmaf(gy, my, gz * mz)) * pc;
        double m = (double)mag[i];
        double c = cos((double)ph) * m, s = sin((double)ph) * m;
        uint32_t k = (label[i] & 0xFFu) + 1u;
        v[0] += c; v[5] += s; v[10] += 1.0;
        for (uint32_t q = 1; q < 5; ++q) {
            bool on = (q == k);
            v[q] += on ? c : 0.0;
            v[5 + q] += on ? s : 0.0;
            v[10 + q] += on ? 1.0 : 0.0;
        }
    }
    block_reduce15(v, partial + ((size_t)j * gridDim.x + blockIdx.x) * 15);
}

hipError_t launch_moment_signal(uint32_t n, const float *mag, const float *mom, const float *x0,
                                float wsum, const uint32_t *label, uint32_t ndir,
                                const float *grads, float pc, double *partial, uint32_t n_blocks,
                                double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_moment_partial, dim3(n_blocks, ndir), dim3(256), 0, s, n, mag, mom, x0,
                       wsum, label, grads, pc, partial);
    hipLaunchKernelGGL(k_partial_final, dim3(ndir), dim3(256), 0, s, partial, n_blocks, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------------ //
// Labelling: findSpinsInFibers (spinSystem.cpp:1615-1852).  Every triangle
// whose box overlaps the ray's box is a candidate (the R-tree query); the
// crossing parity of each (fibre, membrane) group is folded into XOR sums so
// the candidate order does not matter when at most one fibre per membrane
// type is crossed an odd number of times (detected otherwise).
// ------------------------------------------------------------------------ //
// Longest re-cast queue taken one spin per block (k_label_spin).
constexpr uint32_t kRetryPerBlockMax = 512;

__device__ __forceinline__ uint32_t fiber_hash(uint32_t f) { return f * 0x9E3779B1u + 0x7F4A7C15u; }

// The end point of spin i's ray on try `tries`: to 1.1 * max along the
// substrate's ray axis (the physical mode) or along the spin's dominant axis
// (findSpinsInFibers), then for re-casts moved by a Philox jitter of up to
// 0.05 per axis.
__device__ __forceinline__ void label_ray(const LabelParams &P, uint32_t i, uint32_t tries,
                                          float sx, float sy, float sz, float &ex, float &ey,
                                          float &ez) {
    ex = sx; ey = sy; ez = sz;
    if (P.ray_axis == 1) {
        ex = (float)(1.1 * (double)P.xmax);
    } else if (P.ray_axis == 2) {
        ey = (float)(1.1 * (double)P.ymax);
    } else if (P.ray_axis == 3) {
        ez = (float)(1.1 * (double)P.zmax);
    } else if ((fabsf(sx) > fabsf(sy)) & (fabsf(sx) > fabsf(sz))) {
        ex = (float)(((sx >= 0) - (sx < 0)) * 1.1 * (double)P.xmax);
    } else if (fabsf(sy) > fabsf(sz)) {
        ey = (float)(((sy >= 0) - (sy < 0)) * 1.1 * (double)P.ymax);
    } else {
        ez = (float)(((sz >= 0) - (sz < 0)) * 1.1 * (double)P.zmax);
    }
    if (tries > 0) {
        uint4 o = philox(make_uint4(i, tries, 0u, 0u), 0x6A09E667u, 0xBB67AE85u);
        ex = (float)((double)ex + (double)u32_unit(o.x) * 0.05);
        ey = (float)((double)ey + (double)u32_unit(o.y) * 0.05);
        ez = (float)((double)ez + (double)u32_unit(o.z) * 0.05);
    }
}

// One ray's crossing sums: per membrane type the XOR of (fibre + 1) and of the
// fibre hash over crossed triangles, and the count of edge or vertex hits.
struct RaySums {
    uint32_t xf[3], xh[3], amb[3];
};

// Adds triangle t (box lo/hi, fibre and membrane in lo.w) if its box overlaps
// the ray's box and the ray crosses it.
__device__ __forceinline__ void ray_add(const LabelParams &P, uint32_t t, float4 lo, float4 hi,
                                        float sx, float sy, float sz, float ex, float ey,
                                        float ez, RaySums &R) {
    const float rminx = fminf(sx, ex), rminy = fminf(sy, ey), rminz = fminf(sz, ez);
    const float rmaxx = fmaxf(sx, ex), rmaxy = fmaxf(sy, ey), rmaxz = fmaxf(sz, ez);
    if (rminx > hi.x || lo.x > rmaxx || rminy > hi.y || lo.y > rmaxy || rminz > hi.z ||
        lo.z > rmaxz)
        return;
    Hit h;
    if (P.use_mt)
        test_mt(P.tri_mt, t, sx, sy, sz, ex, ey, ez, h);
    else
        test_ref(P.tri_ref, t, sx, sy, sz, ex, ey, ez, h);
    if (h.type == 0) return;
    const uint32_t info = __float_as_uint(lo.w);
    const uint32_t f = info & 0xFFFFu, m = info >> 16;
    if (m > 2) return;
    R.xf[m] ^= f + 1u;
    R.xh[m] ^= fiber_hash(f);
    if (h.type >= 2) R.amb[m] += 1;
}

// The label a ray's sums give: the first membrane type (axon, myelin, glia)
// crossed an odd number of times names the compartment and fibre.  Returns
// false when the ray is ambiguous (an edge or vertex hit, or two fibres of one
// membrane type crossed an odd number of times).
__device__ __forceinline__ bool ray_label(const RaySums &R, uint32_t &comp, uint32_t &fib) {
    comp = 0;
    fib = kFiberNone;
    uint32_t nedge = 0;
    bool found = false, multi = false;
    for (uint32_t m = 0; m < 3 && !found; ++m) {
        nedge += R.amb[m];
        if (R.xf[m] != 0u) {
            const uint32_t f = R.xf[m] - 1u;
            multi |= fiber_hash(f) != R.xh[m];
            found = true;
            fib = f;
            comp = (m == 0) ? 1 : (m == 2) ? 3 : 2;
        } else if (R.xh[m] != 0u) {
            multi = true;
        }
    }
    return nedge == 0 && !multi;
}

// Every spin, every triangle: one thread per spin, triangle boxes staged
// through LDS in tiles of 256.
__global__ __launch_bounds__(256) void k_label(LabelParams P) {
    __shared__ float4 box[2][256];
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const bool live = i < P.n;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    if (live) { sx = P.x[i]; sy = P.y[i]; sz = P.z[i]; }
    uint32_t tries = 0;
    bool done = !live;
    uint32_t out_comp = 0, out_fiber = kFiberNone;
    while (__syncthreads_or(!done)) {
        float ex, ey, ez;
        label_ray(P, i, tries, sx, sy, sz, ex, ey, ez);
        RaySums R = {};
        for (uint32_t t0 = 0; t0 < P.n_tri; t0 += 256) {
            __syncthreads();
            uint32_t tl = t0 + threadIdx.x;
            if (tl < P.n_tri) {
                box[0][threadIdx.x] = P.tri_aabb[2 * tl];
                box[1][threadIdx.x] = P.tri_aabb[2 * tl + 1];
            }
            __syncthreads();
            if (done) continue;
            uint32_t cnt = min(256u, P.n_tri - t0);
            for (uint32_t k = 0; k < cnt; ++k)
                ray_add(P, t0 + k, box[0][k], box[1][k], sx, sy, sz, ex, ey, ez, R);
        }
        if (!done) {
            const bool clear = ray_label(R, out_comp, out_fiber);
            if (clear) {
                done = true;
            } else {
                tries++;
                if (tries > P.max_tries) done = true;
            }
        }
    }
    if (live) {
        P.label[i] = out_comp | (out_fiber << 8);
        P.tries[i] = tries;
    }
}

hipError_t launch_label(const LabelParams &p, hipStream_t s) {
    if (p.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_label, dim3((p.n + 255) / 256), dim3(256), 0, s, p);
    return hipGetLastError();
}

// Is the triangle (box lo, hi) taken by the outside list rather than by the
// row walk?  AABB grids: its box starts past the grid's face on the ray's
// side (no grid row lists it in a cell the ray can reach).  Exact-membership
// grids: its box reaches past that face (the ray may cross it beyond the grid,
// where no row cell lists it).
__device__ __forceinline__ bool past_grid(const LabelParams &P, float4 lo, float4 hi) {
    const int a = (int)P.ray_axis - 1;
    if (P.exact_grid) return (a == 0 ? hi.x : a == 1 ? hi.y : hi.z) > P.hi[a];
    return (a == 0 ? lo.x : a == 1 ? lo.y : lo.z) > P.hi[a];
}

__global__ __launch_bounds__(256) void k_label_outside(LabelParams P) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= P.n_tri) return;
    if (past_grid(P, P.tri_aabb[2 * t], P.tri_aabb[2 * t + 1]))
        P.outside[atomicAdd(P.n_outside, 1u)] = t;
}

// Cell of coordinate v on axis a whose membership box, lo + c cell to
// lo + (c+1) cell in float as the grid builds compute it, holds v.
__device__ __forceinline__ uint32_t label_cell(const LabelParams &P, float v, int a) {
    uint32_t c = clamp_cell(floorf((v - P.lo[a]) * P.inv_cell[a]), P.dims[a]);
    while (c > 0 && v < P.lo[a] + (float)c * P.cell[a]) --c;
    while (c + 1 < P.dims[a] && v > P.lo[a] + (float)(c + 1) * P.cell[a]) ++c;
    return c;
}

// Is triangle t listed in cell c (lists are in ascending triangle order)?
__device__ __forceinline__ bool listed(const LabelParams &P, uint32_t c, uint32_t t) {
    uint32_t a = P.cell_start[c], b = P.cell_start[c + 1];
    while (a < b) {
        const uint32_t m = (a + b) >> 1;
        const uint32_t v = P.cell_tri[m];
        if (v == t) return true;
        if (v < t) a = m + 1; else b = m;
    }
    return false;
}

// Lanes that share one spin's row walk (consecutive lanes of a wave).
constexpr uint32_t kRowLanes = 8;

// The first ray of each spin along the substrate's ray axis A (k_label's
// try 0): candidates are the triangles listed in the cells of the ray's grid
// row from the spin's cell on, each considered once (in the first of those
// cells that lists it), plus the triangles whose boxes start past the grid's
// face on the ray's side.  For an AABB grid these are k_label's candidates:
// those are the triangles whose boxes hold a point of the ray; for one that
// does not start past the grid, the point of the ray at max(box lo_A, s_A)
// lies in the grid box, so the cell of the row holding it meets the
// triangle's box and lists it (the grid lists a triangle in every cell whose
// closed membership box meets its bounding box, and the cells' boxes tile the
// grid box).  An exact-membership grid lists a triangle that lies inside the
// grid box on the ray's side in the cell of the row that holds a crossing,
// and in a contiguous run of the row's cells (a convex triangle meets a
// convex row of boxes in an interval), so every crossing is counted once; a
// triangle reaching past that face is taken by the outside list.  The kRowLanes lanes of a spin take every
// kRowLanes-th cell and outside triangle and combine their sums; the sums do
// not depend on the order, so the result equals k_label's first try.  Spins
// with an ambiguous ray, outside the grid box, or beyond the ray's end are
// queued for the re-cast kernels.
__global__ __launch_bounds__(256) void k_label_row(LabelParams P) {
    const uint32_t gid = blockIdx.x * 256 + threadIdx.x;
    const uint32_t i = gid / kRowLanes, g = gid % kRowLanes;
    const bool live = i < P.n;  // every lane stays for the shuffles below
    const int A = (int)P.ray_axis - 1;
    float s3[3] = {0.f, 0.f, 0.f};
    if (live) { s3[0] = P.x[i]; s3[1] = P.y[i]; s3[2] = P.z[i]; }
    const float sx = s3[0], sy = s3[1], sz = s3[2];
    float ex, ey, ez;
    label_ray(P, i, 0u, sx, sy, sz, ex, ey, ez);
    const float eA = A == 0 ? ex : A == 1 ? ey : ez;
    // the row walk covers rays that start in the grid box and run towards +A
    const bool row = live && sx >= P.lo[0] && sx <= P.hi[0] && sy >= P.lo[1] &&
                     sy <= P.hi[1] && sz >= P.lo[2] && sz <= P.hi[2] && s3[A] <= eA;
    RaySums R = {};
    if (row) {
        uint32_t c3[3] = {label_cell(P, sx, 0), label_cell(P, sy, 1), label_cell(P, sz, 2)};
        const uint32_t stride[3] = {1u, P.dims[0], P.dims[0] * P.dims[1]};
        const uint32_t cA0 = c3[A];
        c3[A] = 0;
        const uint32_t rowc = (c3[2] * P.dims[1] + c3[1]) * P.dims[0] + c3[0];
        const uint32_t sA = stride[A];
        for (uint32_t cA = cA0 + g; cA < P.dims[A]; cA += kRowLanes) {
            const uint32_t c = rowc + cA * sA;
            for (uint32_t k = P.cell_start[c]; k < P.cell_start[c + 1]; ++k) {
                const uint32_t t = P.cell_tri[k];
                const float4 lo = P.tri_aabb[2 * t], hi = P.tri_aabb[2 * t + 1];
                if (past_grid(P, lo, hi)) continue;              // in the outside list
                if (cA > cA0 && listed(P, c - sA, t)) continue;  // considered one cell back
                ray_add(P, t, lo, hi, sx, sy, sz, ex, ey, ez, R);
            }
        }
        const uint32_t no = *P.n_outside;
        for (uint32_t j = g; j < no; j += kRowLanes) {
            const uint32_t t = P.outside[j];
            ray_add(P, t, P.tri_aabb[2 * t], P.tri_aabb[2 * t + 1], sx, sy, sz, ex, ey, ez, R);
        }
    }
    for (uint32_t off = 1; off < kRowLanes; off <<= 1)
        for (uint32_t m = 0; m < 3; ++m) {
            R.xf[m] ^= __shfl_xor(R.xf[m], off, 64);
            R.xh[m] ^= __shfl_xor(R.xh[m], off, 64);
            R.amb[m] += __shfl_xor(R.amb[m], off, 64);
        }
    if (!live || g != 0) return;
    if (!row) {
        P.retry[atomicAdd(P.n_retry, 1u)] = make_uint2(i, 0u);
        return;
    }
    uint32_t comp, fib;
    if (ray_label(R, comp, fib)) {
        P.label[i] = comp | (fib << 8);
        P.tries[i] = 0;
    } else if (P.max_tries == 0) {  // no re-cast allowed: keep the first result
        P.label[i] = comp | (fib << 8);
        P.tries[i] = 1;
    } else {
        P.retry[atomicAdd(P.n_retry, 1u)] = make_uint2(i, 1u);  // re-cast with a jitter
    }
}

// Spins of one block of the long-queue re-cast kernel, and lanes per spin.
constexpr uint32_t kGroupSpins = 16, kGroupLanes = 256 / kGroupSpins;

// Queued spins of the row labeller, from their recorded try on, when the
// queue is long: kGroupSpins spins per block, kGroupLanes lanes per spin
// sharing out each LDS tile of triangle boxes, the lanes' sums combined by
// shuffles after every try.  Blocks take the queue's groups in turn.
__global__ __launch_bounds__(256) void k_label_group(LabelParams P) {
    __shared__ float4 box[2][256];
    const uint32_t count = *P.n_retry;
    if (count <= kRetryPerBlockMax) return;  // k_label_spin takes a short queue
    const uint32_t sp = threadIdx.x / kGroupLanes, g = threadIdx.x % kGroupLanes;
    for (uint32_t base = blockIdx.x * kGroupSpins; base < count;
         base += gridDim.x * kGroupSpins) {  // block-uniform
        const uint32_t slot = base + sp;
        const bool live = slot < count;
        const uint2 e = live ? P.retry[slot] : make_uint2(0u, 0u);
        const uint32_t i = e.x;
        uint32_t tries = e.y;
        float sx = 0.f, sy = 0.f, sz = 0.f;
        if (live) { sx = P.x[i]; sy = P.y[i]; sz = P.z[i]; }
        bool done = !live;  // uniform over the spin's lanes
        uint32_t out_comp = 0, out_fiber = kFiberNone;
        while (__syncthreads_or(!done)) {
            float ex, ey, ez;
            label_ray(P, i, tries, sx, sy, sz, ex, ey, ez);
            RaySums R = {};
            for (uint32_t t0 = 0; t0 < P.n_tri; t0 += 256) {
                __syncthreads();
                const uint32_t tl = t0 + threadIdx.x;
                if (tl < P.n_tri) {
                    box[0][threadIdx.x] = P.tri_aabb[2 * tl];
                    box[1][threadIdx.x] = P.tri_aabb[2 * tl + 1];
                }
                __syncthreads();
                if (done) continue;
                const uint32_t cnt = min(256u, P.n_tri - t0);
                for (uint32_t k = g; k < cnt; k += kGroupLanes)
                    ray_add(P, t0 + k, box[0][k], box[1][k], sx, sy, sz, ex, ey, ez, R);
            }
            for (uint32_t off = 1; off < kGroupLanes; off <<= 1)
                for (uint32_t m = 0; m < 3; ++m) {
                    R.xf[m] ^= __shfl_xor(R.xf[m], off, 64);
                    R.xh[m] ^= __shfl_xor(R.xh[m], off, 64);
                    R.amb[m] += __shfl_xor(R.amb[m], off, 64);
                }
            if (!done) {
                if (ray_label(R, out_comp, out_fiber)) {
                    done = true;
                } else {
                    tries++;
                    if (tries > P.max_tries) done = true;
                }
            }
        }
        if (live && g == 0) {
            P.label[i] = out_comp | (out_fiber << 8);
            P.tries[i] = tries;
        }
    }
}

// Queued spins when the queue is short: one block per spin, its 256 threads
// sharing out the triangles, so a spin's re-casts take microseconds.
__global__ __launch_bounds__(256) void k_label_spin(LabelParams P) {
    __shared__ uint32_t part[4][9];
    const uint32_t count = *P.n_retry;
    if (count > kRetryPerBlockMax) return;  // k_label_group takes a long queue
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    for (uint32_t q = blockIdx.x; q < count; q += gridDim.x) {  // block-uniform
        const uint2 e = P.retry[q];
        const uint32_t i = e.x;
        uint32_t tries = e.y;
        const float sx = P.x[i], sy = P.y[i], sz = P.z[i];
        uint32_t out_comp = 0, out_fiber = kFiberNone;
        for (;;) {
            float ex, ey, ez;
            label_ray(P, i, tries, sx, sy, sz, ex, ey, ez);
            RaySums R = {};
            for (uint32_t t = threadIdx.x; t < P.n_tri; t += 256)
                ray_add(P, t, P.tri_aabb[2 * t], P.tri_aabb[2 * t + 1], sx, sy, sz, ex, ey, ez,
                        R);
            for (uint32_t off = 1; off < 64; off <<= 1)
                for (uint32_t m = 0; m < 3; ++m) {
                    R.xf[m] ^= __shfl_xor(R.xf[m], off, 64);
                    R.xh[m] ^= __shfl_xor(R.xh[m], off, 64);
                    R.amb[m] += __shfl_xor(R.amb[m], off, 64);
                }
            if (lane == 0)
                for (uint32_t m = 0; m < 3; ++m) {
                    part[w][m] = R.xf[m];
                    part[w][3 + m] = R.xh[m];
                    part[w][6 + m] = R.amb[m];
                }
            __syncthreads();
            RaySums T = {};
            for (uint32_t v = 0; v < 4; ++v)
                for (uint32_t m = 0; m < 3; ++m) {
                    T.xf[m] ^= part[v][m];
                    T.xh[m] ^= part[v][3 + m];
                    T.amb[m] += part[v][6 + m];
                }
            __syncthreads();  // every thread has read the partials
            // every thread reaches the same decision from the same sums
            if (ray_label(T, out_comp, out_fiber)) break;
            tries++;
            if (tries > P.max_tries) break;
        }
        if (threadIdx.x == 0) {
            P.label[i] = out_comp | (out_fiber << 8);
            P.tries[i] = tries;
        }
    }
}

hipError_t launch_label_row(const LabelParams &p, hipStream_t s) {
    if (p.n == 0) return hipSuccess;
    if (p.n_tri)
        hipLaunchKernelGGL(k_label_outside, dim3((p.n_tri + 255) / 256), dim3(256), 0, s, p);
    const uint64_t lanes = (uint64_t)p.n * kRowLanes;
    hipLaunchKernelGGL(k_label_row, dim3((uint32_t)((lanes + 255) / 256)), dim3(256), 0, s, p);
    // the re-casts: the kernel that does not fit the queue's length exits at once
    hipLaunchKernelGGL(k_label_spin, dim3(kRetryPerBlockMax), dim3(256), 0, s, p);
    const uint32_t groups = (p.n + kGroupSpins - 1) / kGroupSpins;
    hipLaunchKernelGGL(k_label_group, dim3(groups < 4096u ? groups : 4096u), dim3(256), 0, s, p);
    return hipGetLastError();
}

// MWC seed pairs for spin i: the first two words of Philox block (i, 0, 0, 3);
// a zero word (which would stall its multiply-with-carry half) becomes 1.
__global__ __launch_bounds__(256) void k_init_seeds(uint32_t n, uint32_t *seed, uint32_t k0,
                                                    uint32_t k1) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint4 r = philox(make_uint4(i, 0u, 0u, 3u), k0, k1);
    seed[2 * i] = r.x ? r.x : 1u;
    seed[2 * i + 1] = r.y ? r.y : 1u;
}

hipError_t launch_init_seeds(uint32_t n, uint32_t *seed, uint32_t key0, uint32_t key1,
                             hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_init_seeds, dim3((n + 255) / 256), dim3(256), 0, s, n, seed, key0, key1);
    return hipGetLastError();
}

// resetSpins position draw (spinSystem.cpp:1386-1398) from Philox stream 2.
// Per axis: width * (u - 0.5) + centre, u in (0, 1] (the box [-b, b] has
// width 2b and centre 0, which leaves resetSpins' expression unchanged).
__global__ __launch_bounds__(256) void k_init_uniform(uint32_t n, float *x, float *y, float *z,
                                                      float *mag, float *phase, uint32_t *label,
                                                      uint32_t *id, uint32_t *wrap, float3 width,
                                                      float3 centre, uint32_t k0, uint32_t k1,
                                                      uint32_t d0, uint32_t d1) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    id[i] = i;
    uint4 o = philox(make_uint4(i, d0, d1, 2u), k0, k1);
    x[i] = width.x * (u32_unit(o.x) - 0.5f) + centre.x;
    y[i] = width.y * (u32_unit(o.y) - 0.5f) + centre.y;
    z[i] = width.z * (u32_unit(o.z) - 0.5f) + centre.z;
    mag[i] = 1.0f;
    phase[i] = 0.0f;
    label[i] = kFiberNone << 8;
    wrap[i] = 0u;
}

hipError_t launch_init_uniform(uint32_t n, float *x, float *y, float *z, float *mag, float *phase,
                               uint32_t *label, uint32_t *id, uint32_t *wrap, const float lo[3],
                               const float hi[3], uint32_t key0, uint32_t key1, uint64_t draw,
                               hipStream_t s) {
    if (n == 0) return hipSuccess;
    const float3 w = make_float3(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]);
    const float3 c = make_float3(0.5f * (hi[0] + lo[0]), 0.5f * (hi[1] + lo[1]),
                                 0.5f * (hi[2] + lo[2]));
    hipLaunchKernelGGL(k_init_uniform, dim3((n + 255) / 256), dim3(256), 0, s, n, x, y, z, mag,
                       phase, label, id, wrap, w, c, key0, key1, (uint32_t)draw,
                       (uint32_t)(draw >> 32));
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_moment_reset(uint32_t n, const float *x, const float *y,
                                                      const float *z, float *x0, float *disp,
                                                      float *mom, float *momc, float *mag) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    mag[i] = 1.0f;
    x0[i] = x[i]; x0[n + i] = y[i]; x0[2 * n + i] = z[i];
    for (int k = 0; k < 3; ++k) {
        disp[k * n + i] = 0.0f;
        mom[k * n + i] = 0.0f;
        momc[k * n + i] = 0.0f;
    }
}

hipError_t launch_moment_reset(uint32_t n, const float *x, const float *y, const float *z,
                               float *x0, float *disp, float *mom, float *momc, float *mag,
                               hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_moment_reset, dim3((n + 255) / 256), dim3(256), 0, s, n, x, y, z, x0,
                       disp, mom, momc, mag);
    return hipGetLastError();
}

// ------------------------------------------------------------------------ //
// Spatial reordering of spins (Morton order of their grid cells), so that the
// lanes of a wavefront walk the same cells and triangle records.
// ------------------------------------------------------------------------ //
__device__ __forceinline__ uint32_t spread3(uint32_t v) {  // 10 bits -> every third bit
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Sort key of a spin: its membrane-hit rate since the last reorder in 16
// bands, highest first, then the Morton index of its cell.  Spins that hit at
// like rates share a wavefront, so no lane holds the wave back for long; the
// busiest workgroups are dispatched first; and within a band the Morton order
// keeps neighbouring blocks on neighbouring triangles.  Resets the hit counts.
__global__ __launch_bounds__(256) void k_sort_key(StepParams P, uint32_t *keys, uint32_t *vals) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= P.n) return;
    uint32_t cx, cy, cz;
    cell_of<1>(P, P.x[i], P.y[i], P.z[i], cx, cy, cz);
    uint32_t band = 0;
    if (P.hits) {
        if (P.hit_window)
            band = (uint32_t)min<uint64_t>(15, (uint64_t)P.hits[i] * 16u / P.hit_window);
        P.hits[i] = 0;
    }
    const uint32_t morton = spread3(cx) | (spread3(cy) << 1) | (spread3(cz) << 2);
    keys[i] = ((15u - band) << 28) | (morton >> 2);
    vals[i] = i;
}

// Spin reorder: a stable LSD radix sort of the 32-bit keys with the slot
// index as value, four passes of 8-bit digits.  Each pass: per-tile digit
// histograms (tiles of 4096 keys), one exclusive scan over them in
// digit-major order (every tile's count of digit 0, then of digit 1, ...),
// then each tile scatters its keys to its digit offsets, ranking equal digits
// in key order.  Within a wave the lanes holding the same digit are found
// with one ballot per digit bit; the waves of a tile combine their counts
// through LDS.
constexpr uint32_t kSortThreads = 256, kSortPerThread = 16;
constexpr uint32_t kSortTile = kSortThreads * kSortPerThread;  // keys per tile
constexpr uint32_t kRadix = 256;

__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint32_t *__restrict__ keys,
                                                             uint32_t n, uint32_t shift,
                                                             uint32_t n_tiles,
                                                             uint32_t *__restrict__ hist) {
    __shared__ uint32_t cnt[kRadix];
    const uint32_t t = threadIdx.x, tile = blockIdx.x;
    cnt[t] = 0;
    __syncthreads();
    for (uint32_t k = 0; k < kSortPerThread; ++k) {
        const uint32_t i = tile * kSortTile + k * kSortThreads + t;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    hist[t * n_tiles + tile] = cnt[t];
}

// Exclusive scan of m values in place, in blocks of 256: each block's sum,
// an exclusive scan of the block sums (one workgroup, 256 at a time with a
// running carry), then each block's own scan plus its offset.
constexpr uint32_t kScanBlock = 256;

// Inclusive sum over a 256-thread block (wave scans, wave totals through LDS).
__device__ __forceinline__ uint32_t block_incl_sum(uint32_t v, uint32_t *wsum) {
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    const uint32_t x = wave_incl_sum(v);
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t off = 0;
    for (uint32_t q = 0; q < w; ++q) off += wsum[q];
    __syncthreads();
    return x + off;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const uint32_t *__restrict__ v,
                                                            uint32_t m, uint32_t *__restrict__ part) {
    __shared__ uint32_t wsum[kScanBlock / 64];
    const uint32_t i = blockIdx.x * kScanBlock + threadIdx.x;
    const uint32_t x = block_incl_sum(i < m ? v[i] : 0u, wsum);
    if (threadIdx.x == kScanBlock - 1) part[blockIdx.x] = x;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_parts(uint32_t *__restrict__ part, uint32_t g) {
    __shared__ uint32_t wsum[kScanBlock / 64];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t b = 0; b < g; b += kScanBlock) {
        const uint32_t i = b + threadIdx.x;
        const uint32_t x = i < g ? part[i] : 0u;
        const uint32_t inc = block_incl_sum(x, wsum);
        const uint32_t base = carry;
        if (i < g) part[i] = base + inc - x;
        __syncthreads();
        if (threadIdx.x == kScanBlock - 1) carry = base + inc;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_apply(uint32_t *__restrict__ v, uint32_t m,
                                                           const uint32_t *__restrict__ part) {
    __shared__ uint32_t wsum[kScanBlock / 64];
    const uint32_t i = blockIdx.x * kScanBlock + threadIdx.x;
    const uint32_t x = i < m ? v[i] : 0u;
    const uint32_t inc = block_incl_sum(x, wsum);
    if (i < m) v[i] = part[blockIdx.x] + inc - x;
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(
    const uint32_t *__restrict__ keys_in, const uint32_t *__restrict__ vals_in,
    uint32_t *__restrict__ keys_out, uint32_t *__restrict__ vals_out, uint32_t n, uint32_t shift,
    uint32_t n_tiles, const uint32_t *__restrict__ offsets) {
    __shared__ uint32_t run[kRadix];                        // next free place per digit
    __shared__ uint32_t wave_cnt[kSortThreads / 64][kRadix];  // per-wave digit counts
    const uint32_t t = threadIdx.x, w = t >> 6, tile = blockIdx.x;
    run[t] = offsets[t * n_tiles + tile];
    for (uint32_t q = 0; q < kSortThreads / 64; ++q) wave_cnt[q][t] = 0;
    __syncthreads();
    for (uint32_t k = 0; k < kSortPerThread; ++k) {
        const uint32_t i = tile * kSortTile + k * kSortThreads + t;
        const bool valid = i < n;
        const uint32_t key = valid ? keys_in[i] : 0u;
        const uint32_t val = valid ? vals_in[i] : 0u;
        const uint32_t d = (key >> shift) & (kRadix - 1);
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (uint32_t bit = 0; bit < 8; ++bit) {
            const unsigned long long m = __ballot((d >> bit) & 1u);
            peers &= ((d >> bit) & 1u) ? m : ~m;
        }
        const uint32_t below = lane_rank(peers);
        if (valid && below == 0) wave_cnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + below;
            for (uint32_t q = 0; q < w; ++q) pos += wave_cnt[q][d];
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (uint32_t q = 0; q < kSortThreads / 64; ++q) {
            add += wave_cnt[q][t];
            wave_cnt[q][t] = 0;
        }
        run[t] += add;
        __syncthreads();
    }
}

// Temporary space of the sort: the digit histograms, then the scan's block sums.
static size_t sort_hist_words(uint32_t n) {
    return std::max<size_t>(1, (n + kSortTile - 1) / kSortTile) * kRadix;
}

hipError_t sort_temp_bytes(uint32_t n, size_t *bytes) {
    const size_t h = sort_hist_words(n);
    *bytes = (h + (h + kScanBlock - 1) / kScanBlock) * sizeof(uint32_t);
    return hipSuccess;
}

hipError_t launch_sort_perm(const StepParams &p, uint32_t *keys_a, uint32_t *keys_b,
                            uint32_t *vals_a, uint32_t *vals_b, void *temp, size_t temp_bytes,
                            hipStream_t s, const uint32_t **perm) {
    *perm = vals_a;
    if (p.n == 0) return hipSuccess;
    const uint32_t tiles = (p.n + kSortTile - 1) / kSortTile;
    size_t need = 0;
    if (const hipError_t e = sort_temp_bytes(p.n, &need); e != hipSuccess) return e;
    if (temp_bytes < need) return hipErrorInvalidValue;
    const uint32_t m = tiles * kRadix, g = (m + kScanBlock - 1) / kScanBlock;
    uint32_t *hist = static_cast<uint32_t *>(temp);
    uint32_t *part = hist + sort_hist_words(p.n);
    hipLaunchKernelGGL(k_sort_key, dim3((p.n + 255) / 256), dim3(256), 0, s, p, keys_a, vals_a);
    uint32_t *ki = keys_a, *ko = keys_b, *vi = vals_a, *vo = vals_b;
    for (uint32_t shift = 0; shift < 32; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(kSortThreads), 0, s, ki, p.n, shift,
                           tiles, hist);
        hipLaunchKernelGGL(k_scan_reduce, dim3(g), dim3(kScanBlock), 0, s, hist, m, part);
        hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(kScanBlock), 0, s, part, g);
        hipLaunchKernelGGL(k_scan_apply, dim3(g), dim3(kScanBlock), 0, s, hist, m, part);
        hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kSortThreads), 0, s, ki, vi, ko, vo,
                           p.n, shift, tiles, hist);
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    *perm = vi;  // four passes: back in the first buffers
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_gather(uint32_t n, const uint32_t *perm, Fields f) {
    const uint32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const uint32_t src = perm[j];
    for (uint32_t q = 0; q < f.count; ++q) f.dst[q][j] = f.src[q][src];
}

hipError_t launch_gather(uint32_t n, const uint32_t *perm, const Fields &f, hipStream_t s) {
    if (n == 0 || f.count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, s, n, perm, f);
    return hipGetLastError();
}

}  // namespace dsim
