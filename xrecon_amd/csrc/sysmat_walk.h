// Traversals of the system-matrix builder, shared by the host builder (sysmat.cpp) and
// the GPU emission (sysmat_gpu.hip).
//
// The reference builds A's sparse matrix on the CPU by walking its detector-driven FP32
// traversal and dropping coefficients that are 0 or not finite (src/cpu/system_matrix3d.cpp:
// 23-29, 39-62 -> src/cpu/forward3d.cpp:611-658; src/cpu/system_matrix2d.cpp for 2D).  A
// matrix with the same index and sparsity pattern needs the same FP32 decisions -- where
// each walk starts (floor of FP32 quotients), which boundary comes first, which segments
// round to zero -- so these functions evaluate the same FP32 expressions in the same order
// with expression contraction off.  Every value that needs a transcendental function
// (sinf, cosf, tanf, atanf of the float32 angle sums) is computed once on the host into the
// per-view tables below; the traversals themselves use only IEEE +, -, *, /, sqrt, floor
// and comparisons, which the host and the GPU round identically, so both emit the same
// triplets bit for bit and in the same order.
//
// One traversal "item" is what one GPU thread walks:
//   cone 3D:     (view, detector row, 256-column block, volume line)
//   parallel 3D: (view, voxel)
//   2D:          (view, image line)
// and the host builder visits the items in the reference's loop order, so the
// concatenation of the items' triplets in item order is the reference's triplet order.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/xrecon_amd.h"

#pragma clang fp contract(off)

namespace xr {
namespace sm {

// The reference's column block of the cone traversal (src/cpu/forward3d.cpp:13).
constexpr int kBlockCols = 256;

// std::min / std::max semantics (the first argument on ties).
__host__ __device__ inline float smin(float a, float b) { return (b < a) ? b : a; }
__host__ __device__ inline float smax(float a, float b) { return (a < b) ? b : a; }
__host__ __device__ inline int smin(int a, int b) { return (b < a) ? b : a; }
__host__ __device__ inline int smax(int a, int b) { return (a < b) ? b : a; }

// Coefficients kept in the matrix: the reference drops 0 and non-finite values
// (src/cpu/system_matrix3d.cpp:23-29).
__host__ __device__ inline bool keep_value(float v) { return v != 0.0f && __builtin_isfinite(v); }

struct Src {
    float x, y;
};

// include/xrmath.h:8-38 and :60-62, in FP32.
__host__ __device__ inline float ax_x(Src s, float px, float py) { return (s.x * py - s.y * px) / (py - s.y); }
__host__ __device__ inline float ax_y(Src s, float px, float py) { return (s.y * px - s.x * py) / (px - s.x); }
__host__ __device__ inline float tri_cos(float a, float b, float c) { return fabsf(c) / sqrtf(a * a + b * b + c * c); }

// Per-view values (include/xrmath.h:116-123; src/cpu/forward3d.cpp:309-310).
struct SmView {
    float angle, sinv, cosv;
    Src s;
    int xdrv, positive;
    int pad;
};

// Column boundary of a detector column on the virtual flat detector (no transcendental).
__host__ __device__ inline float flat_u(const xr_geometry3d &g, int col, bool positive, float offset)
{
    return positive ? ((col + offset - g.nu * 0.5f) * g.du + g.shift_u) * g.sod / g.sdd
                    : ((g.nu * 0.5f - col - offset) * g.du + g.shift_u) * g.sod / g.sdd;
}

// Depth walk over the slices a row interval covers (src/cpu/forward3d.cpp:203-227).
template <typename Emit>
__host__ __device__ inline void depth_walk(float z_begin, float z_end, float z0, float dzs, int nz, float acoef,
                                           Emit &emit)
{
    float bound = smax(z0, z_begin);
    const float end = smin(z0 + dzs * nz, z_end);
    if (!(bound < end)) return;
    int iz = static_cast<int>(floorf((bound - z0) / dzs));
    iz = smax(0, smin(nz - 1, iz));
    float next = z0 + (iz + 1) * dzs;
    while (bound < end && iz < nz) {
        const float cn = smin(next, end);
        emit(iz, (cn - bound) * acoef);
        bound = cn;
        ++iz;
        next += dzs;
    }
}

// Start value of the arc traversal of one line (src/cpu/forward3d.cpp:398-411) before the
// block offset: delta = atan((t0 - t1) / (1 + t0 t1)).  Host only (atanf).
inline float arc_line_delta(const SmView &v, float shift_u, float p0);

// One (view, detector row, column block, volume line) of the cone traversal
// (src/cpu/forward3d.cpp:299-494): emit(local column, ix, iy, iz, coefficient).
//   cax    the view's column boundaries on the driving axis, Nu + 1 entries
//   center the view's column centres (the coordinate the row heights are scaled at), Nu
//   delta  arc only: arc_line_delta of this line
template <typename Emit>
__host__ __device__ inline void cone_line(const xr_volume3d &vol, const xr_geometry3d &g, const SmView &v,
                                          const float *cax, const float *center, float delta, int det_row, int col0,
                                          int ncol, int L, Emit &emit)
{
    const bool flat = g.type == XR_FLAT;
    const float fs = g.sod / g.sdd;
    const float rv0 = ((det_row - g.nv * 0.5f) * g.dv + g.shift_v) * (flat ? fs : 1.0f);
    const float rv1 = rv0 + g.dv * (flat ? fs : 1.0f);
    const float S = v.xdrv ? v.s.y : v.s.x;
    const float P = v.xdrv ? v.s.x : v.s.y;
    // the reference's per-column arrays r0, r1 and scale (src/cpu/forward3d.cpp:330-344),
    // evaluated for the column the walk is on
    int cur_col = -1;
    float r0 = 0.0f, r1 = 0.0f, sc = 0.0f;
    auto column = [&](int lc) {
        if (lc == cur_col) return;
        cur_col = lc;
        const float c = center[col0 + lc];
        r0 = S / (S - c) * rv0;
        r1 = S / (S - c) * rv1;
        const float area = (cax[col0 + lc + 1] - cax[col0 + lc]) * (r1 - r0);
        const float cw = tri_cos(0.5f * (cax[col0 + lc + 1] + cax[col0 + lc]) - P, 0.5f * (r1 + r0), S);
        sc = (v.xdrv ? vol.dy : vol.dx) / (area * cw);
    };
    const float z0 = -vol.nz * 0.5f * vol.dz + vol.shift_z;
    const float z1 = vol.nz * 0.5f * vol.dz + vol.shift_z;
    const int nalong = v.xdrv ? vol.nx : vol.ny;
    const float e0 = v.xdrv ? -vol.nx * 0.5f * vol.dx + vol.shift_x : -vol.ny * 0.5f * vol.dy + vol.shift_y;
    const float e1 = v.xdrv ? vol.nx * 0.5f * vol.dx + vol.shift_x : vol.ny * 0.5f * vol.dy + vol.shift_y;
    const float det_interval = flat ? (g.du * g.sod / g.sdd) * fabsf(v.xdrv ? v.cosv : v.sinv) : 0.0f;
    float p0, p1, za0, za1;
    if (v.xdrv) {
        const float py = (vol.ny * 0.5f - L - 0.5f) * vol.dy + vol.shift_y;
        p0 = ax_x(v.s, e0, py);
        p1 = ax_x(v.s, e1, py);
        za0 = v.s.y / (v.s.y - py) * z0;
        za1 = v.s.y / (v.s.y - py) * z1;
    } else {
        const float px = (L - vol.nx * 0.5f + 0.5f) * vol.dx + vol.shift_x;
        p0 = ax_y(v.s, px, e0);
        p1 = ax_y(v.s, px, e1);
        za0 = v.s.x / (v.s.x - px) * z0;
        za1 = v.s.x / (v.s.x - px) * z1;
    }
    const float pint = (p1 - p0) / nalong;
    const float pintz = (za1 - za0) / vol.nz;
    int m = 0, ld = 0;
    float bound = smax(p0, cax[col0]);
    if (flat) {
        if (p0 == bound) {
            const float u0 = flat_u(g, col0, v.positive, 0.0f);
            float dcoord, d0;
            if (v.xdrv) {
                d0 = u0 * v.cosv;
                dcoord = v.s.y * bound / ((bound - v.s.x) * v.sinv / v.cosv + v.s.y);
            } else {
                d0 = u0 * v.sinv;
                dcoord = v.s.x * bound / ((bound - v.s.y) * v.cosv / v.sinv + v.s.x);
            }
            ld = static_cast<int>(floorf((dcoord - d0) / det_interval));
        } else {
            m = static_cast<int>(floorf((bound - p0) / pint));
        }
    } else {
        const float flag = v.positive ? 1.0f : -1.0f;
        ld = static_cast<int>(floorf(g.nu * 0.5f - col0 + delta * flag / g.du));
        if (ld < 0) {
            bound = cax[col0];
            m = static_cast<int>(floorf((bound - p0) / pint));
            ld = 0;
        }
    }
    m = smax(0, m);
    float next_pix = p0 + (m + 1) * pint;
    float next_det = (ld < ncol) ? cax[col0 + ld + 1] : 0.0f;
    while (m < nalong && ld < ncol) {
        column(ld);
        const float zb = smax(za0, r0);
        const float ze = smin(za1, r1);
        const float nxt = smin(next_pix, next_det);
        const float acoef = nxt - bound;
        const int ix = v.xdrv ? m : L;
        const int iy = v.xdrv ? L : vol.ny - 1 - m;
        const int cur = ld;
        const float scale = sc;
        auto slice = [&](int iz, float c) { emit(cur, ix, iy, iz, c * scale); };
        depth_walk(zb, ze, za0, pintz, vol.nz, acoef, slice);
        bound = nxt;
        if (next_pix < next_det) {
            ++m;
            next_pix += pint;
        } else {
            ++ld;
            if (ld < ncol) next_det = cax[col0 + ld + 1];
        }
    }
}

inline float arc_line_delta(const SmView &v, float shift_u, float p0)
{
    const float t = std::tan(shift_u + v.angle);
    const float t1v = (std::abs(t) < 1e-12f) ? (t < 0.0f ? -1e-12f : 1e-12f) : t;
    const float t0 = v.xdrv ? ((v.s.x - p0 == 0.0f) ? 1e10f : v.s.y / (v.s.x - p0)) : (v.s.y - p0) / v.s.x;
    const float t1 = -1.0f / t1v;
    return std::atan((t0 - t1) / (1.0f + t0 * t1));
}

// The first point p0 of a cone line on the driving axis (the value cone_line computes),
// for the host's per-line arc table.
inline float cone_line_p0(const xr_volume3d &vol, const SmView &v, int L)
{
    if (v.xdrv) {
        const float e0 = -vol.nx * 0.5f * vol.dx + vol.shift_x;
        const float py = (vol.ny * 0.5f - L - 0.5f) * vol.dy + vol.shift_y;
        return ax_x(v.s, e0, py);
    }
    const float e0 = -vol.ny * 0.5f * vol.dy + vol.shift_y;
    const float px = (L - vol.nx * 0.5f + 0.5f) * vol.dx + vol.shift_x;
    return ax_y(v.s, px, e0);
}

// Parallel-beam 3D coefficients of one voxel (src/cpu/forward3d.cpp:37-109):
// emit(detector row, physical column, coefficient).
template <typename Emit>
__host__ __device__ inline void parallel_voxel(const xr_volume3d &vol, const xr_geometry3d &g, float sinv, float cosv,
                                               int ix, int iy, int iz, Emit &emit)
{
    const bool xdrv = cosv * cosv > 0.5f;
    const bool keep = (!xdrv && sinv >= 0.0f) || (xdrv && cosv >= 0.0f);
    const float i0 = keep ? -0.5f * g.nu : 0.5f * g.nu;
    const float da0 = xdrv ? (i0 * g.du + g.shift_u) / cosv : (i0 * g.du + g.shift_u) / sinv;
    const float dint = xdrv ? g.du / fabsf(cosv) : g.du / fabsf(sinv);
    const float pint = xdrv ? vol.dx : vol.dy;
    const float xys = xdrv ? vol.dy / g.du : vol.dx / g.du;
    float va0;
    if (xdrv) {
        const float px0 = (ix - 0.5f * vol.nx) * vol.dx + vol.shift_x;
        const float py = (0.5f * vol.ny - iy - 0.5f) * vol.dy + vol.shift_y;
        va0 = (sinv / cosv) * py + px0;
    } else {
        const float px = (ix - 0.5f * vol.nx + 0.5f) * vol.dx + vol.shift_x;
        const float py0 = (0.5f * vol.ny - iy - 1.0f) * vol.dy + vol.shift_y;
        va0 = (cosv / sinv) * px + py0;
    }
    const float va1 = va0 + pint;
    float ab = smax(va0, da0);
    const float ae = smin(va1, da0 + g.nu * dint);
    if (!(ab < ae)) return;
    int det = static_cast<int>(floorf((ab - da0) / dint));
    det = smax(0, smin(g.nu - 1, det));
    float dnext = da0 + (det + 1) * dint;
    const float vz0 = (iz - 0.5f * vol.nz) * vol.dz + vol.shift_z;
    const float vz1 = vz0 + vol.dz;
    const float dz0 = -0.5f * g.nv * g.dv + g.shift_v;
    const float zb = smax(vz0, dz0);
    const float ze = smin(vz1, dz0 + g.nv * g.dv);
    if (!(zb < ze)) return;
    int row0 = static_cast<int>(floorf((zb - dz0) / g.dv));
    row0 = smax(0, smin(g.nv - 1, row0));
    while (ab < ae && det >= 0 && det < g.nu) {
        const float an = smin(dnext, ae);
        const float acoef = (an - ab) * xys;
        float rb = zb;
        int row = row0;
        float rnext = dz0 + (row + 1) * g.dv;
        while (rb < ze && row >= 0 && row < g.nv) {
            const float zn = smin(rnext, ze);
            emit(row, keep ? det : (g.nu - 1 - det), acoef * (zn - rb) / g.dv);
            rb = zn;
            ++row;
            rnext += g.dv;
        }
        ab = an;
        ++det;
        dnext += dint;
    }
}

// ---- 2D (src/cpu/system_matrix2d.cpp:87-357)

// Per-view values of the 2D traversal.  tc: the arc start's -1/tan(shift_u + beta)
// (computed on the host).
struct SmView2 {
    float angle, sinv, cosv;
    Src s;
    int xdrv, keep;
    float tc;
};

__host__ __device__ inline float cos2d(Src s, bool xdrv, float a1, float a0)
{
    const float a = (a1 + a0) * 0.5f - (xdrv ? s.x : s.y);
    const float b = xdrv ? s.y : s.x;
    return fabsf(b) / sqrtf(a * a + b * b);
}

// The first point of a 2D fan line on the driving axis (as line2d computes it).
__host__ __device__ inline float line2d_p0(const xr_volume2d &vol, const SmView2 &v, int L)
{
    const float e0 = v.xdrv ? -vol.nx * 0.5f * vol.dx + vol.shift_x : -vol.ny * 0.5f * vol.dy + vol.shift_y;
    if (v.xdrv) {
        const float y = (vol.ny * 0.5f - L - 0.5f) * vol.dy + vol.shift_y;
        return ax_x(v.s, e0, y);
    }
    const float x = (L - vol.nx * 0.5f + 0.5f) * vol.dx + vol.shift_x;
    return ax_y(v.s, x, e0);
}

// The arc start delta of a 2D line (host only: atanf).
inline float arc2d_line_delta(const SmView2 &v, float p0)
{
    const float t0 = v.xdrv ? (((v.s.x - p0) == 0.f) ? 1e10f : (v.s.y / (v.s.x - p0))) : (v.s.y - p0) / v.s.x;
    return std::atan((t0 - v.tc) / (1.0f + t0 * v.tc));
}

// One (view, image line) of the 2D traversal: emit(physical detector, image row, image
// column, coefficient) for each in-range segment, in the reference's order.
//   dax    fan geometries: the view's detector boundaries on the driving axis, Nu + 1
//   delta  arc only: arc2d_line_delta of this line
template <typename Emit>
__host__ __device__ inline void line2d(const xr_volume2d &vol, const xr_geometry2d &g, const SmView2 &v,
                                       const float *dax, float delta, int L, Emit &emit)
{
    const bool xdrv = v.xdrv != 0, keep = v.keep != 0;
    const float sinv = v.sinv, cosv = v.cosv;
    const Src s = v.s;
    const int nalong = xdrv ? vol.nx : vol.ny;
    auto put = [&](int row, int col, int det, float seg, float scale) {
        if (det < 0 || det >= g.nu || row < 0 || row >= vol.ny || col < 0 || col >= vol.nx) return;
        emit(keep ? det : (g.nu - 1 - det), row, col, seg * scale);
    };
    if (g.type == XR_PARALLEL) {
        const float i0 = keep ? -g.nu * 0.5f : g.nu * 0.5f;
        const float d0 = xdrv ? (i0 * g.du + g.shift_u) / cosv : (i0 * g.du + g.shift_u) / sinv;
        const float dint = xdrv ? g.du / fabsf(cosv) : g.du / fabsf(sinv);
        const float pint = xdrv ? vol.dx : vol.dy;
        const float scale = xdrv ? vol.dy / g.du : vol.dx / g.du;
        const float e0 = xdrv ? -vol.nx * 0.5f * vol.dx + vol.shift_x : -vol.ny * 0.5f * vol.dy + vol.shift_y;
        const float p0 = xdrv ? (sinv / cosv) * ((vol.ny * 0.5f - L - 0.5f) * vol.dy + vol.shift_y) + e0
                              : (cosv / sinv) * ((L - vol.nx * 0.5f + 0.5f) * vol.dx + vol.shift_x) + e0;
        float bound = smax(p0, d0);
        int m, det;
        if (p0 == bound) {
            m = 0;
            det = static_cast<int>(floorf((bound - d0) / dint));
        } else {
            m = static_cast<int>(floorf((bound - p0) / pint));
            det = 0;
        }
        float pn = (m + 1) * pint + p0;
        float dn = (det + 1) * dint + d0;
        while (m < nalong && det < g.nu) {
            const int row = xdrv ? L : vol.ny - 1 - m, col = xdrv ? m : L;
            if (pn < dn) {
                put(row, col, det, pn - bound, scale);
                bound = pn;
                ++m;
                pn += pint;
            } else {
                put(row, col, det, dn - bound, scale);
                bound = dn;
                ++det;
                dn += dint;
            }
        }
        return;
    }
    const bool flat = g.type == XR_FLAT;
    const float e0 = xdrv ? -vol.nx * 0.5f * vol.dx + vol.shift_x : -vol.ny * 0.5f * vol.dy + vol.shift_y;
    const float e1 = xdrv ? vol.nx * 0.5f * vol.dx + vol.shift_x : vol.ny * 0.5f * vol.dy + vol.shift_y;
    const float det_step = g.du * fabsf(xdrv ? cosv : sinv);
    const float det0c = ((keep ? -g.nu * 0.5f : g.nu * 0.5f) * g.du + g.shift_u) * (xdrv ? cosv : sinv);
    const int flag = keep ? 1 : -1;
    float p0, p1;
    if (xdrv) {
        const float y = (vol.ny * 0.5f - L - 0.5f) * vol.dy + vol.shift_y;
        p0 = ax_x(s, e0, y);
        p1 = ax_x(s, e1, y);
    } else {
        const float x = (L - vol.nx * 0.5f + 0.5f) * vol.dx + vol.shift_x;
        p0 = ax_y(s, x, e0);
        p1 = ax_y(s, x, e1);
    }
    const float pstep = (p1 - p0) / nalong;
    float bound;
    int m, det;
    if (flat) {
        bound = smax(p0, dax[0]);
        if (p0 == bound) {
            const float b2d = xdrv ? s.y * bound / ((bound - s.x) * sinv / cosv + s.y)
                                   : s.x * bound / ((bound - s.y) * cosv / sinv + s.x);
            m = 0;
            det = static_cast<int>(floorf((b2d - det0c) / det_step));
        } else {
            m = static_cast<int>(floorf((bound - p0) / pstep));
            det = 0;
        }
    } else {
        det = static_cast<int>(floorf(g.nu * 0.5f + delta * flag / g.du));
        if (det < 0) {
            bound = dax[0];
            m = static_cast<int>(floorf((bound - p0) / pstep));
            det = 0;
        } else {
            bound = p0;
            m = 0;
        }
    }
    float pn = (m + 1) * pstep + p0;
    float dn = (det < g.nu) ? dax[det + 1] : 0.f;
    while (m < nalong && det < g.nu) {
        const float scale =
            (xdrv ? vol.dy : vol.dx) / ((dax[det + 1] - dax[det]) * cos2d(s, xdrv, dax[det + 1], dax[det]));
        const int row = xdrv ? L : vol.ny - 1 - m, col = xdrv ? m : L;
        if (pn < dn) {
            put(row, col, det, pn - bound, scale);
            bound = pn;
            ++m;
            pn += pstep;
        } else {
            put(row, col, det, dn - bound, scale);
            bound = dn;
            ++det;
            if (det < g.nu) dn = dax[det + 1];
        }
    }
}

}  // namespace sm
}  // namespace xr
