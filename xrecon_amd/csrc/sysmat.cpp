// Host-side system-matrix builder: COO triplets of the forward projector A.
//
// The reference builds A's sparse matrix on the CPU only (src/cpu/system_matrix3d.cpp:
// 39-62 -> src/cpu/forward3d.cpp:611-658 for 3D; src/cpu/system_matrix2d.cpp for 2D) by
// walking its detector-driven FP32 traversal and dropping coefficients that are 0 or not
// finite (system_matrix3d.cpp:23-29).  A matrix with the same index and sparsity pattern
// needs the same FP32 decisions -- where each walk starts (floor of FP32 quotients),
// which boundary comes first, which segments round to zero -- so this file evaluates the
// same FP32 expressions in the same order.  Expression contraction into FMAs is disabled
// to keep every rounding where the reference has it.  The GPU operators do not use this
// traversal; they compute the same coefficients from FP64 tables (plan.cpp).
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "xr_internal.h"

namespace {

struct Coo {
    std::vector<int64_t> rows, cols;
    std::vector<float> vals;
    void add(int64_t r, int64_t c, float v)
    {
        if (v == 0.0f || !std::isfinite(v)) return;
        rows.push_back(r);
        cols.push_back(c);
        vals.push_back(v);
    }
};

// ---- FP32 geometry helpers (include/xrmath.h:8-62, src/cpu/forward3d.cpp:196-297)

struct Src {
    float x, y;
};

inline float ax_x(Src s, float px, float py) { return (s.x * py - s.y * px) / (py - s.y); }
inline float ax_y(Src s, float px, float py) { return (s.y * px - s.x * py) / (px - s.x); }
inline float ax_x_ang(Src s, float a) { return s.x + s.y * tanf(a); }
inline float ax_y_ang(Src s, float a) { return s.y + s.x / tanf(a); }
inline float tri_cos(float a, float b, float c) { return fabsf(c) / sqrtf(a * a + b * b + c * c); }

inline float tan_away_from_zero(float v)
{
    const float t = std::tan(v);
    if (std::abs(t) < 1e-12f) return t < 0.0f ? -1e-12f : 1e-12f;
    return t;
}

struct View {
    float angle, sinv, cosv;
    Src s;
    bool xdrv, positive;
};

inline View make_view(float ang, float sod)
{
    View v;
    v.angle = ang;
    v.sinv = sinf(ang);
    v.cosv = cosf(ang);
    v.s = Src{-v.sinv * sod, v.cosv * sod};
    v.xdrv = v.cosv * v.cosv > 0.5f;
    v.positive = v.xdrv ? (v.cosv >= 0.0f) : (v.sinv >= 0.0f);
    return v;
}

// Column boundary of (block-relative) detector column on the driving axis, flat detector.
inline float flat_u(const xr_geometry3d &g, int col, bool positive, float offset)
{
    return positive ? ((col + offset - g.nu * 0.5f) * g.du + g.shift_u) * g.sod / g.sdd
                    : ((g.nu * 0.5f - col - offset) * g.du + g.shift_u) * g.sod / g.sdd;
}

inline float flat_axis(const View &v, const xr_geometry3d &g, int col)
{
    const float u = flat_u(g, col, v.positive, 0.0f);
    const float x = u * v.cosv, y = u * v.sinv;
    return v.xdrv ? ax_x(v.s, x, y) : ax_y(v.s, x, y);
}

inline float arc_angle(const View &v, const xr_geometry3d &g, int col, float offset)
{
    return v.positive ? (col + offset - g.nu * 0.5f) * g.du + g.shift_u + v.angle
                      : (g.nu * 0.5f - col - offset) * g.du + g.shift_u + v.angle;
}

inline float arc_axis(const View &v, const xr_geometry3d &g, int col)
{
    const float a = arc_angle(v, g, col, 0.0f);
    return v.xdrv ? ax_x_ang(v.s, a) : ax_y_ang(v.s, a);
}

// Depth walk over the slices a row interval covers (src/cpu/forward3d.cpp:203-227).
template <typename Emit>
void depth_walk(float z_begin, float z_end, float z0, float dzs, int nz, float acoef, Emit &&emit)
{
    float bound = std::max(z0, z_begin);
    const float end = std::min(z0 + dzs * nz, z_end);
    if (!(bound < end)) return;
    int iz = static_cast<int>(std::floor((bound - z0) / dzs));
    iz = std::max(0, std::min(nz - 1, iz));
    float next = z0 + (iz + 1) * dzs;
    while (bound < end && iz < nz) {
        const float cn = std::min(next, end);
        emit(iz, (cn - bound) * acoef);
        bound = cn;
        ++iz;
        next += dzs;
    }
}

// One (view, detector row, 256-column block) of the cone traversal
// (src/cpu/forward3d.cpp:299-494): emit(local_col, ix, iy, iz, coefficient * scale).
template <typename Emit>
void cone_block(const xr_volume3d &vol, const xr_geometry3d &g, const View &v, int det_row, int col0, int ncol,
                Emit &&emit)
{
    const bool flat = g.type == XR_FLAT;
    std::vector<float> cax(ncol + 1), r0(ncol), r1(ncol), sc(ncol);
    const float fs = g.sod / g.sdd;
    const float rv0 = ((det_row - g.nv * 0.5f) * g.dv + g.shift_v) * (flat ? fs : 1.0f);
    const float rv1 = rv0 + g.dv * (flat ? fs : 1.0f);
    for (int lc = 0; lc < ncol; ++lc) {
        const int col = col0 + lc;
        if (lc == 0) cax[0] = flat ? flat_axis(v, g, col0) : arc_axis(v, g, col0);
        cax[lc + 1] = flat ? flat_axis(v, g, col + 1) : arc_axis(v, g, col + 1);
        float center;
        if (flat) {
            const float u = flat_u(g, col, v.positive, 0.5f);
            center = v.xdrv ? u * v.sinv : u * v.cosv;
        } else {
            const float a = arc_angle(v, g, col, 0.5f);
            center = v.xdrv ? -std::cos(a) * g.sdd + v.s.y : std::sin(a) * g.sdd + v.s.x;
        }
        const float S = v.xdrv ? v.s.y : v.s.x;
        const float P = v.xdrv ? v.s.x : v.s.y;
        r0[lc] = S / (S - center) * rv0;
        r1[lc] = S / (S - center) * rv1;
        const float area = (cax[lc + 1] - cax[lc]) * (r1[lc] - r0[lc]);
        const float cw = tri_cos(0.5f * (cax[lc + 1] + cax[lc]) - P, 0.5f * (r1[lc] + r0[lc]), S);
        sc[lc] = (v.xdrv ? vol.dy : vol.dx) / (area * cw);
    }
    const float z0 = -vol.nz * 0.5f * vol.dz + vol.shift_z;
    const float z1 = vol.nz * 0.5f * vol.dz + vol.shift_z;
    const int nlines = v.xdrv ? vol.ny : vol.nx;
    const int nalong = v.xdrv ? vol.nx : vol.ny;
    const float e0 = v.xdrv ? -vol.nx * 0.5f * vol.dx + vol.shift_x : -vol.ny * 0.5f * vol.dy + vol.shift_y;
    const float e1 = v.xdrv ? vol.nx * 0.5f * vol.dx + vol.shift_x : vol.ny * 0.5f * vol.dy + vol.shift_y;
    const float det_interval = flat ? (g.du * g.sod / g.sdd) * std::abs(v.xdrv ? v.cosv : v.sinv) : 0.0f;
    for (int L = 0; L < nlines; ++L) {
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
        float bound = std::max(p0, cax[0]);
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
                ld = static_cast<int>(std::floor((dcoord - d0) / det_interval));
            } else {
                m = static_cast<int>(std::floor((bound - p0) / pint));
            }
        } else {
            const float t0 = v.xdrv ? ((v.s.x - p0 == 0.0f) ? 1e10f : v.s.y / (v.s.x - p0)) : (v.s.y - p0) / v.s.x;
            const float t1 = -1.0f / tan_away_from_zero(g.shift_u + v.angle);
            const float delta = std::atan((t0 - t1) / (1.0f + t0 * t1));
            const float flag = v.positive ? 1.0f : -1.0f;
            ld = static_cast<int>(std::floor(g.nu * 0.5f - col0 + delta * flag / g.du));
            if (ld < 0) {
                bound = cax[0];
                m = static_cast<int>(std::floor((bound - p0) / pint));
                ld = 0;
            }
        }
        m = std::max(0, m);
        float next_pix = p0 + (m + 1) * pint;
        float next_det = (ld < ncol) ? cax[ld + 1] : 0.0f;
        while (m < nalong && ld < ncol) {
            const float zb = std::max(za0, r0[ld]);
            const float ze = std::min(za1, r1[ld]);
            const float nxt = std::min(next_pix, next_det);
            const float acoef = nxt - bound;
            const int ix = v.xdrv ? m : L;
            const int iy = v.xdrv ? L : vol.ny - 1 - m;
            const int cur = ld;
            depth_walk(zb, ze, za0, pintz, vol.nz, acoef,
                       [&](int iz, float c) { emit(cur, ix, iy, iz, c * sc[cur]); });
            bound = nxt;
            if (next_pix < next_det) {
                ++m;
                next_pix += pint;
            } else {
                ++ld;
                if (ld < ncol) next_det = cax[ld + 1];
            }
        }
    }
}

// Parallel-beam 3D coefficients of one voxel (src/cpu/forward3d.cpp:37-109).
template <typename Emit>
void parallel_voxel(const xr_volume3d &vol, const xr_geometry3d &g, float ang, int ix, int iy, int iz, Emit &&emit)
{
    const float sinv = std::sin(ang), cosv = std::cos(ang);
    const bool xdrv = cosv * cosv > 0.5f;
    const bool keep = (!xdrv && sinv >= 0.0f) || (xdrv && cosv >= 0.0f);
    const float i0 = keep ? -0.5f * g.nu : 0.5f * g.nu;
    const float da0 = xdrv ? (i0 * g.du + g.shift_u) / cosv : (i0 * g.du + g.shift_u) / sinv;
    const float dint = xdrv ? g.du / std::abs(cosv) : g.du / std::abs(sinv);
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
    float ab = std::max(va0, da0);
    const float ae = std::min(va1, da0 + g.nu * dint);
    if (!(ab < ae)) return;
    int det = static_cast<int>(std::floor((ab - da0) / dint));
    det = std::max(0, std::min(g.nu - 1, det));
    float dnext = da0 + (det + 1) * dint;
    const float vz0 = (iz - 0.5f * vol.nz) * vol.dz + vol.shift_z;
    const float vz1 = vz0 + vol.dz;
    const float dz0 = -0.5f * g.nv * g.dv + g.shift_v;
    const float zb = std::max(vz0, dz0);
    const float ze = std::min(vz1, dz0 + g.nv * g.dv);
    if (!(zb < ze)) return;
    int row0 = static_cast<int>(std::floor((zb - dz0) / g.dv));
    row0 = std::max(0, std::min(g.nv - 1, row0));
    while (ab < ae && det >= 0 && det < g.nu) {
        const float an = std::min(dnext, ae);
        const float acoef = (an - ab) * xys;
        float rb = zb;
        int row = row0;
        float rnext = dz0 + (row + 1) * g.dv;
        while (rb < ze && row >= 0 && row < g.nv) {
            const float zn = std::min(rnext, ze);
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

void build3d(const xr_volume3d &vol, const xr_geometry3d &g, const float *ang, int nview, Coo &coo)
{
    if (g.type == XR_PARALLEL) {
        for (int v = 0; v < nview; ++v)
            for (int iz = 0; iz < vol.nz; ++iz)
                for (int iy = 0; iy < vol.ny; ++iy)
                    for (int ix = 0; ix < vol.nx; ++ix) {
                        const int64_t icol = ((int64_t)iz * vol.ny + iy) * vol.nx + ix;
                        parallel_voxel(vol, g, ang[v], ix, iy, iz, [&](int row, int col, float c) {
                            coo.add(((int64_t)v * g.nv + row) * g.nu + col, icol, c);
                        });
                    }
        return;
    }
    const int kBlockCols = 256;  // the reference's column block (src/cpu/forward3d.cpp:13)
    const int nblk = (g.nu - 1) / kBlockCols + 1;
    for (int v = 0; v < nview; ++v) {
        const View vw = make_view(ang[v], g.sod);
        for (int r = 0; r < g.nv; ++r)
            for (int bk = 0; bk < nblk; ++bk) {
                const int c0 = bk * kBlockCols;
                const int nc = std::min(kBlockCols, g.nu - c0);
                cone_block(vol, g, vw, r, c0, nc, [&](int lc, int ix, int iy, int iz, float c) {
                    const int col = c0 + lc;
                    const int out = vw.positive ? col : (g.nu - 1 - col);
                    coo.add(((int64_t)v * g.nv + r) * g.nu + out, ((int64_t)iz * vol.ny + iy) * vol.nx + ix, c);
                });
            }
    }
}

// ---- 2D (src/cpu/system_matrix2d.cpp:87-357)

inline float cos2d(Src s, bool xdrv, float a1, float a0)
{
    const float a = (a1 + a0) * 0.5f - (xdrv ? s.x : s.y);
    const float b = xdrv ? s.y : s.x;
    return std::abs(b) / std::sqrt(a * a + b * b);
}

void build2d(const xr_volume2d &vol, const xr_geometry2d &g, const float *ang, int nview, Coo &coo)
{
    auto put = [&](int v, int row, int col, int det, bool keep, float seg, float scale) {
        if (det < 0 || det >= g.nu || row < 0 || row >= vol.ny || col < 0 || col >= vol.nx) return;
        const int dout = keep ? det : (g.nu - 1 - det);
        coo.add((int64_t)v * g.nu + dout, (int64_t)row * vol.nx + col, seg * scale);
    };
    std::vector<float> dax(g.nu + 1);
    for (int v = 0; v < nview; ++v) {
        const float b = ang[v];
        const float sinv = std::sin(b), cosv = std::cos(b);
        const Src s{-sinv * g.sod, cosv * g.sod};
        const bool xdrv = cosv * cosv > 0.5f;
        const bool keep = (cosv * cosv <= 0.5f && sinv >= 0.f) || (cosv * cosv > 0.5f && cosv >= 0.f);
        const int nlines = xdrv ? vol.ny : vol.nx;
        const int nalong = xdrv ? vol.nx : vol.ny;
        if (g.type == XR_PARALLEL) {
            const float i0 = keep ? -g.nu * 0.5f : g.nu * 0.5f;
            const float d0 = xdrv ? (i0 * g.du + g.shift_u) / cosv : (i0 * g.du + g.shift_u) / sinv;
            const float dint = xdrv ? g.du / std::abs(cosv) : g.du / std::abs(sinv);
            const float pint = xdrv ? vol.dx : vol.dy;
            const float scale = xdrv ? vol.dy / g.du : vol.dx / g.du;
            const float e0 = xdrv ? -vol.nx * 0.5f * vol.dx + vol.shift_x : -vol.ny * 0.5f * vol.dy + vol.shift_y;
            for (int L = 0; L < nlines; ++L) {
                const float p0 = xdrv ? (sinv / cosv) * ((vol.ny * 0.5f - L - 0.5f) * vol.dy + vol.shift_y) + e0
                                      : (cosv / sinv) * ((L - vol.nx * 0.5f + 0.5f) * vol.dx + vol.shift_x) + e0;
                float bound = std::max(p0, d0);
                int m, det;
                if (p0 == bound) {
                    m = 0;
                    det = static_cast<int>(std::floor((bound - d0) / dint));
                } else {
                    m = static_cast<int>(std::floor((bound - p0) / pint));
                    det = 0;
                }
                float pn = (m + 1) * pint + p0;
                float dn = (det + 1) * dint + d0;
                while (m < nalong && det < g.nu) {
                    const int row = xdrv ? L : vol.ny - 1 - m, col = xdrv ? m : L;
                    if (pn < dn) {
                        put(v, row, col, det, keep, pn - bound, scale);
                        bound = pn;
                        ++m;
                        pn += pint;
                    } else {
                        put(v, row, col, det, keep, dn - bound, scale);
                        bound = dn;
                        ++det;
                        dn += dint;
                    }
                }
            }
            continue;
        }
        const bool flat = g.type == XR_FLAT;
        for (int d = 0; d <= g.nu; ++d) {
            const float idx = keep ? (d - g.nu * 0.5f) : (g.nu * 0.5f - d);
            if (flat) {
                const float px = (idx * g.du + g.shift_u) * cosv, py = (idx * g.du + g.shift_u) * sinv;
                dax[d] = xdrv ? ax_x(s, px, py) : ax_y(s, px, py);
            } else {
                const float a = idx * g.du + g.shift_u + b;
                dax[d] = xdrv ? ax_x_ang(s, a) : ax_y_ang(s, a);
            }
        }
        const float e0 = xdrv ? -vol.nx * 0.5f * vol.dx + vol.shift_x : -vol.ny * 0.5f * vol.dy + vol.shift_y;
        const float e1 = xdrv ? vol.nx * 0.5f * vol.dx + vol.shift_x : vol.ny * 0.5f * vol.dy + vol.shift_y;
        const float det_step = g.du * std::abs(xdrv ? cosv : sinv);
        const float det0c = ((keep ? -g.nu * 0.5f : g.nu * 0.5f) * g.du + g.shift_u) * (xdrv ? cosv : sinv);
        const int flag = keep ? 1 : -1;
        for (int L = 0; L < nlines; ++L) {
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
                bound = std::max(p0, dax[0]);
                if (p0 == bound) {
                    const float b2d = xdrv ? s.y * bound / ((bound - s.x) * sinv / cosv + s.y)
                                           : s.x * bound / ((bound - s.y) * cosv / sinv + s.x);
                    m = 0;
                    det = static_cast<int>(std::floor((b2d - det0c) / det_step));
                } else {
                    m = static_cast<int>(std::floor((bound - p0) / pstep));
                    det = 0;
                }
            } else {
                float t0, tc;
                if (xdrv) {
                    t0 = ((s.x - p0) == 0.f) ? 1e10f : (s.y / (s.x - p0));
                    tc = (g.shift_u + b == 0.f) ? 1e10f : (-1.0f / std::tan(g.shift_u + b));
                } else {
                    t0 = (s.y - p0) / s.x;
                    tc = -1.0f / std::tan(g.shift_u + b);
                }
                const float delta = std::atan((t0 - tc) / (1.0f + t0 * tc));
                det = static_cast<int>(std::floor(g.nu * 0.5f + delta * flag / g.du));
                if (det < 0) {
                    bound = dax[0];
                    m = static_cast<int>(std::floor((bound - p0) / pstep));
                    det = 0;
                } else {
                    bound = p0;
                    m = 0;
                }
            }
            float pn = (m + 1) * pstep + p0;
            float dn = (det < g.nu) ? dax[det + 1] : 0.f;
            while (m < nalong && det < g.nu) {
                const float scale = (xdrv ? vol.dy : vol.dx) / ((dax[det + 1] - dax[det]) * cos2d(s, xdrv, dax[det + 1], dax[det]));
                const int row = xdrv ? L : vol.ny - 1 - m, col = xdrv ? m : L;
                if (pn < dn) {
                    put(v, row, col, det, keep, pn - bound, scale);
                    bound = pn;
                    ++m;
                    pn += pstep;
                } else {
                    put(v, row, col, det, keep, dn - bound, scale);
                    bound = dn;
                    ++det;
                    if (det < g.nu) dn = dax[det + 1];
                }
            }
        }
    }
}

}  // namespace

struct xr_coo {
    Coo c;
};

using xr::fail;

extern "C" {

int xr_system_matrix3d(const xr_volume3d *vol, const xr_geometry3d *geo, const float *angles_host, int num_view,
                       xr_coo **out)
{
    xr::set_error("");
    if (!vol || !geo || !out || (num_view > 0 && !angles_host)) return fail(XR_EINVAL, "null argument");
    if (vol->nx <= 0 || vol->ny <= 0 || vol->nz <= 0 || geo->nu <= 0 || geo->nv <= 0 || num_view < 0)
        return fail(XR_EINVAL, "sizes must be positive");
    if (geo->type != XR_PARALLEL && geo->type != XR_FLAT && geo->type != XR_ARC)
        return fail(XR_EINVAL, "scan type must be parallel (0), flat (1) or arc (2)");
    // The reference emits int32 row and column indices (src/cpu/forward3d.cpp:651-652).
    if ((int64_t)num_view * geo->nv * geo->nu >= ((int64_t)1 << 31) ||
        (int64_t)vol->nx * vol->ny * vol->nz >= ((int64_t)1 << 31))
        return fail(XR_EINVAL, "system matrix too large for 32-bit indices");
    std::unique_ptr<xr_coo> c(new xr_coo());
    build3d(*vol, *geo, angles_host, num_view, c->c);
    *out = c.release();
    return XR_OK;
}

int xr_system_matrix2d(const xr_volume2d *vol, const xr_geometry2d *geo, const float *angles_host, int num_view,
                       xr_coo **out)
{
    xr::set_error("");
    if (!vol || !geo || !out || (num_view > 0 && !angles_host)) return fail(XR_EINVAL, "null argument");
    if (vol->nx <= 0 || vol->ny <= 0 || geo->nu <= 0 || num_view < 0) return fail(XR_EINVAL, "sizes must be positive");
    if (geo->type != XR_PARALLEL && geo->type != XR_FLAT && geo->type != XR_ARC)
        return fail(XR_EINVAL, "scan type must be parallel (0), flat (1) or arc (2)");
    std::unique_ptr<xr_coo> c(new xr_coo());
    build2d(*vol, *geo, angles_host, num_view, c->c);
    *out = c.release();
    return XR_OK;
}

int64_t xr_coo_nnz(const xr_coo *coo) { return coo ? (int64_t)coo->c.vals.size() : XR_EINVAL; }

int xr_coo_copy(const xr_coo *coo, int64_t *rows, int64_t *cols, float *vals)
{
    if (!coo) return fail(XR_EINVAL, "null triplets");
    const size_t n = coo->c.vals.size();
    if (n && (!rows || !cols || !vals)) return fail(XR_EINVAL, "null output buffer");
    if (n) {
        std::memcpy(rows, coo->c.rows.data(), n * sizeof(int64_t));
        std::memcpy(cols, coo->c.cols.data(), n * sizeof(int64_t));
        std::memcpy(vals, coo->c.vals.data(), n * sizeof(float));
    }
    return XR_OK;
}

void xr_coo_destroy(xr_coo *coo) { delete coo; }

}  // extern "C"
