// Host-side system-matrix builder: COO triplets of the forward projector A.
//
// The reference builds A's sparse matrix on the CPU only (src/cpu/system_matrix3d.cpp:
// 39-62 -> src/cpu/forward3d.cpp:611-658 for 3D; src/cpu/system_matrix2d.cpp for 2D) by
// walking its detector-driven FP32 traversal and dropping coefficients that are 0 or not
// finite (system_matrix3d.cpp:23-29).  The traversals live in sysmat_walk.h, shared with
// the GPU emission (sysmat_gpu.hip); this file computes their per-view tables (every
// transcendental of the traversal: sinf/cosf of the angle, tanf of the arc column angles,
// cosf/sinf of the arc column centres, atanf of the arc line starts) and visits the
// traversal items in the reference's loop order.  Expression contraction into FMAs is
// disabled to keep every rounding where the reference has it.  The GPU operators do not use
// this traversal; they compute the same coefficients from FP64 tables (plan.cpp).
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "sysmat_internal.h"
#include "xr_internal.h"

namespace xr {
namespace sm {

namespace {

inline float ax_x_ang(Src s, float a) { return s.x + s.y * tanf(a); }
inline float ax_y_ang(Src s, float a) { return s.y + s.x / tanf(a); }

SmView make_view(float ang, float sod)
{
    SmView v{};
    v.angle = ang;
    v.sinv = sinf(ang);
    v.cosv = cosf(ang);
    v.s = Src{-v.sinv * sod, v.cosv * sod};
    v.xdrv = v.cosv * v.cosv > 0.5f;
    v.positive = v.xdrv ? (v.cosv >= 0.0f) : (v.sinv >= 0.0f);
    return v;
}

inline float flat_axis(const SmView &v, const xr_geometry3d &g, int col)
{
    const float u = flat_u(g, col, v.positive, 0.0f);
    const float x = u * v.cosv, y = u * v.sinv;
    return v.xdrv ? ax_x(v.s, x, y) : ax_y(v.s, x, y);
}

inline float arc_angle(const SmView &v, const xr_geometry3d &g, int col, float offset)
{
    return v.positive ? (col + offset - g.nu * 0.5f) * g.du + g.shift_u + v.angle
                      : (g.nu * 0.5f - col - offset) * g.du + g.shift_u + v.angle;
}

inline float arc_axis(const SmView &v, const xr_geometry3d &g, int col)
{
    const float a = arc_angle(v, g, col, 0.0f);
    return v.xdrv ? ax_x_ang(v.s, a) : ax_y_ang(v.s, a);
}

}  // namespace

void build_tables3d(const xr_volume3d &vol, const xr_geometry3d &g, const float *ang, int nview, Tables3d &t)
{
    const int nu = g.nu;
    t.nview = nview;
    t.maxline = std::max(vol.nx, vol.ny);
    t.views.assign(nview, SmView{});
    t.cax.assign(kTablePad + (size_t)nview * (nu + 1), 0.0f);
    t.center.assign(kTablePad + (size_t)nview * nu, 0.0f);
    t.delta.assign(g.type == XR_ARC ? (size_t)nview * t.maxline : 0, 0.0f);
    if (g.type == XR_PARALLEL) {
        for (int v = 0; v < nview; ++v) {
            t.views[v].angle = ang[v];
            t.views[v].sinv = std::sin(ang[v]);
            t.views[v].cosv = std::cos(ang[v]);
        }
        return;
    }
    const bool flat = g.type == XR_FLAT;
    for (int v = 0; v < nview; ++v) {
        const SmView vw = make_view(ang[v], g.sod);
        t.views[v] = vw;
        float *cax = t.cax.data() + kTablePad + (size_t)v * (nu + 1);
        float *center = t.center.data() + kTablePad + (size_t)v * nu;
        for (int col = 0; col <= nu; ++col) cax[col] = flat ? flat_axis(vw, g, col) : arc_axis(vw, g, col);
        for (int col = 0; col < nu; ++col) {
            if (flat) {
                const float u = flat_u(g, col, vw.positive, 0.5f);
                center[col] = vw.xdrv ? u * vw.sinv : u * vw.cosv;
            } else {
                const float a = arc_angle(vw, g, col, 0.5f);
                center[col] = vw.xdrv ? -std::cos(a) * g.sdd + vw.s.y : std::sin(a) * g.sdd + vw.s.x;
            }
        }
        if (!flat) {
            const int nlines = vw.xdrv ? vol.ny : vol.nx;
            for (int L = 0; L < nlines; ++L)
                t.delta[(size_t)v * t.maxline + L] = arc_line_delta(vw, g.shift_u, cone_line_p0(vol, vw, L));
        }
    }
}

void build_tables2d(const xr_volume2d &vol, const xr_geometry2d &g, const float *ang, int nview, Tables2d &t)
{
    const int nu = g.nu;
    t.nview = nview;
    t.maxline = std::max(vol.nx, vol.ny);
    t.views.assign(nview, SmView2{});
    t.dax.assign(kTablePad + (size_t)nview * (nu + 1), 0.0f);
    t.delta.assign(g.type == XR_ARC ? (size_t)nview * t.maxline : 0, 0.0f);
    for (int v = 0; v < nview; ++v) {
        const float b = ang[v];
        SmView2 &w = t.views[v];
        w.angle = b;
        w.sinv = std::sin(b);
        w.cosv = std::cos(b);
        w.s = Src{-w.sinv * g.sod, w.cosv * g.sod};
        w.xdrv = w.cosv * w.cosv > 0.5f;
        w.keep = (w.cosv * w.cosv <= 0.5f && w.sinv >= 0.f) || (w.cosv * w.cosv > 0.5f && w.cosv >= 0.f);
        if (g.type == XR_PARALLEL) continue;
        float *dax = t.dax.data() + kTablePad + (size_t)v * (nu + 1);
        for (int d = 0; d <= nu; ++d) {
            const float idx = w.keep ? (d - nu * 0.5f) : (nu * 0.5f - d);
            if (g.type == XR_FLAT) {
                const float px = (idx * g.du + g.shift_u) * w.cosv, py = (idx * g.du + g.shift_u) * w.sinv;
                dax[d] = w.xdrv ? ax_x(w.s, px, py) : ax_y(w.s, px, py);
            } else {
                const float a = idx * g.du + g.shift_u + b;
                dax[d] = w.xdrv ? ax_x_ang(w.s, a) : ax_y_ang(w.s, a);
            }
        }
        if (g.type == XR_ARC) {
            if (w.xdrv) w.tc = (g.shift_u + b == 0.f) ? 1e10f : (-1.0f / std::tan(g.shift_u + b));
            else w.tc = -1.0f / std::tan(g.shift_u + b);
            const int nlines = w.xdrv ? vol.ny : vol.nx;
            for (int L = 0; L < nlines; ++L)
                t.delta[(size_t)v * t.maxline + L] = arc2d_line_delta(w, line2d_p0(vol, w, L));
        }
    }
}

}  // namespace sm
}  // namespace xr

namespace {

using xr::sm::keep_value;

struct Coo {
    std::vector<int64_t> rows, cols;
    std::vector<float> vals;
    void add(int64_t r, int64_t c, float v)
    {
        if (!keep_value(v)) return;
        rows.push_back(r);
        cols.push_back(c);
        vals.push_back(v);
    }
};

// The reference's loop order: views, then (parallel) voxels z-y-x, or (cone) detector
// rows, 256-column blocks and volume lines.
void build3d(const xr_volume3d &vol, const xr_geometry3d &g, const float *ang, int nview, Coo &coo)
{
    using namespace xr::sm;
    Tables3d t;
    build_tables3d(vol, g, ang, nview, t);
    if (g.type == XR_PARALLEL) {
        for (int v = 0; v < nview; ++v)
            for (int iz = 0; iz < vol.nz; ++iz)
                for (int iy = 0; iy < vol.ny; ++iy)
                    for (int ix = 0; ix < vol.nx; ++ix) {
                        const int64_t icol = ((int64_t)iz * vol.ny + iy) * vol.nx + ix;
                        auto emit = [&](int row, int col, float c) {
                            coo.add(((int64_t)v * g.nv + row) * g.nu + col, icol, c);
                        };
                        parallel_voxel(vol, g, t.views[v].sinv, t.views[v].cosv, ix, iy, iz, emit);
                    }
        return;
    }
    const int nblk = (g.nu - 1) / kBlockCols + 1;
    for (int v = 0; v < nview; ++v) {
        const SmView &vw = t.views[v];
        const float *cax = t.cax.data() + kTablePad + (size_t)v * (g.nu + 1);
        const float *center = t.center.data() + kTablePad + (size_t)v * g.nu;
        const int nlines = vw.xdrv ? vol.ny : vol.nx;
        for (int r = 0; r < g.nv; ++r)
            for (int bk = 0; bk < nblk; ++bk) {
                const int c0 = bk * kBlockCols;
                const int nc = std::min(kBlockCols, g.nu - c0);
                auto emit = [&](int lc, int ix, int iy, int iz, float c) {
                    const int col = c0 + lc;
                    const int out = vw.positive ? col : (g.nu - 1 - col);
                    coo.add(((int64_t)v * g.nv + r) * g.nu + out, ((int64_t)iz * vol.ny + iy) * vol.nx + ix, c);
                };
                for (int L = 0; L < nlines; ++L) {
                    const float delta = g.type == XR_ARC ? t.delta[(size_t)v * t.maxline + L] : 0.0f;
                    cone_line(vol, g, vw, cax, center, delta, r, c0, nc, L, emit);
                }
            }
    }
}

void build2d(const xr_volume2d &vol, const xr_geometry2d &g, const float *ang, int nview, Coo &coo)
{
    using namespace xr::sm;
    Tables2d t;
    build_tables2d(vol, g, ang, nview, t);
    for (int v = 0; v < nview; ++v) {
        const SmView2 &w = t.views[v];
        const float *dax = t.dax.data() + kTablePad + (size_t)v * (g.nu + 1);
        const int nlines = w.xdrv ? vol.ny : vol.nx;
        auto emit = [&](int det, int row, int col, float c) {
            coo.add((int64_t)v * g.nu + det, (int64_t)row * vol.nx + col, c);
        };
        for (int L = 0; L < nlines; ++L) {
            const float delta = g.type == XR_ARC ? t.delta[(size_t)v * t.maxline + L] : 0.0f;
            line2d(vol, g, w, dax, delta, L, emit);
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
