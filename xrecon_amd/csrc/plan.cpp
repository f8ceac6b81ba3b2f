// Host-side plan builder: per-view geometry tables computed in FP64, stored as FP32.
//
// The reference recomputes this geometry per voxel (or per detector row block) per view
// in FP32 on the device (include/xrmath.h:116-123, src/kernels/backward3d_flat.cu:30-61).
// Here it is computed once per (geometry, angle set), in double precision, and the
// kernels only read it.  The discrete decisions that select the operator (x-driven vs
// y-driven and the column order) are made exactly as the reference makes them: FP32
// sinf/cosf of the float32 angle, squared in FP32, compared with 0.5f
// (include/cpu3d_common.h:65, :87; src/cpu/forward3d.cpp:309-310).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "dd_body.h"

namespace xr {

namespace {
thread_local std::string g_error;

inline double axis_map(bool xdrv, double sx, double sy, double px, double py)
{
    // include/xrmath.h:8-22 (map2x / map2y), in double precision.
    return xdrv ? (sx * py - sy * px) / (py - sy) : (sy * px - sx * py) / (px - sx);
}
}  // namespace

void set_error(const std::string &msg) { g_error = msg; }
const char *last_error() { return g_error.c_str(); }

int fail(int code, const std::string &msg)
{
    set_error(msg);
    return code;
}

int build_tables(Plan3d &p, const float *angles, PlanTables &t, std::string &err)
{
    const xr_volume3d &vol = p.vol;
    const xr_geometry3d &geo = p.geo;
    const int nx = vol.nx, ny = vol.ny, nz = vol.nz, nu = geo.nu, nv = geo.nv;
    const int V = p.nview;
    const double dx = vol.dx, dy = vol.dy, dz = vol.dz;
    const double shx = vol.shift_x, shy = vol.shift_y, shz = vol.shift_z;
    const double du = geo.du, dv = geo.dv, shu = geo.shift_u, shv = geo.shift_v;
    const double sod = geo.sod, sdd = geo.sdd;
    const bool cone = p.kind != GEO_PAR;
    const bool is2d = p.is2d != 0;

    // Virtual scale of the flat detector (include/xrmath.h:175-183): 3D flat detectors are
    // given at their physical pitch and scaled to the iso plane; 2D pitches arrive
    // already effective (xrecon/projector/geometry.py:106-109).
    const double mag = (p.kind == GEO_FLAT && !is2d) ? sod / sdd : 1.0;
    const double rowscale = (p.kind == GEO_FLAT && !is2d) ? sod / sdd : 1.0;

    // Row boundaries v_r = v0 + r * dvv on the iso plane (src/cpu/forward3d.cpp:318-320).
    // A 2D operator is a one-slice volume seen by a one-row detector whose row interval
    // is centred and exactly covered: v = -1/2, +1/2 with sigma = zeta = 1, so that every
    // depth overlap is 1 and the ray cosine sees hbar = 0 (src/cpu/system_matrix2d.cpp:
    // 58-68 cos_weight_x has no height term).
    double v0, dvv, zb, dzz;
    if (is2d) {
        v0 = -0.5; dvv = 1.0; zb = -0.5; dzz = 1.0;
    } else {
        v0 = ((0 - nv * 0.5) * dv + shv) * rowscale;
        dvv = dv * rowscale;
        zb = -nz * 0.5 * dz + shz;
        dzz = dz;
    }
    p.v0 = (float)v0;
    p.dvv = (float)dvv;
    p.rv0 = (float)(v0 / dvv);
    p.zb = (float)zb;
    p.zc = (float)(zb + 0.5 * (is2d ? 1 : nz) * dzz);
    p.dz = (float)dzz;
    p.zoff = (float)(-zb / dzz);
    p.sod = (float)sod;

    // Backprojection column width w (include/xrmath.h:188-198; 2D fan-flat uses du,
    // src/cpu/backward2d.cpp:86; 2D arc 2*SOD*tan(du/2), :109; 2D parallel du, :187).
    double colw;
    if (p.kind == GEO_FLAT) colw = du * mag;
    else if (p.kind == GEO_ARC) colw = 2.0 * sod * std::tan(0.5 * du);
    else colw = du;

    p.maxline = nx > ny ? nx : ny;
    const int ML = p.maxline;
    t.hdr.assign(V, ViewHdr{});
    t.lineC.assign((size_t)V * ML, float4{0, 0, 0, 0});
    t.lineV.assign((size_t)V * ML, float4{0, 0, 0, 0});
    t.col.assign((size_t)V * (nu + 1), float4{0, 0, 0, 0});
    p.xdrv.assign(V, 0);
    p.keep.assign(V, 0);
    p.any_x = p.any_y = 0;
    p.big = ((int64_t)(nz + 1) * nx * ny * 8 >= ((int64_t)1 << 32) || (int64_t)nx * ny * 8 >= ((int64_t)1 << 24)) ? 1 : 0;

    const int ncg = ceil_div(nu, kWave);
    std::vector<int> gspan((size_t)V * ncg, 0);
    std::vector<double> gwmax((size_t)V * ncg, 0.0);
    std::vector<double> d(nu + 1), sig(nu);
    for (int v = 0; v < V; ++v) {
        const float af = angles[v];
        if (!std::isfinite(af)) {
            err = "angle " + std::to_string(v) + " is not finite";
            return XR_EINVAL;
        }
        const float sf = std::sin(af), cf = std::cos(af);  // float overloads: sinf/cosf
        const bool xdrv = cf * cf > 0.5f;
        const bool keep = xdrv ? (cf >= 0.0f) : (sf >= 0.0f);
        p.xdrv[v] = xdrv;
        p.keep[v] = keep;
        (xdrv ? p.any_x : p.any_y) = 1;
        const double a = (double)af;
        const double sb = std::sin(a), cb = std::cos(a);
        const double sx = -sod * sb, sy = sod * cb;
        const double sk = keep ? 1.0 : -1.0;

        ViewHdr &h = t.hdr[v];
        h.xdrv = xdrv;
        h.keep = keep;
        h.nline = xdrv ? ny : nx;
        h.nalong = xdrv ? nx : ny;
        h.delta = (float)(xdrv ? dy : dx);
        h.colw = (float)colw;
        h.sx = (float)sx;
        h.sy = (float)sy;
        h.P = (float)(xdrv ? sx : sy);
        h.S = (float)(xdrv ? sy : sx);

        // Lines: pixel edges on the axis are P0 + m * Delta along a line, voxel heights
        // on the axis plane scale by zeta (include/cpu3d_common.h:68-78).
        for (int L = 0; L < h.nline; ++L) {
            double P0, Dl, zeta;
            if (xdrv) {
                const double py = (ny * 0.5 - L - 0.5) * dy + shy;
                const double xl = -nx * 0.5 * dx + shx;
                if (cone) {
                    zeta = sy / (sy - py);
                    P0 = axis_map(true, sx, sy, xl, py);
                    Dl = dx * zeta;
                } else {
                    // src/cpu/forward3d.cpp:61-64: axis = tan(beta) * py + x
                    zeta = 1.0;
                    P0 = (sb / cb) * py + xl;
                    Dl = dx;
                }
            } else {
                const double px = (L - nx * 0.5 + 0.5) * dx + shx;
                const double yb = -ny * 0.5 * dy + shy;
                if (cone) {
                    zeta = sx / (sx - px);
                    P0 = axis_map(false, sx, sy, px, yb);
                    Dl = dy * zeta;
                } else {
                    // src/cpu/forward3d.cpp:65-68: axis = cot(beta) * px + y
                    zeta = 1.0;
                    P0 = (cb / sb) * px + yb;
                    Dl = dy;
                }
            }
            if (is2d) zeta = 1.0;  // the one-slice embedding has no depth magnification
            if (!(Dl > 0.0) || !(zeta > 0.0)) {
                err = "view " + std::to_string(v) + ": source inside the volume or degenerate pixel footprint";
                return XR_EINVAL;
            }
            const size_t li = (size_t)v * ML + L;
            t.lineC[li] = float4{(float)(-P0 / Dl), (float)(1.0 / Dl), (float)(1.0 / (zeta * dzz)),
                                 (float)(Dl * zeta * dzz)};
            t.lineV[li] = float4{(float)P0, (float)Dl, (float)zeta, (float)(1.0 / (Dl * zeta))};
        }

        // Columns: boundaries on the axis in internal order; per-column depth scale sigma
        // from the column centre (include/cpu3d_common.h:94-131 flat, :150-182 arc;
        // src/cpu/forward3d.cpp:37-58 parallel).
        for (int c = 0; c <= nu; ++c) {
            const double b = keep ? c : (nu - c);
            if (p.kind == GEO_FLAT) {
                const double u = ((b - nu * 0.5) * du + shu) * mag;
                d[c] = axis_map(xdrv, sx, sy, u * cb, u * sb);
            } else if (p.kind == GEO_ARC) {
                const double g = (b - nu * 0.5) * du + shu + a;
                d[c] = xdrv ? sx + sy * std::tan(g) : sy + sx / std::tan(g);
            } else {
                d[c] = ((b - nu * 0.5) * du + shu) / (xdrv ? cb : sb);
            }
        }
        for (int c = 0; c < nu; ++c) {
            if (!cone || is2d) {
                sig[c] = 1.0;
                continue;
            }
            const double bm = keep ? (c + 0.5) : (nu - c - 0.5);
            double mx, my;
            if (p.kind == GEO_FLAT) {
                const double u = ((bm - nu * 0.5) * du + shu) * mag;
                mx = u * cb;
                my = u * sb;
            } else {
                const double g = (bm - nu * 0.5) * du + shu + a;
                mx = std::sin(g) * sdd + sx;
                my = -std::cos(g) * sdd + sy;
            }
            sig[c] = xdrv ? sy / (sy - my) : sx / (sx - mx);
        }
        for (int c = 0; c < nu; ++c) {
            if (!(d[c + 1] > d[c]) || !(sig[c] > 0.0)) {
                err = "view " + std::to_string(v) + ": detector column " + std::to_string(c) +
                      " does not map to an increasing interval (detector behind the source?)";
                return XR_EINVAL;
            }
        }
        const size_t cb0 = (size_t)v * (nu + 1);
        // W is the width that normalises the forward projection: the column's axis width
        // for the cone geometries (src/cpu/forward3d.cpp:333-336), and the detector pitch
        // for parallel beams, whose forward coefficient is ov * (dy or dx) / du
        // (src/cpu/forward3d.cpp:58, :92).
        for (int c = 0; c < nu; ++c)
            t.col[cb0 + c] = float4{(float)d[c], (float)sig[c], (float)(1.0 / (sig[c] * dvv)),
                                    (float)(cone ? d[c + 1] - d[c] : du)};
        t.col[cb0 + nu] = float4{(float)d[nu], 0.0f, 0.0f, 0.0f};
        h.dlo = (float)d[0];
        h.dhi = (float)d[nu];
        // Span class: the widest column, measured in pixel widths of the line with the
        // smallest pixel footprint, below one (below two): a column overlaps at most two
        // (three) consecutive pixels of any line, and the cell gather takes a fixed two-
        // (three-) pixel path.  A small margin keeps FP32 rounding of the boundaries on the
        // safe side.  Classified per 64-column group of the cell gather (the view's own
        // class is the widest).
        {
            double dmin = 1e300;
            for (int L = 0; L < h.nline; ++L) dmin = std::min(dmin, 1.0 / (double)t.lineC[(size_t)v * ML + L].y);
            auto cls = [dmin](double w) { return (w < 0.99 * dmin) ? 2 : (w < 1.99 * dmin) ? 3 : 0; };
            double wmax = 0.0;
            for (int g = 0; g < ncg; ++g) {
                double wg = 0.0;
                for (int c = g * kWave; c < std::min(nu, (g + 1) * kWave); ++c) wg = std::max(wg, d[c + 1] - d[c]);
                wmax = std::max(wmax, wg);
                gspan[(size_t)v * ncg + g] = cls(wg);
                gwmax[(size_t)v * ncg + g] = wg;
            }
            h.span = cls(wmax);
        }

        // Inverse map axis point -> fractional internal column, used by the voxel kernels
        // to find a footprint's first column (then corrected against the table).
        if (p.kind == GEO_FLAT) {
            // Virtual-detector coordinate of the ray through the axis point a:
            //   x-driven u = SOD a cos / (SOD + a sin), y-driven u = SOD a sin / (SOD - a cos);
            // internal column = Nu/2 + sk * (u / mag - shu) / du.
            if (xdrv) { h.tp1 = (float)(sod * cb); h.tp0 = 0; h.tq1 = (float)sb; h.tq0 = (float)sod; }
            else      { h.tp1 = (float)(sod * sb); h.tp0 = 0; h.tq1 = (float)(-cb); h.tq0 = (float)sod; }
            h.tk0 = (float)(nu * 0.5 - sk * shu / du);
            h.tk1 = (float)(sk / (mag * du));
        } else if (p.kind == GEO_ARC) {
            // Fan angle of the ray relative to the column-centre direction beta + shu:
            //   tan(gamma_rel) = T(a); internal column = Nu/2 + sk * atan(T) / du.
            const double bp = a + shu, cp = std::cos(bp), sp = std::sin(bp);
            const double s_sh = std::sin(shu), c_sh = std::cos(shu);
            if (xdrv) { h.tp1 = (float)cp; h.tp0 = (float)(-sod * s_sh); h.tq1 = (float)sp; h.tq0 = (float)(sod * c_sh); }
            else      { h.tp1 = (float)sp; h.tp0 = (float)(-sod * s_sh); h.tq1 = (float)(-cp); h.tq0 = (float)(sod * c_sh); }
            h.tk0 = (float)(nu * 0.5);
            h.tk1 = (float)(sk / du);
        } else {
            // d_c = ((b - Nu/2) du + shu) / cos  ->  c = Nu/2 + sk (a cos - shu) / du
            h.tp1 = 1; h.tp0 = 0; h.tq1 = 0; h.tq0 = 1;
            h.tk0 = (float)(nu * 0.5 - sk * shu / du);
            h.tk1 = (float)(sk * (xdrv ? cb : sb) / du);
        }
    }

    // Voxel gathers on volumes too small to fill the GPU split their views into sets.
    const int gx = ceil_div(nx, kVoxTile), gy = ceil_div(ny, kVoxTile);
    {
        const int64_t groups = (int64_t)gx * gy * ceil_div(nz, chunk_slices(default_chunk(p)));
        int sp = groups >= kVoxMinGroups ? 1 : (int)std::min<int64_t>(16, ceil_div(kVoxMinGroups, groups));
        p.vsplit = std::max(1, std::min(sp, V));
    }

    // Cell-gather column groups: depth-scale range of each group's columns per view, and
    // the span class from above.
    t.cgrp.assign((size_t)V * ncg, CellGroup{0.0f, 0.0f, 0, 0.0f});
    for (int v = 0; v < V; ++v)
        for (int g = 0; g < ncg; ++g) {
            float lo = 3.0e38f, hi = -3.0e38f;
            for (int c = g * kWave; c < std::min(nu, (g + 1) * kWave); ++c) {
                const float sg = t.col[(size_t)v * (nu + 1) + c].y;
                lo = std::min(lo, sg);
                hi = std::max(hi, sg);
            }
            t.cgrp[(size_t)v * ncg + g] = CellGroup{lo, hi, gspan[(size_t)v * ncg + g], (float)gwmax[(size_t)v * ncg + g]};
        }

    // The cell gather launches one grid per span class over that class's groups, so that
    // the narrow classes run with the registers of their own path only.  Within a class
    // the groups are listed column group by column group, views in order: a cell
    // workgroup takes 4 consecutive entries, i.e. one group of columns on (mostly) 4
    // adjacent views, which read nearly the same z-prefix entries (512^2 x 256 cone-arc
    // A per item 40.1 -> 35.7 ms against view-major lists, where the 4 waves took 4
    // neighbouring groups of one view; the 256^2 x 64 workload is unchanged).
    t.citems.clear();
    const int classes[3] = {2, 3, 0};
    for (int k = 0; k < 3; ++k) {
        t.ncls[k] = 0;
        for (int g = 0; g < ncg; ++g)
            for (int v = 0; v < V; ++v)
                if (t.cgrp[(size_t)v * ncg + g].span == classes[k]) {
                    t.citems.push_back(v * ncg + g);
                    ++t.ncls[k];
                }
        p.ncls[k] = t.ncls[k];
    }

    // LDS staging bounds of the cell gather's workgroups of classes 2 and 3 (QuadLine):
    // workgroup q of a class takes the class's entries 4q .. 4q+3.  A column's pixel range
    // on a line comes from its boundaries t = d lc.y + lc.x, evaluated as the kernel
    // evaluates them; the boundaries d are monotone in the column, so the group's first and
    // last boundaries bound all of its columns.
    {
        const int ML = p.maxline;
        int nq = 0;
        for (int k = 0; k < 2; ++k) {
            p.cquad_off[k] = nq;
            nq += (int)ceil_div(t.ncls[k], kWavesPerBlock);
        }
        t.cquad.assign((size_t)nq * ML, QuadLine{0, 0, 0.0f, 0.0f});
        const int ntile = cell_tile_count(nv);
        t.cqstage.assign((size_t)nq * ntile, 0);
        const CellArgs ca = make_cell_args(p, 1);
        std::vector<int> pitch((size_t)ceil_div(ML, kCellChunk));
        int start = 0;
        for (int k = 0; k < 2; ++k) {
            const int extra = k == 0 ? 1 : 2;  // entries a fixed-span read takes past its first pixel
            for (int q = 0; q * kWavesPerBlock < t.ncls[k]; ++q) {
                QuadLine *ql = t.cquad.data() + (size_t)(p.cquad_off[k] + q) * ML;
                const int i0 = start + q * kWavesPerBlock, i1 = std::min(start + t.ncls[k], i0 + kWavesPerBlock);
                const ViewHdr &h0 = t.hdr[t.citems[i0] / ncg];
                bool mixed = false;
                for (int i = i0; i < i1; ++i) mixed = mixed || t.hdr[t.citems[i] / ncg].xdrv != h0.xdrv;
                if (mixed) continue;  // not staged: the waves' line counts and tables may differ
                const float fal = (float)h0.nalong;
                for (int L = 0; L < h0.nline; ++L) {
                    int mlo = h0.nalong, mhi = -1;
                    float slo = 3.0e38f, shi = -3.0e38f;
                    for (int i = i0; i < i1; ++i) {
                        const int v = t.citems[i] / ncg, g = t.citems[i] - v * ncg;
                        const float4 lc = t.lineC[(size_t)v * ML + L];
                        const float4 *ct = t.col.data() + (size_t)v * (nu + 1);
                        const float ta = std::fmaf(ct[g * kWave].x, lc.y, lc.x);
                        const float tb = std::fmaf(ct[std::min(nu, (g + 1) * kWave)].x, lc.y, lc.x);
                        const float tlo = std::min(ta, tb), thi = std::max(ta, tb);
                        if (!(thi > 0.0f && tlo < fal)) continue;
                        mlo = std::min(mlo, std::max((int)std::floor(std::max(tlo, -1.0f)) - 1, 0));
                        mhi = std::max(mhi, std::min((int)std::floor(std::min(thi, fal)) + 1, h0.nalong - 1) + extra);
                        const CellGroup &cg = t.cgrp[(size_t)v * ncg + g];
                        slo = std::min(slo, std::min(cg.sig_lo * lc.z, cg.sig_hi * lc.z));
                        shi = std::max(shi, std::max(cg.sig_lo * lc.z, cg.sig_hi * lc.z));
                    }
                    if (mhi >= mlo) ql[L] = QuadLine{mlo, mhi - mlo + 1, slo, shi};
                }
                // Staged (group quad, row tile) pairs for tiles of 4, 8 and 16 rows: those
                // whose every block fits a slot, evaluated as the kernel evaluates them.
                cell_chunk_pitches(ql, h0.nline, pitch.data());
                int *flag = t.cqstage.data() + (size_t)(p.cquad_off[k] + q) * ntile;
                for (int R = 4; R <= 16; R *= 2)
                    for (int rt = 0; rt * R < nv; ++rt) {
                        bool fits = true;
                        for (int L = 0; fits && L < h0.nline; ++L) {
                            if (ql[L].nm <= 0) continue;
                            const int pt = pitch[L / kCellChunk];
                            const VoxBox bx = R == 4    ? cell_box<4>(ql[L], ca, rt * R, pt, h0.nalong)
                                              : R == 8 ? cell_box<8>(ql[L], ca, rt * R, pt, h0.nalong)
                                                       : cell_box<16>(ql[L], ca, rt * R, pt, h0.nalong);
                            fits = bx.nc > 0;
                        }
                        flag[cell_tile_offset(nv, R) + rt] = fits ? 1 : 0;
                    }
            }
            start += t.ncls[k];
        }
    }

    // Voxel-gather staging estimates per (view, 16 x 16 tile of voxel columns).
    t.vtile.assign((size_t)V * gx * gy, VoxTile{0, 0, 0.0f, 0.0f});
    for (int v = 0; v < V; ++v) {
        const float4 *ct = t.col.data() + (size_t)v * (nu + 1);
        for (int by = 0; by < gy; ++by)
            for (int bx = 0; bx < gx; ++bx) {
                const int i0 = bx * kVoxTile, j0 = by * kVoxTile;
                const int i1 = std::min(i0 + kVoxTile - 1, nx - 1), j1 = std::min(j0 + kVoxTile - 1, ny - 1);
                VoxTile &vt = t.vtile[((size_t)v * gy + by) * gx + bx];
                if (p.kind == GEO_FLAT) vt = voxel_tile<GEO_FLAT>(t.hdr[v], t.lineV.data(), ct, ny, nu, ML, v, i0, i1, j0, j1);
                else if (p.kind == GEO_ARC) vt = voxel_tile<GEO_ARC>(t.hdr[v], t.lineV.data(), ct, ny, nu, ML, v, i0, i1, j0, j1);
                else vt = voxel_tile<GEO_PAR>(t.hdr[v], t.lineV.data(), ct, ny, nu, ML, v, i0, i1, j0, j1);
            }
    }
    return XR_OK;
}

}  // namespace xr
